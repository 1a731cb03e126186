"""MI355X backend for highuvlith-core's Hopkins/SOCS aerial-image path and its
ILT adjoint: FP64 host kernel construction (C++) + HIP kernels for gfx950,
behind the C ABI of include/huv.h."""
from ._lib import HuvError  # noqa: F401
from .engine import (  # noqa: F401
    AerialImageEngine, DenseKernelSet, DeviceContext, EuvProjectionOptics, GridConfig,
    IlluminationShape, ImagingSettings, ProjectionOptics, Source, SparseKernelSet,
    build_kernel_set, pupil_function, rasterize_rects, sample_source, select_kernel_count)
