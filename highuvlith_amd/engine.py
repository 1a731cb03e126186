"""Host-side mirror of highuvlith-core's imaging interface for the SOCS path.

Names, argument meaning and error behaviour follow the reference:
  - ``ProjectionOptics`` / ``EuvProjectionOptics``   optics/mod.rs:201-360, optics/euv.rs:49-190
  - illumination shapes + ``Source``               source.rs:163-186, 1178-1248
  - ``GridConfig``                                 types.rs:210-319
  - ``ImagingSettings``                            aerial.rs:166-206
  - ``sample_source``                              aerial.rs:941
  - ``AerialImageEngine``                          aerial.rs:1472-2023
Kernel sets are built on the host in FP64 (C++, libhuv.so); images are
computed on the GPU through the C ABI (include/huv.h).
"""
import collections
import ctypes as C
import math
import threading
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import HuvError, check, dptr, i64ptr

# ---------------------------------------------------------------------------
# Optics
# ---------------------------------------------------------------------------
WATER_INDEX_193NM = 1.437
ARF_WAVELENGTH_NM = 193.368
KRF_WAVELENGTH_NM = 248.3
F2_WAVELENGTH_NM = 157.63
EUV_WAVELENGTH_NM = 13.5


@dataclass
class ProjectionOptics:
    """Refractive / immersion projection optics (optics/mod.rs:201-360)."""
    na: float
    immersion_index: float = 1.0
    flare_fraction: float = 0.02
    paraxial_defocus: bool = False

    def __post_init__(self):
        if not (self.immersion_index >= 1.0 and math.isfinite(self.immersion_index)):
            raise HuvError(1, "immersion_index: must be finite and >= 1")
        limit = 1.0 if self.immersion_index == 1.0 else 0.95 * self.immersion_index
        ok = self.na > 0 and (self.na < 1.0 if self.immersion_index == 1.0 else self.na <= limit)
        if not ok or math.isnan(self.na):
            raise HuvError(1, "numerical_aperture: out of range")

    @staticmethod
    def immersion_193i():
        """NA 1.35 in water, 2 % flare (optics/mod.rs:285)."""
        return ProjectionOptics(1.35, WATER_INDEX_193NM, 0.02)

    def to_c(self):
        return _lib.huv_optics(0, self.na, self.immersion_index, self.flare_fraction, 0.0, 1.0,
                               int(self.paraxial_defocus))

    @property
    def flare(self):
        return self.flare_fraction


@dataclass
class EuvProjectionOptics:
    """EUV mirror optics (optics/euv.rs:49-190): unit transmission, no flare by default."""
    numerical_aperture: float
    central_obscuration: float = 0.0
    flare: float = 0.0
    transmission: float = 1.0
    paraxial_defocus: bool = False

    @staticmethod
    def nxe_033():
        return EuvProjectionOptics(0.33, 0.0)

    @property
    def na(self):
        return self.numerical_aperture

    @property
    def flare_fraction(self):
        return self.flare

    def to_c(self):
        return _lib.huv_optics(1, self.numerical_aperture, 1.0, self.flare,
                               self.central_obscuration, self.transmission,
                               int(self.paraxial_defocus))


def pupil_function(optics, fx, fy, defocus_nm, wavelength_nm):
    lib = _lib.load()
    re, im = C.c_double(), C.c_double()
    o = optics.to_c()
    check(lib.huv_pupil(C.byref(o), fx, fy, defocus_nm, wavelength_nm, C.byref(re), C.byref(im)))
    return complex(re.value, im.value)


# ---------------------------------------------------------------------------
# Illumination
# ---------------------------------------------------------------------------
@dataclass
class IlluminationShape:
    kind: int
    params: tuple

    @staticmethod
    def Conventional(sigma):
        return IlluminationShape(0, (sigma,))

    @staticmethod
    def Annular(sigma_inner, sigma_outer):
        return IlluminationShape(1, (sigma_inner, sigma_outer))

    @staticmethod
    def Quadrupole(sigma_center, sigma_radius, opening_angle_deg):
        return IlluminationShape(2, (sigma_center, sigma_radius, opening_angle_deg))

    @staticmethod
    def Dipole(sigma_center, sigma_radius, orientation_deg=0.0):
        return IlluminationShape(3, (sigma_center, sigma_radius, orientation_deg))

    @staticmethod
    def CoherentGaussian(sigma):
        return IlluminationShape(4, (sigma,))

    @staticmethod
    def SectorPoles(sigma_inner, sigma_outer, opening_deg, first_pole_deg, poles):
        """Annular-sector poles: x-dipole = (.., 0, 2), quasar = (.., 45, 4)."""
        return IlluminationShape(5, (sigma_inner, sigma_outer, opening_deg, first_pole_deg,
                                     float(poles)))

    def to_c(self):
        p = (C.c_double * 6)(*(list(self.params) + [0.0] * (6 - len(self.params))))
        return _lib.huv_illumination(self.kind, p)

    def intensity_at(self, sx, sy):
        ill = self.to_c()
        return _lib.load().huv_illumination_at(C.byref(ill), sx, sy)


@dataclass
class Source:
    """Monochromatic source: centre wavelength and pupil fill."""
    wavelength_nm: float
    illumination: IlluminationShape

    @staticmethod
    def arf(illumination):
        return Source(ARF_WAVELENGTH_NM, illumination)

    @staticmethod
    def krf(sigma):
        return Source(KRF_WAVELENGTH_NM, IlluminationShape.Conventional(sigma))

    @staticmethod
    def f2(sigma):
        return Source(F2_WAVELENGTH_NM, IlluminationShape.Conventional(sigma))


@dataclass
class GridConfig:
    size: int
    pixel_nm: float

    def __post_init__(self):
        if self.size < 1 or (self.size & (self.size - 1)) != 0:
            raise HuvError(1, f"grid size {self.size} is not a power of two")
        if not (self.pixel_nm > 0 and math.isfinite(self.pixel_nm)):
            raise HuvError(1, "pixel_nm: must be positive")

    def field_size_nm(self):
        return self.size * self.pixel_nm

    def freq_step(self):
        return 1.0 / self.field_size_nm()


METHODS = {"dense_tcc": 0, "dense_gram": 1, "randomized": 2}
METHOD_NAMES = {v: k for k, v in METHODS.items()}


@dataclass
class ImagingSettings:
    """aerial.rs:166-206 (scalar, exact defocus, clear-field normalisation)."""
    max_kernels: int = 48
    kernel_energy_fraction: float = 1.0
    source_points_per_axis: Optional[int] = None
    kernel_cache_capacity: int = 32
    method: Optional[str] = None  # None = automatic choice of the reference


def sample_source(illumination, points_per_axis=None, df_sigma=0.5):
    """Source points (N, 3) = (sx, sy, weight), weights summing to 1 (aerial.rs:941)."""
    lib = _lib.load()
    ill = illumination.to_c()
    cnt = C.c_int64()
    ppa = 0 if points_per_axis is None else int(points_per_axis)
    check(lib.huv_sample_source(C.byref(ill), ppa, df_sigma, None, 0, C.byref(cnt)))
    out = np.zeros((cnt.value, 3), dtype=np.float64)
    check(lib.huv_sample_source(C.byref(ill), ppa, df_sigma, dptr(out), cnt.value, C.byref(cnt)))
    return out


def select_kernel_count(values, trace, max_kernels, fraction):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return _lib.load().huv_select_kernel_count(dptr(v), len(v), trace, max_kernels, fraction)


def rasterize_rects(rects, grid, background=1.0 + 0j):
    """Area-coverage raster of painted rectangles [(x, y, w, h, value)], complex (n, n)."""
    arr = np.zeros((max(len(rects), 1), 6))
    for i, (x, y, w, h, v) in enumerate(rects):
        arr[i] = (x, y, w, h, complex(v).real, complex(v).imag)
    out = np.zeros((grid.size, grid.size, 2))
    check(_lib.load().huv_rasterize_rects(grid.size, grid.pixel_nm, len(rects), dptr(arr),
                                          complex(background).real, complex(background).imag,
                                          dptr(out)))
    return out[..., 0] + 1j * out[..., 1]


# ---------------------------------------------------------------------------
# Kernel sets
# ---------------------------------------------------------------------------
@dataclass
class SparseKernelSet:
    """Kernels on the mask-frequency support only (aerial.rs:955-973)."""
    n: int
    support: np.ndarray          # (S,) int64 flat indices i*n + j
    nonzero_rows: np.ndarray     # (R,) int64
    eigenvalues: np.ndarray      # (K,) absolute
    values: np.ndarray           # (K, S) complex128
    captured_energy_fraction: float
    clear_field: float
    pupil_peak: float
    normalization_divisor: float
    method: str
    num_columns: int
    support_exceeds_nyquist: bool
    defocus_nm: float
    wavelength_nm: float

    @property
    def num_kernels(self):
        return len(self.eigenvalues)

    def dense(self):
        """KernelSet of aerial.rs:1912-1942: dense n x n kernels, eigenvalues / divisor."""
        k = np.zeros((self.num_kernels, self.n * self.n), dtype=np.complex128)
        k[:, self.support] = self.values
        return DenseKernelSet(self.eigenvalues / self.normalization_divisor,
                              k.reshape(self.num_kernels, self.n, self.n), self.defocus_nm,
                              self.wavelength_nm, self.captured_energy_fraction)


@dataclass
class DenseKernelSet:
    eigenvalues: np.ndarray
    kernels: np.ndarray
    defocus_nm: float
    wavelength_nm: float
    captured_energy_fraction: float


def build_kernel_set(optics, points, grid, defocus_nm, wavelength_nm, max_kernels,
                     kernel_energy_fraction=1.0, method=None):
    lib = _lib.load()
    o = optics.to_c()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    h = C.c_void_p()
    m = -1 if method is None else METHODS[method]
    check(lib.huv_kernels_build(C.byref(o), dptr(pts), len(pts), grid.size, grid.pixel_nm,
                                defocus_nm, wavelength_nm, max_kernels, kernel_energy_fraction,
                                m, C.byref(h)))
    return _take_kernel_set(lib, h, grid, defocus_nm, wavelength_nm)


def build_kernel_sets_device(ctx, optics, points, grid, defocus_nm, wavelength_nm, max_kernels,
                             kernel_energy_fraction=1.0, method=None):
    """Kernel sets of several focus planes with the TCC factor, its Gram product
    and the kernels computed on the GPU in FP64 and the eigendecompositions on
    host threads (dense TCC / dense Gram paths; huv_kernels_build_device)."""
    lib = _lib.load()
    o = optics.to_c()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    zs = np.ascontiguousarray(np.atleast_1d(defocus_nm), dtype=np.float64)
    m = -1 if method is None else METHODS[method]
    hs = (C.c_void_p * len(zs))()
    check(lib.huv_kernels_build_device(ctx.handle, C.byref(o), dptr(pts), len(pts), grid.size,
                                       grid.pixel_nm, len(zs), dptr(zs), wavelength_nm,
                                       max_kernels, kernel_energy_fraction, m, hs))
    return [_take_kernel_set(lib, C.c_void_p(hs[i]), grid, float(zs[i]), wavelength_nm)
            for i in range(len(zs))]


def _take_kernel_set(lib, h, grid, defocus_nm, wavelength_nm):
    """SparseKernelSet copied out of a huv_host_kernels handle, which is freed."""
    try:
        info = _lib.huv_kernels_info()
        check(lib.huv_kernels_info_get(h, C.byref(info)))
        K, S = info.num_kernels, info.num_support
        support = np.zeros(S, dtype=np.int64)
        rows = np.zeros(info.num_rows, dtype=np.int64)
        eig = np.zeros(K, dtype=np.float64)
        vals = np.zeros((K, S, 2), dtype=np.float64)
        check(lib.huv_kernels_copy(h, i64ptr(support), i64ptr(rows), dptr(eig), dptr(vals)))
    finally:
        lib.huv_kernels_free(h)
    return SparseKernelSet(
        n=grid.size, support=support, nonzero_rows=rows, eigenvalues=eig,
        values=vals[..., 0] + 1j * vals[..., 1],
        captured_energy_fraction=info.captured_energy_fraction, clear_field=info.clear_field,
        pupil_peak=info.pupil_peak, normalization_divisor=info.normalization_divisor,
        method=METHOD_NAMES[info.method], num_columns=info.num_columns,
        support_exceeds_nyquist=bool(info.support_exceeds_nyquist), defocus_nm=defocus_nm,
        wavelength_nm=wavelength_nm)


# ---------------------------------------------------------------------------
# GPU context
# ---------------------------------------------------------------------------
class DeviceContext:
    """One huv_ctx (one HIP stream) on a GPU; raises when no MI355X is present."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.huv_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def set_stream(self, stream_ptr):
        check(self.lib.huv_ctx_set_stream(self.handle, C.c_void_p(stream_ptr)))

    def synchronize(self):
        check(self.lib.huv_ctx_synchronize(self.handle))

    STAGES = ("rows_band", "cols_band", "socs_fields", "coarse_rows", "coarse_cols",
              "band_image", "ilt_pointwise", "ilt_adjoint", "ilt_other", "rect_tables",
              "rect_band", "poly_band", "coarse")

    def set_profiling(self, on=True):
        check(self.lib.huv_ctx_set_profiling(self.handle, int(bool(on))))

    def stage_times(self, reset=True):
        """{stage: (total_ms, launches)} from HIP events on the context stream."""
        ms = np.zeros(len(self.STAGES))
        calls = np.zeros(len(self.STAGES), dtype=np.int64)
        check(self.lib.huv_ctx_stage_times(self.handle, dptr(ms), i64ptr(calls), len(ms),
                                           int(bool(reset))))
        return {s: (float(ms[i]), int(calls[i])) for i, s in enumerate(self.STAGES)}

    def set_autotune(self, on=True):
        """Time every launch shape on the first call of each new shape class."""
        check(self.lib.huv_ctx_set_autotune(self.handle, int(bool(on))))

    def tuning(self, n, band, nplanes, kmax):
        """(m, n1, xnb, xlanes, chunks) chosen for this shape class, or None."""
        t = _lib.huv_socs_tuning()
        found = C.c_int()
        check(self.lib.huv_ctx_tuning_get(self.handle, n, band, nplanes, kmax, C.byref(t),
                                          C.byref(found)))
        return (t.m, t.n1, t.xnb, t.xlanes, t.chunks) if found.value else None

    def set_tuning(self, n, band, nplanes, kmax, m=0, n1=0, xnb=0, xlanes=0, chunks=0):
        t = _lib.huv_socs_tuning(int(m), int(n1), int(xnb), int(xlanes), int(chunks))
        check(self.lib.huv_ctx_tuning_set(self.handle, n, band, nplanes, kmax, C.byref(t)))

    def upload(self, ks, flare):
        """Upload a SparseKernelSet (eigenvalues normalised) -> DeviceKernelSet."""
        eig = np.ascontiguousarray(ks.eigenvalues / ks.normalization_divisor)
        vals = np.ascontiguousarray(np.stack([ks.values.real, ks.values.imag], -1))
        sup = np.ascontiguousarray(ks.support, dtype=np.int64)
        h = C.c_void_p()
        check(self.lib.huv_kset_upload(self.handle, ks.n, ks.num_kernels, len(sup), i64ptr(sup),
                                       dptr(eig), dptr(vals), float(flare), C.byref(h)))
        return DeviceKernelSet(self, h, ks)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.huv_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceKernelSet:
    def __init__(self, ctx, handle, host):
        self.ctx, self.handle, self.host = ctx, handle, host
        b, m = C.c_int(), C.c_int()
        check(ctx.lib.huv_kset_shape(handle, C.byref(b), C.byref(m)))
        self.band, self.coarse = b.value, m.value

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.ctx.lib.huv_kset_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


def socs_images_device(ctx, masks, planes, out, pixel_transfer=False, graph=False):
    """Launch batched images on torch CUDA tensors (masks (N, n, n) f32, out (N, P, n, n) f32).

    Runs on torch's current stream.  graph=True records the launch sequence
    in a HIP graph on the first call and replays it for later calls on the
    same tensors.
    """
    import torch
    assert masks.is_cuda and masks.dtype == torch.float32 and masks.is_contiguous()
    assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
    nclips, n = masks.shape[0], masks.shape[-1]
    if out.shape != (nclips, len(planes), n, n):
        raise HuvError(2, f"output shape {tuple(out.shape)} != {(nclips, len(planes), n, n)}")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    opts = _lib.huv_socs_opts(int(bool(pixel_transfer)), int(bool(graph)))
    check(ctx.lib.huv_socs_images(ctx.handle, nclips, n, C.c_void_p(masks.data_ptr()),
                                  len(planes), arr, C.c_void_p(out.data_ptr()), C.byref(opts)))
    return out


def socs_images_complex_device(ctx, masks, planes, out, pixel_transfer=False, graph=False):
    """Images of complex transmittance rasters: masks (N, n, n) complex64 CUDA tensor,
    out (N, P, n, n) f32; on torch's current stream (aerial.rs:1789-1799)."""
    import torch
    assert masks.is_cuda and masks.dtype == torch.complex64 and masks.is_contiguous()
    assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
    nclips, n = masks.shape[0], masks.shape[-1]
    if out.shape != (nclips, len(planes), n, n):
        raise HuvError(2, f"output shape {tuple(out.shape)} != {(nclips, len(planes), n, n)}")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    opts = _lib.huv_socs_opts(int(bool(pixel_transfer)), int(bool(graph)))
    check(ctx.lib.huv_socs_images_complex(ctx.handle, nclips, n, C.c_void_p(masks.data_ptr()),
                                          len(planes), arr, C.c_void_p(out.data_ptr()),
                                          C.byref(opts)))
    return out


def socs_images_spectrum(ctx, spectra, planes, out):
    """Images from host mask spectra: spectra (N, n, n) complex128 numpy in
    Fft2D::forward's layout and scale, out (N, P, n, n) f32 CUDA tensor
    (compute_from_spectrum, aerial.rs:1805-1818).  Synchronises torch's current
    stream before staging."""
    import torch
    spectra = np.ascontiguousarray(spectra, dtype=np.complex128)
    if spectra.ndim == 2:
        spectra = spectra[None]
    nclips, n = spectra.shape[0], spectra.shape[-1]
    assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
    if out.shape != (nclips, len(planes), n, n) or spectra.shape[1] != n:
        raise HuvError(2, f"spectrum {spectra.shape} / output {tuple(out.shape)} do not match "
                          f"{len(planes)} planes")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    check(ctx.lib.huv_socs_images_spectrum(ctx.handle, nclips, n,
                                           spectra.view(np.float64).ctypes.data_as(_lib.PD),
                                           len(planes), arr, C.c_void_p(out.data_ptr())))
    return out


def socs_images_batch_device(ctx, groups, pixel_transfer=False, graph=False):
    """Uneven batch: groups = [(masks (N_i, n_i, n_i), planes_i, out (N_i, P_i, n_i, n_i)), ...]
    of torch CUDA f32 tensors, all issued on torch's current stream; with
    graph=True the whole batch replays as one graph launch."""
    import torch
    keep = []
    arr = (_lib.huv_clip_group * len(groups))()
    for i, (masks, planes, out) in enumerate(groups):
        assert masks.is_cuda and masks.dtype == torch.float32 and masks.is_contiguous()
        assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
        nclips, n = masks.shape[0], masks.shape[-1]
        if out.shape != (nclips, len(planes), n, n):
            raise HuvError(2, f"group {i}: output shape {tuple(out.shape)} != "
                              f"{(nclips, len(planes), n, n)}")
        pl = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
        keep.append(pl)
        arr[i] = _lib.huv_clip_group(n, nclips, masks.data_ptr(), len(planes),
                                     C.cast(pl, C.POINTER(C.c_void_p)), out.data_ptr())
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    opts = _lib.huv_socs_opts(int(bool(pixel_transfer)), int(bool(graph)))
    check(ctx.lib.huv_socs_images_batch(ctx.handle, len(groups), arr, C.byref(opts)))


def band_spectrum_device(ctx, masks, band, pixel_transfer=False):
    """M[kx+B][ky+B] band of FFT_n(mask) per clip (torch complex64 (N, W, W))."""
    import torch
    nclips, n = masks.shape[0], masks.shape[-1]
    W = 2 * band + 1
    out = torch.empty((nclips, W, W, 2), dtype=torch.float32, device=masks.device)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    check(ctx.lib.huv_band_spectrum(ctx.handle, nclips, n, band, C.c_void_p(masks.data_ptr()),
                                    C.c_void_p(out.data_ptr()), int(bool(pixel_transfer))))
    return torch.view_as_complex(out)


_MAX_GRATING_PERIODS = float(1 << 20)  # mask.rs:101


def _grating_intervals(width, pitch, offset, L):
    """Intervals of a grating (lines of `width` every `pitch`, one centred at
    `offset`) inside the field [-L/2, L/2), and the amplitude fraction
    (mask.rs grating axes; more than 2^20 periods become their mean)."""
    half = L / 2.0
    if not (math.isfinite(width) and math.isfinite(pitch) and math.isfinite(offset)):
        return [], 1.0
    if width <= 0.0 or pitch <= 0.0:
        return [], 1.0
    if width >= pitch:
        return [(-half, half)], 1.0
    if L / pitch > _MAX_GRATING_PERIODS:
        return [(-half, half)], width / pitch
    k_lo = math.floor((-half - offset - width / 2.0) / pitch)
    k_hi = math.ceil((half - offset + width / 2.0) / pitch)
    out = []
    for k in range(k_lo, k_hi + 1):
        c = offset + k * pitch
        lo, hi = max(c - width / 2.0, -half), min(c + width / 2.0, half)
        if hi > lo:
            out.append((lo, hi))
    return out, 1.0


def mask_feature_rects(features, grid, dark_field=False, absorber=0j):
    """The reference's rectangle-type mask features (mask.rs:120-180) as painted
    rectangles for rect_arrays / compute_rects: returns (rects, background).

    features: ("rect", x, y, w, h) | ("gray", x, y, w, h, T) |
    ("lines", cd, pitch, orientation, offset) (orientation 0: lines along y,
    repeating along x) | ("array", w, h, pitch_x, pitch_y, offset_x, offset_y).
    A feature is clear (1) on a dark-field mask and the absorber amplitude on a
    bright-field one, where the background is the other; a gray rectangle has
    amplitude sqrt(T).  Gratings are expanded into their periods inside the
    field (exact: the Fourier series of a field-periodic grating is the sum of
    its clipped periods)."""
    L = grid.size * grid.pixel_nm
    bg = complex(absorber) if dark_field else 1.0 + 0j
    fg = 1.0 + 0j if dark_field else complex(absorber)
    rects = []
    for f in features:
        kind = f[0]
        if kind == "rect":
            x, y, w, h = map(float, f[1:5])
            rects.append((x, y, w, h, fg))
        elif kind == "gray":
            x, y, w, h, t = map(float, f[1:6])
            # mask.rs:360 rejects a transmittance outside [0, 1] when the mask is
            # built; the spectrum drops a non-finite one and clamps (mask.rs:1009-1012)
            if not math.isfinite(t):
                continue
            if not 0.0 <= t <= 1.0:
                raise HuvError(1, f"gray transmittance {t} must lie in [0, 1]")
            rects.append((x, y, w, h, math.sqrt(t)))
        elif kind == "lines":
            cd, pitch, orient, offset = float(f[1]), float(f[2]), int(f[3]), float(f[4])
            ivs, frac = _grating_intervals(cd, pitch, offset, L)
            v = bg + (fg - bg) * frac
            for lo, hi in ivs:
                if orient == 0:
                    rects.append(((lo + hi) / 2, 0.0, hi - lo, L, v))
                else:
                    rects.append((0.0, (lo + hi) / 2, L, hi - lo, v))
        elif kind == "array":
            w, h, px, py, ox, oy = map(float, f[1:7])
            ix, fx = _grating_intervals(w, px, ox, L)
            iy, fy = _grating_intervals(h, py, oy, L)
            v = bg + (fg - bg) * (fx * fy)
            for ylo, yhi in iy:
                for xlo, xhi in ix:
                    rects.append(((xlo + xhi) / 2, (ylo + yhi) / 2, xhi - xlo, yhi - ylo, v))
        else:
            raise HuvError(2, f"unsupported mask feature {kind!r} for the rectangle path")
    return rects, bg


def mask_feature_lists(features, grid, dark_field=False, absorber=0j):
    """Like mask_feature_rects, and ("polygon", [(x, y), ...]) features as
    (vertices, amplitude) pairs: returns (rects, polygons, background)."""
    rect_feats, polys = [], []
    fg = 1.0 + 0j if dark_field else complex(absorber)
    for f in features:
        if f[0] == "polygon":
            polys.append(([(float(x), float(y)) for (x, y) in f[1]], fg))
        else:
            rect_feats.append(f)
    rects, bg = mask_feature_rects(rect_feats, grid, dark_field, absorber)
    return rects, polys, bg


class PolygonArrays:
    """Host arrays of a batch's polygons for the C ABI (huv_polygons):
    poly_clips = [[(vertices, value), ...], ...], one list per clip."""

    def __init__(self, poly_clips):
        nclips = len(poly_clips)
        self.poly_offsets = np.zeros(nclips + 1, np.int64)
        verts, vals, voff = [], [], [0]
        for i, polys in enumerate(poly_clips):
            self.poly_offsets[i + 1] = self.poly_offsets[i] + len(polys)
            for vertices, value in polys:
                verts.extend(vertices)
                voff.append(len(verts))
                v = complex(value)
                vals.append((v.real, v.imag))
        self.vertex_offsets = np.asarray(voff, np.int64)
        self.vertices = np.ascontiguousarray(np.asarray(verts, np.float64).reshape(-1, 2)
                                             if verts else np.zeros((1, 2)))
        self.values = np.ascontiguousarray(np.asarray(vals, np.float64).reshape(-1, 2)
                                           if vals else np.zeros((1, 2)))
        self.struct = _lib.huv_polygons(i64ptr(self.poly_offsets), i64ptr(self.vertex_offsets),
                                        dptr(self.vertices), dptr(self.values))


def rect_arrays(clips, backgrounds=0.0):
    """Host arrays (offsets, rects, background) of a batch of rectangle masks for the C ABI:
    clips = [[(x, y, w, h, value), ...], ...] (centre and size in nm, later rectangles
    win), backgrounds = one amplitude or one per clip."""
    nclips = len(clips)
    if np.ndim(backgrounds) == 0:
        backgrounds = [backgrounds] * nclips
    if len(backgrounds) != nclips:
        raise HuvError(2, "one background per clip")
    offsets = np.zeros(nclips + 1, np.int64)
    for i, c in enumerate(clips):
        offsets[i + 1] = offsets[i] + len(c)
    rects = np.zeros((max(int(offsets[-1]), 1), 6))
    r = 0
    for c in clips:
        for (x, y, w, h, v) in c:
            v = complex(v)
            rects[r] = (x, y, w, h, v.real, v.imag)
            r += 1
    bg = np.array([[complex(b).real, complex(b).imag] for b in backgrounds],
                  np.float64).reshape(nclips, 2)
    return offsets, rects, np.ascontiguousarray(bg)


def band_spectrum_features_device(ctx, clips, poly_clips, grid, band, backgrounds=0.0):
    """Exact band spectrum of masks of rectangles (clips, as for rect_arrays) and simple
    polygons (poly_clips, as for PolygonArrays); torch complex64 (N, W, W)."""
    import torch
    if len(poly_clips) != len(clips):
        raise HuvError(2, "one polygon list per clip")
    offsets, rects, bg = rect_arrays(clips, backgrounds)
    pa = PolygonArrays(poly_clips)
    nclips, W = len(clips), 2 * band + 1
    out = torch.empty((nclips, W, W, 2), dtype=torch.float32, device="cuda")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    check(ctx.lib.huv_band_spectrum_features(
        ctx.handle, nclips, grid.size, grid.pixel_nm, band, i64ptr(offsets), dptr(rects),
        C.byref(pa.struct), dptr(bg), C.c_void_p(out.data_ptr())))
    return torch.view_as_complex(out)


def band_spectrum_rects_device(ctx, clips, grid, band, backgrounds=0.0):
    """Exact band spectrum S[kx+B][ky+B] of rectangle masks (torch complex64 (N, W, W)),
    in the layout and scale of band_spectrum_device (mask.rs:616-657 analytic path)."""
    import torch
    offsets, rects, bg = rect_arrays(clips, backgrounds)
    nclips, W = len(clips), 2 * band + 1
    out = torch.empty((nclips, W, W, 2), dtype=torch.float32, device="cuda")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    check(ctx.lib.huv_band_spectrum_rects(
        ctx.handle, nclips, grid.size, grid.pixel_nm, band,
        i64ptr(offsets), dptr(rects), dptr(bg),
        C.c_void_p(out.data_ptr())))
    return torch.view_as_complex(out)


def socs_images_rects_device(ctx, grid, arrays, planes, out):
    """Images of rectangle masks (arrays = rect_arrays(...)) into out (N, P, n, n) f32 CUDA,
    on torch's current stream; the call synchronises that stream before staging."""
    import torch
    offsets, rects, bg = arrays
    nclips, n = len(offsets) - 1, grid.size
    assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
    if out.shape != (nclips, len(planes), n, n):
        raise HuvError(2, f"output shape {tuple(out.shape)} != {(nclips, len(planes), n, n)}")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    check(ctx.lib.huv_socs_images_rects(
        ctx.handle, nclips, n, grid.pixel_nm, i64ptr(offsets), dptr(rects), dptr(bg),
        len(planes), arr, C.c_void_p(out.data_ptr()), None))
    return out


def socs_images_features_device(ctx, grid, arrays, polygons, planes, out):
    """Images of masks of rectangles (arrays = rect_arrays(...)) and simple polygons
    (polygons = PolygonArrays(...)) into out (N, P, n, n) f32 CUDA, on torch's
    current stream; the call synchronises that stream before staging."""
    import torch
    offsets, rects, bg = arrays
    nclips, n = len(offsets) - 1, grid.size
    assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()
    if out.shape != (nclips, len(planes), n, n):
        raise HuvError(2, f"output shape {tuple(out.shape)} != {(nclips, len(planes), n, n)}")
    if len(polygons.poly_offsets) != nclips + 1:
        raise HuvError(2, "one polygon list per clip")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    check(ctx.lib.huv_socs_images_features(
        ctx.handle, nclips, n, grid.pixel_nm, i64ptr(offsets), dptr(rects),
        C.byref(polygons.struct), dptr(bg), len(planes), arr, C.c_void_p(out.data_ptr()), None))
    return out


# ---------------------------------------------------------------------------
# Engine
# ---------------------------------------------------------------------------
class AerialImageEngine:
    """Hopkins/SOCS imaging engine (aerial.rs:1472-2023), GPU-backed.

    Construction samples the source and builds the in-focus kernel set on the
    host in FP64; kernel sets for other focus planes are built on demand and
    kept in a bounded LRU cache.  Images are computed on the GPU.
    """

    def __init__(self, source, optics, grid, max_kernels=None, settings=None, device=0,
                 context=None):
        """context: an existing DeviceContext to share (engines of several grid
        sizes feeding one uneven batch must share one)."""
        if settings is None:
            settings = ImagingSettings()
        if max_kernels is not None:
            settings = ImagingSettings(max_kernels, settings.kernel_energy_fraction,
                                       settings.source_points_per_axis,
                                       settings.kernel_cache_capacity, settings.method)
        if settings.max_kernels < 1:
            raise HuvError(1, "max_kernels: must be at least 1")
        if not (0.0 < settings.kernel_energy_fraction <= 1.0):
            raise HuvError(1, "kernel_energy_fraction: must be in (0, 1]")
        if settings.source_points_per_axis == 0:
            raise HuvError(1, "source_points_per_axis: must be at least 1")
        self.source, self.optics, self.grid, self.settings = source, optics, grid, settings
        self.device = device
        cutoff = optics.na / source.wavelength_nm
        self.df_sigma = grid.freq_step() / cutoff
        self._points = sample_source(source.illumination, settings.source_points_per_axis,
                                     self.df_sigma)
        self._lock = threading.Lock()
        self._cache = collections.OrderedDict()
        self._focus = self._build(0.0)
        self._ctx = context
        if context is not None:
            self.device = context.device
        self._dev = {}

    # -- kernel sets --------------------------------------------------------
    def _build(self, z):
        s = self.settings
        return build_kernel_set(self.optics, self._points, self.grid, z, self.source.wavelength_nm,
                                s.max_kernels, s.kernel_energy_fraction, s.method)

    @staticmethod
    def _key(z):
        return 0.0 if z == 0.0 else float(z)

    def kernel_set(self, defocus_nm=0.0):
        z = self._key(defocus_nm)
        if z == 0.0:
            return self._focus
        with self._lock:
            if z in self._cache:
                self._cache.move_to_end(z)
                return self._cache[z]
        ks = self._build(z)
        with self._lock:
            if self.settings.kernel_cache_capacity > 0:
                self._cache[z] = ks
                while len(self._cache) > self.settings.kernel_cache_capacity:
                    self._cache.popitem(last=False)
        return ks

    def build_planes_device(self, defocus_nm):
        """Build the kernel sets of several focus planes at once with the GPU's help
        (huv_kernels_build_device) and keep them in the cache; planes already
        cached are kept.  Returns the sets in order."""
        zs = [self._key(z) for z in defocus_nm]
        with self._lock:
            todo = sorted({z for z in zs if z != 0.0 and z not in self._cache})
        if todo:
            s = self.settings
            sets = build_kernel_sets_device(self.context(), self.optics, self._points, self.grid,
                                            todo, self.source.wavelength_nm, s.max_kernels,
                                            s.kernel_energy_fraction, s.method)
            with self._lock:
                for z, ks in zip(todo, sets):
                    if self.settings.kernel_cache_capacity > 0:
                        self._cache[z] = ks
                while len(self._cache) > self.settings.kernel_cache_capacity:
                    self._cache.popitem(last=False)
            built = dict(zip(todo, sets))
        else:
            built = {}
        return [built[z] if z in built else self.kernel_set(z) for z in zs]

    def kernels(self, defocus_nm=0.0):
        """Dense kernel set (eigenvalues / clear field) for ILT (aerial.rs:1912-1942)."""
        return self.kernel_set(defocus_nm).dense()

    def kernel_diagnostics(self, defocus_nm=0.0):
        ks = self.kernel_set(defocus_nm)
        return dict(num_kernels=ks.num_kernels, captured_energy_fraction=ks.captured_energy_fraction,
                    num_frequencies=len(ks.support), num_columns=ks.num_columns, method=ks.method,
                    support_exceeds_nyquist=ks.support_exceeds_nyquist,
                    clear_field_intensity=ks.clear_field,
                    normalized=ks.normalization_divisor != 1.0)

    def clear_field_intensity(self, defocus_nm=0.0):
        return self.kernel_set(defocus_nm).clear_field

    def num_kernels(self):
        return self._focus.num_kernels

    def captured_energy_fraction(self):
        return self._focus.captured_energy_fraction

    @property
    def source_points(self):
        return self._points

    def flare_fraction(self):
        return self.optics.flare_fraction

    def wavelength_nm(self):
        return self.source.wavelength_nm

    # -- GPU ----------------------------------------------------------------
    def context(self):
        if self._ctx is None:
            self._ctx = DeviceContext(self.device)
        return self._ctx

    def device_kernel_set(self, defocus_nm=0.0):
        z = self._key(defocus_nm)
        if z not in self._dev:
            self._dev[z] = self.context().upload(self.kernel_set(z), self.flare_fraction())
        return self._dev[z]

    def compute_batch(self, masks, defocus_nm=(0.0,), pixel_transfer=False, out=None,
                      graph=False):
        """Images of a batch of real masks at several focus planes on the GPU.

        masks: torch float32 CUDA tensor (N, n, n) (or numpy, copied).  Returns
        a torch tensor (N, P, n, n) on the same device.
        """
        import torch
        if not isinstance(masks, torch.Tensor):
            masks = torch.as_tensor(np.ascontiguousarray(masks, dtype=np.float32),
                                    device=f"cuda:{self.device}")
        if masks.dim() == 2:
            masks = masks[None]
        n = self.grid.size
        if masks.shape[-2:] != (n, n):
            raise HuvError(2, f"transmittance map must match the engine grid ({n}x{n})")
        masks = masks.contiguous()
        planes = [self.device_kernel_set(z) for z in defocus_nm]
        if out is None:
            out = torch.empty((masks.shape[0], len(planes), n, n), dtype=torch.float32,
                              device=masks.device)
        return socs_images_device(self.context(), masks, planes, out, pixel_transfer, graph)

    def compute_rects(self, clips, defocus_nm=(0.0,), backgrounds=0.0, out=None):
        """Images of rectangle masks from their exact spectrum (compute(mask),
        aerial.rs:1775-1777, for masks of rectangles): clips = [[(x, y, w, h, value),
        ...], ...] in nm, later rectangles win, or the tuple of rect_arrays;
        backgrounds = one amplitude or one per clip (0 = dark field).  Returns a
        torch tensor (N, P, n, n)."""
        import torch
        n = self.grid.size
        planes = [self.device_kernel_set(z) for z in defocus_nm]
        arrays = clips if isinstance(clips, tuple) else rect_arrays(clips, backgrounds)
        nclips = len(arrays[0]) - 1
        if out is None:
            out = torch.empty((nclips, len(planes), n, n), dtype=torch.float32,
                              device=f"cuda:{self.device}")
        return socs_images_rects_device(self.context(), self.grid, arrays, planes, out)

    def compute_features(self, masks, defocus_nm=(0.0,), dark_field=False, absorber=0j):
        """compute(mask) (aerial.rs:1775-1777) for a batch of masks given as feature
        lists (rectangles, gray rectangles, gratings, arrays and simple polygons; see
        mask_feature_lists), from their exact spectrum.  Returns a torch tensor
        (N, P, n, n)."""
        import torch
        clips, polys, bgs = [], [], []
        for feats in masks:
            r, p, b = mask_feature_lists(feats, self.grid, dark_field, absorber)
            clips.append(r)
            polys.append(p)
            bgs.append(b)
        n = self.grid.size
        planes = [self.device_kernel_set(z) for z in defocus_nm]
        out = torch.empty((len(clips), len(planes), n, n), dtype=torch.float32,
                          device=f"cuda:{self.device}")
        return socs_images_features_device(self.context(), self.grid, rect_arrays(clips, bgs),
                                           PolygonArrays(polys), planes, out)

    def compute_batch_complex(self, masks, defocus_nm=(0.0,), pixel_transfer=False, out=None,
                              graph=False):
        """compute_batch for complex transmittance maps: masks (N, n, n) complex
        (torch complex64 CUDA tensor, or numpy, copied).  Returns (N, P, n, n)."""
        import torch
        if not isinstance(masks, torch.Tensor):
            masks = torch.as_tensor(np.ascontiguousarray(masks, dtype=np.complex64),
                                    device=f"cuda:{self.device}")
        if masks.dim() == 2:
            masks = masks[None]
        n = self.grid.size
        if masks.shape[-2:] != (n, n):
            raise HuvError(2, f"transmittance map must match the engine grid ({n}x{n})")
        masks = masks.to(torch.complex64).contiguous()
        planes = [self.device_kernel_set(z) for z in defocus_nm]
        if out is None:
            out = torch.empty((masks.shape[0], len(planes), n, n), dtype=torch.float32,
                              device=masks.device)
        return socs_images_complex_device(self.context(), masks, planes, out, pixel_transfer,
                                          graph)

    def compute_from_spectrum_batch(self, spectra, defocus_nm=(0.0,)):
        """compute_from_spectrum (aerial.rs:1805-1818) for (N, n, n) host spectra in
        Fft2D::forward's layout and scale; returns a torch tensor (N, P, n, n)."""
        import torch
        spectra = np.asarray(spectra)
        if spectra.ndim == 2:
            spectra = spectra[None]
        n = self.grid.size
        if spectra.shape[-2:] != (n, n):
            raise HuvError(2, f"mask spectrum must match the engine grid ({n}x{n})")
        planes = [self.device_kernel_set(z) for z in defocus_nm]
        out = torch.empty((spectra.shape[0], len(planes), n, n), dtype=torch.float32,
                          device=f"cuda:{self.device}")
        return socs_images_spectrum(self.context(), spectra, planes, out)

    def compute_from_spectrum(self, spectrum, defocus_nm=0.0):
        return self.compute_from_spectrum_batch(spectrum, (defocus_nm,))[0, 0].cpu().numpy()

    def compute_from_transmittance(self, transmittance, defocus_nm=0.0):
        """aerial.rs:1789-1799: real maps take the real-raster path, complex ones
        (any nonzero imaginary part) the complex-raster path."""
        return self.compute_through_focus_from_transmittance(transmittance, (defocus_nm,))[0]

    def compute_through_focus_from_transmittance(self, transmittance, defocus_nm):
        t = np.asarray(transmittance)
        if np.iscomplexobj(t) and np.any(t.imag != 0):
            imgs = self.compute_batch_complex(t, tuple(defocus_nm))[0].cpu().numpy()
        else:
            imgs = self.compute_batch(np.real(t), tuple(defocus_nm))[0].cpu().numpy()
        return [imgs[i] for i in range(len(defocus_nm))]
