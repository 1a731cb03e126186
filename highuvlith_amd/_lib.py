"""ctypes binding of the C ABI in include/huv.h (libhuv.so, built in-tree).

Loading fails loudly when the library has not been built: there is no
Python or CPU fallback for the device entry points.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhuv.so")

HUV_OK = 0
HUV_INVALID_PARAMETER = 1
HUV_UNSUPPORTED = 8
STATUS_NAMES = {
    1: "InvalidParameter",
    2: "DimensionMismatch",
    3: "NumericalError",
    4: "InternalError",
    5: "DeviceError",
    6: "NoDiffractionOrders",
    7: "PupilSamplingTooDense",
    8: "Unsupported",
}


class HuvError(RuntimeError):
    """A non-zero huv_status, with the library's message."""

    def __init__(self, status, message):
        self.status = status
        self.kind = STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{self.kind}: {message}")


class huv_optics(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("na", C.c_double),
        ("immersion_index", C.c_double),
        ("flare", C.c_double),
        ("central_obscuration", C.c_double),
        ("transmission", C.c_double),
        ("paraxial_defocus", C.c_int),
    ]


class huv_illumination(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 6)]


class huv_kernels_info(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("pixel_nm", C.c_double),
        ("num_kernels", C.c_int),
        ("num_support", C.c_int64),
        ("num_rows", C.c_int64),
        ("method", C.c_int),
        ("num_columns", C.c_int64),
        ("support_exceeds_nyquist", C.c_int),
        ("captured_energy_fraction", C.c_double),
        ("clear_field", C.c_double),
        ("pupil_peak", C.c_double),
        ("normalization_divisor", C.c_double),
        ("defocus_nm", C.c_double),
        ("wavelength_nm", C.c_double),
    ]


class huv_socs_opts(C.Structure):
    _fields_ = [("pixel_transfer", C.c_int), ("graph", C.c_int)]


class huv_polygons(C.Structure):
    _fields_ = [("poly_offsets", C.POINTER(C.c_int64)), ("vertex_offsets", C.POINTER(C.c_int64)),
                ("vertices", C.POINTER(C.c_double)), ("values", C.POINTER(C.c_double))]


class huv_socs_tuning(C.Structure):
    _fields_ = [("m", C.c_int), ("n1", C.c_int), ("xnb", C.c_int), ("xlanes", C.c_int),
                ("chunks", C.c_int)]


class huv_clip_group(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nclips", C.c_int),
        ("masks", C.c_void_p),
        ("nplanes", C.c_int),
        ("planes", C.POINTER(C.c_void_p)),
        ("out", C.c_void_p),
    ]


class huv_ilt_config(C.Structure):
    _fields_ = [
        ("cost_kind", C.c_int),
        ("threshold", C.c_double),
        ("steepness", C.c_double),
        ("beta", C.c_double),
        ("tv_weight", C.c_double),
        ("tv_epsilon", C.c_double),
        ("bin_weight", C.c_double),
        ("absorber_re", C.c_double),
        ("absorber_im", C.c_double),
        ("min_feature_nm", C.c_double),
        ("pixel_nm", C.c_double),
        ("gradient", C.c_int),
    ]


P = C.c_void_p
PD = C.POINTER(C.c_double)
PF = C.c_void_p
PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "huv_last_error": (C.c_char_p, []),
    "huv_abi_version": (C.c_int, []),
    "huv_illumination_at": (C.c_double, [C.POINTER(huv_illumination), C.c_double, C.c_double]),
    "huv_sample_source": (C.c_int, [C.POINTER(huv_illumination), C.c_int, C.c_double, PD,
                                    C.c_int64, PI64]),
    "huv_pupil": (C.c_int, [C.POINTER(huv_optics), C.c_double, C.c_double, C.c_double,
                            C.c_double, PD, PD]),
    "huv_select_kernel_count": (C.c_int, [PD, C.c_int64, C.c_double, C.c_int64, C.c_double]),
    "huv_kernels_build": (C.c_int, [C.POINTER(huv_optics), PD, C.c_int64, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                    C.POINTER(P)]),
    "huv_kernels_build_device": (C.c_int, [P, C.POINTER(huv_optics), PD, C.c_int64, C.c_int,
                                           C.c_double, C.c_int, PD, C.c_double, C.c_int,
                                           C.c_double, C.c_int, C.POINTER(P)]),
    "huv_kernels_info_get": (C.c_int, [P, C.POINTER(huv_kernels_info)]),
    "huv_kernels_copy": (C.c_int, [P, PI64, PI64, PD, PD]),
    "huv_kernels_free": (None, [P]),
    "huv_rasterize_rects": (C.c_int, [C.c_int, C.c_double, C.c_int, PD, C.c_double, C.c_double,
                                      PD]),
    "huv_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "huv_ctx_destroy": (None, [P]),
    "huv_ctx_set_stream": (C.c_int, [P, P]),
    "huv_ctx_stream": (P, [P]),
    "huv_ctx_own_stream": (P, [P]),
    "huv_ctx_synchronize": (C.c_int, [P]),
    "huv_ctx_set_profiling": (C.c_int, [P, C.c_int]),
    "huv_ctx_stage_times": (C.c_int, [P, PD, PI64, C.c_int, C.c_int]),
    "huv_kset_upload": (C.c_int, [P, C.c_int, C.c_int, C.c_int64, PI64, PD, PD, C.c_double,
                                  C.POINTER(P)]),
    "huv_kset_destroy": (None, [P]),
    "huv_kset_shape": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "huv_socs_images": (C.c_int, [P, C.c_int, C.c_int, PF, C.c_int, C.POINTER(P), PF,
                                  C.POINTER(huv_socs_opts)]),
    "huv_socs_images_host": (C.c_int, [P, C.c_int, C.c_int, PF, C.c_int, C.POINTER(P), PF,
                                       C.POINTER(huv_socs_opts)]),
    "huv_socs_images_noalloc": (C.c_int, [P, C.c_int, C.c_int, PF, C.c_int, C.POINTER(P), PF,
                                          C.c_int]),
    "huv_socs_images_complex": (C.c_int, [P, C.c_int, C.c_int, PF, C.c_int, C.POINTER(P), PF,
                                          C.POINTER(huv_socs_opts)]),
    "huv_socs_images_spectrum": (C.c_int, [P, C.c_int, C.c_int, PD, C.c_int, C.POINTER(P),
                                           PF]),
    "huv_band_spectrum": (C.c_int, [P, C.c_int, C.c_int, C.c_int, PF, PF, C.c_int]),
    "huv_band_spectrum_rects": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_int, PI64, PD, PD,
                                          PF]),
    "huv_socs_images_rects": (C.c_int, [P, C.c_int, C.c_int, C.c_double, PI64, PD, PD, C.c_int,
                                        C.POINTER(P), PF, C.POINTER(huv_socs_opts)]),
    "huv_band_spectrum_features": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_int, PI64, PD,
                                             C.POINTER(huv_polygons), PD, PF]),
    "huv_socs_images_features": (C.c_int, [P, C.c_int, C.c_int, C.c_double, PI64, PD,
                                           C.POINTER(huv_polygons), PD, C.c_int, C.POINTER(P), PF,
                                           C.POINTER(huv_socs_opts)]),
    "huv_socs_images_batch": (C.c_int, [P, C.c_int, C.POINTER(huv_clip_group),
                                        C.POINTER(huv_socs_opts)]),
    "huv_ctx_set_autotune": (C.c_int, [P, C.c_int]),
    "huv_ctx_tuning_get": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(huv_socs_tuning), C.POINTER(C.c_int)]),
    "huv_ctx_tuning_set": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(huv_socs_tuning)]),
    "huv_ilt_create": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(P), PD, PD,
                                 C.POINTER(huv_ilt_config), C.POINTER(P)]),
    "huv_ilt_destroy": (None, [P]),
    "huv_unet_create": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]),
    "huv_unet_destroy": (None, [P]),
    "huv_unet_num_params": (C.c_int, [P, PI64]),
    "huv_unet_get_params": (C.c_int, [P, PF]),
    "huv_unet_set_params": (C.c_int, [P, PF]),
    "huv_unet_forward": (C.c_int, [P, C.c_int, C.c_int, C.c_int, PF, PF]),
    "huv_unet_loss_grad": (C.c_int, [P, C.c_int, C.c_int, C.c_int, PF, PF, PD, PF]),
    "huv_unet_adam": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.c_double]),
    "huv_ilt_cost_grad": (C.c_int, [P, C.c_int, PF, PF, PD, PF]),
    "huv_ilt_optimize": (C.c_int, [P, C.c_int, PF, PF, C.c_int, C.c_double, C.c_double, C.c_int,
                                   PD, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "huv_ilt_optimize_to": (C.c_int, [P, C.c_int, PF, PF, C.c_int, C.c_double, C.c_double,
                                      C.c_int, PD, PD, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "huv_ilt_run_stage": (C.c_int, [P, C.c_int, PF, PF, C.c_int, C.c_double, C.c_double, C.c_int,
                                    PD, PD, PD, C.c_int, C.POINTER(C.c_int), C.c_int, PF,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "huv_ilt_set_binarization": (C.c_int, [P, C.c_double]),
    "huv_ilt_mask": (C.c_int, [P, C.c_int, PF, PF]),
}

_lib = None


def load():
    """Load libhuv.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m highuvlith_amd.build` "
            "(the GPU path has no fallback)")
    # torch ships its own libamdhip64.so.7 (same SONAME as /opt/rocm's).  Load
    # it first so the process holds exactly one HIP runtime, shared by torch's
    # allocator/streams and this library's kernels.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != HUV_OK:
        msg = load().huv_last_error()
        raise HuvError(status, msg.decode() if msg else "")


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array."""
    return a.ctypes.data_as(PD)


def i64ptr(a):
    return a.ctypes.data_as(PI64)
