"""ORACLE — test infrastructure only.

Python binding of oracle/liboracle.so, the FP64 CPU restatement of the
reference's SOCS image path and ILT adjoint (see oracle/oracle.cpp for the
file:line each function follows).  Only tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py may import this package, and only as the
checker / baseline — never as the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_fft2": (None, [PD, C.c_int, C.c_int, C.c_int]),
        "orc_ifft2_pruned": (None, [PD, C.c_int, PI64, C.c_int64]),
        "orc_dft2_naive": (None, [PD, PD, C.c_int, C.c_int]),
        "orc_socs_intensity": (C.c_int, [C.c_int, C.c_int, C.c_int64, PI64, PI64, C.c_int64, PD,
                                         PD, PD, PD, C.c_int]),
        "orc_normalize_flare": (None, [PD, C.c_int64, C.c_double, C.c_double]),
        "orc_polygon_spectrum": (C.c_int, [C.c_int, C.c_double, C.c_int, PD, C.c_double,
                                           C.c_double, PD]),
        "orc_polygon_is_simple": (C.c_int, [C.c_int, PD]),
        "orc_mask_spectrum": (C.c_int, [C.c_int, C.c_double, C.c_int, PD, C.c_int, C.c_double,
                                        C.c_double, PD]),
        "orc_pixel_transfer": (None, [PD, C.c_int]),
        "orc_abbe_image": (None, [C.c_int, C.c_double, PD, PD, C.c_int64, C.c_int, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, PD]),
        "orc_ilt_create": (C.c_void_p, [C.c_int, PD, C.c_int, C.c_double, C.c_double, C.c_int, PD,
                                        PD, PD, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int]),
        "orc_ilt_destroy": (None, [C.c_void_p]),
        "orc_ilt_set_bin_weight": (None, [C.c_void_p, C.c_double]),
        "orc_ilt_cost": (C.c_double, [C.c_void_p, PD]),
        "orc_ilt_regularization": (C.c_double, [C.c_void_p, PD, PD]),
        "orc_ilt_cost_grad": (C.c_double, [C.c_void_p, PD, PD]),
        "orc_ilt_mask": (None, [C.c_void_p, PD, PD]),
        "orc_ilt_image": (None, [C.c_void_p, PD, C.c_int, PD]),
        "orc_ilt_run_stage": (C.c_int, [C.c_void_p, PD, C.c_int, C.c_double, C.c_double, C.c_int,
                                        PD, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(PD)


def _c2(a):
    """complex (..) -> contiguous float64 (.., 2) view buffer."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(a.shape + (2,)).copy()


def _from2(b):
    return b[..., 0] + 1j * b[..., 1]


# ---------------------------------------------------------------------------
def fft2(x, inverse=False):
    """Unnormalised forward / 1/N inverse 2-D DFT (fft2d.rs:151-160)."""
    b = _c2(x)
    ny, nx = x.shape
    load().orc_fft2(_d(b), ny, nx, int(inverse))
    return _from2(b)


def ifft2_pruned(x, rows):
    b = _c2(x)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    load().orc_ifft2_pruned(_d(b), x.shape[0], r.ctypes.data_as(PI64), len(r))
    return _from2(b)


def dft2_naive(x, inverse=False):
    b = _c2(x)
    o = np.zeros_like(b)
    load().orc_dft2_naive(_d(b), _d(o), x.shape[0], int(inverse))
    return _from2(o)


def socs_intensity(n, support, nonzero_rows, eigenvalues, values, spectrum, threads=1):
    """sum_k lambda_k |IFFT(K_k . M)|^2 in the structure of aerial.rs:1682-1721."""
    sup = np.ascontiguousarray(support, dtype=np.int64)
    rows = np.ascontiguousarray(nonzero_rows, dtype=np.int64)
    eig = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    vals = _c2(values)
    spec = _c2(spectrum)
    out = np.zeros((n, n))
    load().orc_socs_intensity(n, len(eig), len(sup), sup.ctypes.data_as(PI64),
                              rows.ctypes.data_as(PI64), len(rows), _d(eig), _d(vals), _d(spec),
                              _d(out), int(threads))
    return out


def normalize_flare(img, divisor, flare):
    out = np.ascontiguousarray(img, dtype=np.float64).copy()
    load().orc_normalize_flare(_d(out), out.size, float(divisor), float(flare))
    return out


def socs_image(ks, spectrum, flare, threads=1):
    """compute_from_spectrum (aerial.rs:1805-1818) for a host kernel set."""
    img = socs_intensity(ks.n, ks.support, ks.nonzero_rows, ks.eigenvalues, ks.values, spectrum,
                         threads)
    return normalize_flare(img, ks.normalization_divisor, flare)


# Mask features for mask_spectrum: ("rect", x, y, w, h) | ("lines", cd, pitch,
# orientation, offset) | ("array", w, h, px, py, ox, oy)
def mask_spectrum(n, pixel_nm, features, dark_field=False, absorber=0j):
    rec = np.zeros((max(len(features), 1), 8))
    for i, f in enumerate(features):
        kind = {"rect": 0, "lines": 1, "array": 2}[f[0]]
        vals = list(f[1:])
        rec[i, 0] = kind
        rec[i, 1:1 + len(vals)] = vals
    out = np.zeros((n, n, 2))
    st = load().orc_mask_spectrum(n, pixel_nm, len(features), _d(rec), int(dark_field),
                                  complex(absorber).real, complex(absorber).imag, _d(out))
    if st != 0:
        raise RuntimeError("overlapping features beyond the exact cell budget")
    return _from2(out)


def polygon_spectrum(n, pixel_nm, vertices, value=1.0, spectrum=None):
    """Add value x the exact spectrum of a simple polygon (clipped to the field) to
    `spectrum` (n x n complex, FFT layout; zeros if None); raises for a
    non-simple polygon."""
    xy = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 2))
    out = np.zeros((n, n, 2)) if spectrum is None else _c2(spectrum)
    v = complex(value)
    st = load().orc_polygon_spectrum(n, pixel_nm, len(xy), _d(xy), v.real, v.imag, _d(out))
    if st != 0:
        raise RuntimeError("self-intersecting polygon")
    return _from2(out)


def pixel_transfer(spec):
    b = _c2(spec)
    load().orc_pixel_transfer(_d(b), spec.shape[0])
    return _from2(b)


def abbe_image(n, pixel_nm, spectrum, points, optics_kind, na, medium_index, obscuration,
               transmission, defocus_nm, wavelength_nm):
    """Direct Abbe sum (tests/imaging_validation.rs:167-198), absolute intensity."""
    spec = _c2(spectrum)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros((n, n))
    load().orc_abbe_image(n, pixel_nm, _d(spec), _d(pts), len(pts), optics_kind, na,
                          medium_index, obscuration, transmission, defocus_nm, wavelength_nm,
                          _d(out))
    return out


class IltProblem:
    """ilt.rs:369-739 in FP64.  conditions: list of (dose, weight, DenseKernelSet)."""

    def __init__(self, n, target, conditions, flare, cost="resist", threshold=0.25,
                 steepness=50.0, beta=4.0, tv_weight=0.01, tv_eps=0.01, bin_weight=0.0,
                 absorber=0j, min_feature_nm=0.0, pixel_nm=1.0, proxy=False):
        self.n = n
        lib = load()
        tgt = np.ascontiguousarray(target, dtype=np.float64)
        rec = np.array([[d, w, len(ks.eigenvalues)] for d, w, ks in conditions], dtype=np.float64)
        lam = np.ascontiguousarray(np.concatenate([ks.eigenvalues for _, _, ks in conditions]))
        kern = _c2(np.concatenate([ks.kernels.reshape(len(ks.eigenvalues), -1)
                                   for _, _, ks in conditions]))
        self._h = lib.orc_ilt_create(n, _d(tgt), 1 if cost == "resist" else 0, threshold,
                                     steepness, len(conditions), _d(rec), _d(lam), _d(kern),
                                     flare, beta, tv_weight, tv_eps, bin_weight,
                                     complex(absorber).real, complex(absorber).imag,
                                     min_feature_nm, pixel_nm, int(proxy))
        self.lr = 0.5

    def __del__(self):
        try:
            load().orc_ilt_destroy(self._h)
        except Exception:
            pass

    def cost(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        return load().orc_ilt_cost(self._h, _d(t))

    def cost_and_gradient(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros_like(t)
        c = load().orc_ilt_cost_grad(self._h, _d(t), _d(g))
        return c, g

    def regularization(self, m):
        mm = np.ascontiguousarray(m, dtype=np.float64)
        g = np.zeros_like(mm)
        c = load().orc_ilt_regularization(self._h, _d(mm), _d(g))
        return c, g

    def mask(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        m = np.zeros_like(t)
        load().orc_ilt_mask(self._h, _d(t), _d(m))
        return m

    def image(self, m, condition=0):
        mm = np.ascontiguousarray(m, dtype=np.float64)
        out = np.zeros_like(mm)
        load().orc_ilt_image(self._h, _d(mm), condition, _d(out))
        return out

    def set_bin_weight(self, w):
        load().orc_ilt_set_bin_weight(self._h, w)

    def run_stage(self, theta, max_iterations, learning_rate=0.5, convergence_tol=1e-4,
                  use_cg=True):
        t = np.ascontiguousarray(theta, dtype=np.float64).copy()
        hist = np.zeros(max_iterations + 1)
        it = C.c_int()
        term = load().orc_ilt_run_stage(self._h, _d(t), max_iterations, learning_rate,
                                        convergence_tol, int(use_cg), _d(hist), C.byref(it))
        return t, hist[:it.value + 1], term
