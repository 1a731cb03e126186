"""The C ABI library loads on a CPU-only host and exports every symbol
include/huv.h declares; host-engine error paths return status + message."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import highuvlith_amd as H
from highuvlith_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(huv_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_python_signature():
    assert declared_functions() <= set(_lib.SIGNATURES)


def test_abi_version_and_errors_are_reported():
    lib = _lib.load()
    assert lib.huv_abi_version() == 102
    o = H.ProjectionOptics(0.75).to_c()
    pts = np.array([[0.0, 0.0, 1.0]])
    h = C.c_void_p()
    st = lib.huv_kernels_build(C.byref(o), _lib.dptr(pts), 1, 32, 4.0, 0.0, 157.63, 0, 1.0, -1,
                               C.byref(h))
    assert st == 1 and b"max_kernels" in lib.huv_last_error()
    st = lib.huv_kernels_build(C.byref(o), _lib.dptr(pts), 1, 32, -4.0, 0.0, 157.63, 4, 1.0, -1,
                               C.byref(h))
    assert st == 1 and b"grid" in lib.huv_last_error()
    bad = H.ProjectionOptics(0.75).to_c()
    bad.na = 1.2
    st = lib.huv_kernels_build(C.byref(bad), _lib.dptr(pts), 1, 32, 4.0, 0.0, 157.63, 4, 1.0, -1,
                               C.byref(h))
    assert st == 1 and b"numerical_aperture" in lib.huv_last_error()


def test_no_gpu_means_a_device_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(H.HuvError) as e:
        H.DeviceContext(0)
    assert e.value.kind == "DeviceError"
