"""Build the in-tree shared libraries.

libhuv.so (product): the C ABI of include/huv.h — host engine (FP64 C++,
compiled with g++) plus the HIP kernels for gfx950 (hipcc).  Always
compiled from source when called.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhuv.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) stays off in every device build.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]

HOST_SRC = ["host/optics_source.cpp", "host/linalg.cpp", "host/kernels_build.cpp"]
HIP_SRC = ["hip/socs_kernels.hip", "hip/ilt_kernels.hip", "hip/unet_kernels.hip",
           "hip/mask_kernels.hip", "hip/kbuild_kernels.hip", "capi.cpp", "capi_kbuild.cpp",
           "capi_ilt.cpp", "capi_unet.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SRC:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SRC:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        jobs.append([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
