"""Builds the HIP engine into cppirlba_amd/lib/libirlba_hip.so (gfx950).

hipcc compiles every HIP source of csrc/ to an object in _build/ (in
parallel), g++ the host-only sources (the host SVD, built for two x86 ISA
levels with run-time dispatch), and hipcc links one shared library.  All HIP code is compiled with packed FP32
instructions disabled (-target-feature -packed-fp32-ops).  The libtorch
autograd binding (csrc/torch_svd.cpp, host code only) is compiled with g++
against torch's headers into lib/libirlba_torch.so, linked to the engine.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(os.path.dirname(HERE), "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libirlba_hip.so")
TORCH_LIB = os.path.join(LIBDIR, "libirlba_torch.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("IRLBA_HIP_ARCH", "gfx950")

SOURCES = [
    "engine.cpp",
    "kernels_sparse.hip",
    "kernels_dense.hip",
    "kernels_blas.hip",
    "kernels_generate.hip",
    "kernels_backward.hip",
    "kernels_spmm.hip",
    "noalloc.hip",
]

# Host-only C++ (no device code), compiled by g++.
HOST_SOURCES = ["host_svd.cpp"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src in HOST_SOURCES:
        cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [hipcc(), *COMMON]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, SOURCES + HOST_SOURCES))
    for obj, err in results:
        if verbose and err.strip():
            sys.stderr.write(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_torch()
    return LIB


def build_torch():
    """g++ the libtorch autograd binding against torch's headers and the engine."""
    import torch
    from torch.utils import cpp_extension

    tlib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    # torch's HIP stream header (the binding asks for torch's current stream)
    # needs the HIP runtime's declarations; no HIP call is made from here.
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-D__HIP_PLATFORM_AMD__", "-Wall", "-Wno-unused-parameter",
           "-I", INCLUDE, "-I", os.path.join(HERE, "include"),
           *[f"-I{p}" for p in cpp_extension.include_paths()], "-I", os.path.join(rocm, "include"),
           os.path.join(CSRC, "torch_svd.cpp"),
           "-o", TORCH_LIB + ".tmp",
           "-L", LIBDIR, "-lirlba_hip", "-Wl,-rpath,$ORIGIN",
           "-L", tlib, "-ltorch", "-ltorch_cpu", "-lc10_hip", "-lc10", f"-Wl,-rpath,{tlib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on torch_svd.cpp:\n{r.stderr}")
    os.replace(TORCH_LIB + ".tmp", TORCH_LIB)
    return TORCH_LIB


if __name__ == "__main__":
    print(build(verbose=True))
