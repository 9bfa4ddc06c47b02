"""Build libtkmi.so (HIP, gfx950) in-tree: tensorkit.jl_amd/tkmi/libtkmi.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "tkmi", "libtkmi.so")
SOURCES = ["tkmi_runtime.hip", "tkmi_transform.hip", "tkmi_gemm.hip", "tkmi_reduce.hip", "tkmi_svd.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions stay off (hipcc emits them from plain float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# rocSOLVER/rocBLAS only for the FP64 block SVD of svd_trunc
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build_library(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES] + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtkmi.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
