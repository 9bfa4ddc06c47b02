"""Build libtkmi.so (HIP, gfx950) in-tree: tensorkit.jl_amd/tkmi/libtkmi.so."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "tkmi", "libtkmi.so")
SOURCES = ["tkmi_runtime.hip", "tkmi_transform.hip", "tkmi_gemm.hip", "tkmi_reduce.hip", "tkmi_svd.hip",
           "tkmi_jacobi.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (hipcc emits them from plain float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# rocSOLVER/rocBLAS only for the FP64 block SVD of svd_trunc
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build_library(verbose=False):
    """Compile every source to an object (in parallel, one hipcc per file), then link."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    objdir = tempfile.mkdtemp(prefix="tkmi_build_")
    try:
        procs = []
        for src in SOURCES:
            obj = os.path.join(objdir, src.replace(".hip", ".o"))
            cmd = [hipcc] + HIPCC_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)]
            procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                                     text=True)))
        failed = False
        for src, obj, pr in procs:
            out, err = pr.communicate()
            if pr.returncode != 0:
                sys.stderr.write(out + err)
                failed = True
            elif verbose:
                sys.stderr.write(err)
        if failed:
            raise RuntimeError("hipcc failed building libtkmi.so")
        cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
            [obj for _, obj, _ in procs] + LIBS
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libtkmi.so")
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
