"""CPU: the transform and grouped-GEMM kernels, compiled from their unchanged sources as
host C++ against a CPU stand-in of the HIP runtime (tests/emu/hip/hip_runtime.h: a
workgroup's threads are host threads, __syncthreads a barrier, each MFMA builtin a
rendezvous of the wave's 64 lanes with the documented lane layouts) and run under
AddressSanitizer: an out-of-bounds read or write anywhere in a kernel or its host
plumbing fails the run.  tests/emu/emu_check.cpp drives every kernel class through the C
ABI on small random problems and checks each result against a direct loop (details in its
header).  Needs the ROCm clang++ for the host (its ASan runtime); skipped without it."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(not os.path.exists(CLANG), reason="host clang++ with AddressSanitizer not available")
def test_kernels_bounds_checked_on_cpu():
    emu = os.path.join(ROOT, "tests", "emu")
    csrc = os.path.join(ROOT, "tensorkit.jl_amd", "csrc")
    tmp = tempfile.mkdtemp(prefix="tkmi_emu_")
    try:
        jobs = [
            (["-I", emu, "-x", "c++", os.path.join(csrc, "tkmi_transform.hip")], "transform.o"),
            (["-I", emu, "-x", "c++", os.path.join(csrc, "tkmi_gemm.hip")], "gemm.o"),
            (["-I", emu, "-x", "c++", os.path.join(csrc, "tkmi_jacobi.hip")], "jacobi.o"),
            (["-I", emu, os.path.join(emu, "emu_runtime.cpp")], "runtime.o"),
            ([os.path.join(emu, "emu_check.cpp")], "check.o"),
        ]
        procs = [(obj, subprocess.Popen([CLANG] + FLAGS + ["-c"] + args + ["-o", os.path.join(tmp, obj)],
                                        stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
                 for args, obj in jobs]
        for obj, p in procs:
            out, err = p.communicate(timeout=900)
            assert p.returncode == 0, f"{obj}: {err[-4000:]}"
        exe = os.path.join(tmp, "emu_check")
        r = subprocess.run([CLANG, "-fsanitize=address"] + [os.path.join(tmp, o) for _, o in jobs]
                           + ["-lpthread", "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=1200, env=env)
        assert r.returncode == 0 and "EMU OK" in r.stdout, (r.stdout[-4000:] + r.stderr[-8000:])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
