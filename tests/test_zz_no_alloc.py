"""GPU: tkmi_dot_noalloc, the entry point that allocates nothing (its partial sums
live in a static device array), against tkmi_dot on the same inputs: same kernels and
summation order, so the same bits.  Two calls in a row, the second on a smaller input,
so a stale partial from the first call would show."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from tkmi import _lib, tensors
    tensors._device()
    return _lib


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dot_noalloc_matches_dot(lib, dtype):
    L = lib.load()
    dt = lib.TKMI_F64 if dtype == torch.float64 else lib.TKMI_F32
    s = torch.cuda.current_stream().cuda_stream
    work = torch.empty(1025, dtype=torch.float64, device="cuda")
    res = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(3)
    # the first call fills all 1024 partials; the second uses 7 of them
    for n in (5_000_011, 12_345, 0):
        a = torch.from_numpy(rng.standard_normal(n)).to(dtype).cuda()
        b = torch.from_numpy(rng.standard_normal(n)).to(dtype).cuda()
        lib.check(L.tkmi_dot(dt, a.data_ptr(), b.data_ptr(), n, work.data_ptr(), s), "dot")
        lib.check(L.tkmi_dot_noalloc(dt, a.data_ptr(), b.data_ptr(), n, res.data_ptr(), s), "dot_noalloc")
        torch.cuda.synchronize()
        got, ref = float(res[0].item()), float(work[0].item())
        assert got == ref, (n, got, ref)
        exact = math.fsum((a.double().cpu().numpy() * b.double().cpu().numpy()).tolist()) if n else 0.0
        assert abs(got - exact) <= 1e-12 * max(1.0, math.sqrt(n)), (n, got, exact)


def test_dot_noalloc_rejects_bad_arguments(lib):
    L = lib.load()
    a = torch.ones(8, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert L.tkmi_dot_noalloc(lib.TKMI_F64, a.data_ptr(), a.data_ptr(), 8, None, s) == 1
    res = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert L.tkmi_dot_noalloc(lib.TKMI_F64, a.data_ptr(), a.data_ptr(), -1, res.data_ptr(), s) == 1
    assert L.tkmi_dot_noalloc(7, a.data_ptr(), a.data_ptr(), 8, res.data_ptr(), s) == 6
