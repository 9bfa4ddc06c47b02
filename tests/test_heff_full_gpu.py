"""GPU: the two-site H_eff workload at its full size (U1, D = 2052, 42 sector GEMMs,
7.8 GFLOP) and its precision variants, each against the oracle.

  * FP64 chain, with the tile configurations the autotuner keeps (split-K on ·R), against
    the oracle's host replay of the same kernel sequence (the C permute loops over the same
    tables, one OpenBLAS GEMM per coupled sector, fused contractions replayed unfused):
    relative Frobenius error ≤ 1e-13; the captured graph replays to the eager result bit
    for bit.
  * FP32 and 3×BF16 GEMMs, on the FP64 chain and on the float32 chain (operands and x
    rounded to FP32 once, every intermediate FP32): against the dense einsum at D = 96, and
    against the FP64 chain at D = 2052, relative Frobenius error ≤ 1e-6.
  * svd_trunc(truncrank 1024 ∩ trunctol 1e-10) of the two-site result as (V⊗P) ← (P⊗V)′
    (src/factorizations/truncation.jl:171-322): singular values per block against numpy's
    LAPACK, the kept set and ε against a sequential restatement of the truncation; the
    3×BF16 result keeps the same set, and its ε differs from FP64's by at most
    ‖y_bf16 − y_64‖_F (Mirsky: the singular values of a perturbed matrix move by at most
    the perturbation's Frobenius norm, in 2-norm over all of them).
"""
import numpy as np
import pytest
import torch

from oracle import abelian as O
from helpers import oracle_legs, dense_heff

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def heff(T):
    """The FP64 chain at D = 2052 (seed 1) with autotuned tiles, its trace, and the
    graph-captured reduced-precision results."""
    from tkmi import _lib
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=2048)
    assert V.dim() == 2052
    H = TwoSiteHeff(V, P, W, np.random.default_rng(1))
    T.start_trace()
    H.apply()
    trace = T.stop_trace()
    T.autotune_trace(trace, precisions=(_lib.PREC_F64, _lib.PREC_F32, _lib.PREC_BF16X3))
    y64 = H.apply()
    torch.cuda.synchronize()
    out = {"H": H, "trace": trace, "y64": y64.numpy().copy(), "space": y64.space}
    for name, prec in (("fp32", _lib.PREC_F32), ("bf16x3", _lib.PREC_BF16X3)):
        H2 = H.astype(torch.float64)
        H2.capture(precision=prec)
        H2.replay()
        torch.cuda.synchronize()
        out[name] = H2.graph_y.numpy().copy()
        out[name + "_t"] = H2.graph_y
    # the float32 chain, its GEMMs tuned on FP32 data
    H32 = H.astype(torch.float32)
    T.start_trace()
    H32.apply(precision=_lib.PREC_F32)
    T.autotune_trace(T.stop_trace(), precisions=(_lib.PREC_F32, _lib.PREC_BF16X3))
    for name, prec in (("fp32_f32data", _lib.PREC_F32), ("bf16x3_f32data", _lib.PREC_BF16X3)):
        H32.capture(precision=prec)
        for _ in range(2):
            H32.replay()
            torch.cuda.synchronize()
            y = H32.graph_y
            assert y.dtype == torch.float32
            if name in out:
                np.testing.assert_array_equal(y.numpy(), out[name])   # replays agree bit for bit
            out[name] = y.numpy().copy()
        # and the eager float32 chain gives the graph's bits
        H32.outs = [None] * 4
        np.testing.assert_array_equal(H32.apply(precision=prec).numpy(), out[name])
    return out


def test_full_fp64_chain_vs_host_replay(T, heff):
    from oracle import replay as R
    H, trace = heff["H"], heff["trace"]
    splits = []
    from tkmi import _lib
    import ctypes as C
    lib = _lib.load()
    for op in trace:
        if op[0] == "gemm":
            cfg, sk, slots = C.c_int32(), C.c_int64(), C.c_int64()
            lib.tkmi_gemm_plan_query(op[1].handle, _lib.PREC_F64, C.byref(cfg), None)
            lib.tkmi_gemm_plan_splitk(op[1].handle, cfg.value, C.byref(sk), C.byref(slots))
            splits.append(sk.value)
    assert max(splits) > 1  # the ·R GEMM runs split-K: its reduction is covered
    init = {t.data.data_ptr(): t.numpy().copy() for t in (H.L, H.M1, H.M2, H.R, H.x)}
    bufs = R.cpu_replay(R.prepare_cpu_replay(trace), init, 16)
    y_cpu = bufs[H.outs[-1].data.data_ptr()]
    assert relerr(heff["y64"], y_cpu) <= 1e-13
    # the captured FP64 graph replays to the eager result, bit for bit
    H.capture(precision=_lib.PREC_F64)
    for _ in range(2):
        H.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(H.graph_y.numpy(), heff["y64"])


def test_full_reduced_precision_vs_fp64(heff):
    for name in ("fp32", "bf16x3", "fp32_f32data", "bf16x3_f32data"):
        err = relerr(heff[name], heff["y64"])
        assert err <= 1e-6, (name, err)


def test_reduced_precision_vs_dense_d96(T):
    from tkmi import _lib
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=96, qmax=4)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(1))

    def dense(t):
        return O.to_dense(t.numpy(), *oracle_legs(t.space), O.U1Kind)

    ref = dense_heff(dense(H.L), dense(H.x), dense(H.M1), dense(H.M2), dense(H.R))
    for prec in (_lib.PREC_F32, _lib.PREC_BF16X3):
        H.outs = [None] * 4
        y = H.apply(precision=prec)
        err = relerr(dense(y), ref)
        assert err <= 1e-6, (prec, err)
        y32 = H.astype(torch.float32).apply(precision=prec)
        assert y32.dtype == torch.float32
        err = relerr(O.to_dense(y32.numpy().astype(np.float64), *oracle_legs(y32.space), O.U1Kind), ref)
        assert err <= 1e-6, (prec, "float32 chain", err)


def _sequential_truncation(sv, howmany, atol):
    """truncrank(howmany) ∩ trunctol(atol) by a scan of all values sorted descending
    (ties by position), dim(c) = 1 for U1: returns (kept count per block, ε)."""
    allv = sorted(((s, c, i) for c, v in sv.items() for i, s in enumerate(v)), key=lambda x: (-x[0], x[1], x[2]))
    kept = {c: 0 for c in sv}
    disc = 0.0
    for rank, (s, c, i) in enumerate(allv):
        if rank < howmany and s > atol:
            kept[c] += 1
        else:
            disc += s * s
    return kept, float(np.sqrt(disc))


def test_svd_trunc_of_two_site_result(T, heff):
    from tkmi.factorizations import svd_trunc
    y64 = T.TensorMap.from_numpy(heff["y64"], heff["space"])
    y64t = T.permute(y64, (0, 1), (2, 3), copy=True)   # (V⊗P) ← (P⊗V)′
    ybft = T.permute(heff["bf16x3_t"], (0, 1), (2, 3), copy=True)
    Ut, Sk, Vtt, eps64 = svd_trunc(y64t, 1024, 1e-10)
    _, Skb, _, epsbf = svd_trunc(ybft, 1024, 1e-10)
    blocks = {c: b for c, b in y64t.blocks_host().items() if b.size}
    sv = {c: np.linalg.svd(b, compute_uv=False) for c, b in blocks.items()}
    kept, eps_ref = _sequential_truncation(sv, 1024, 1e-10)
    got = {c: int(v.numel()) for c, v in Sk.items()}
    assert got == {c: k for c, k in kept.items() if k}
    assert sum(got.values()) == 1024
    for c, v in Sk.items():
        s = v.cpu().numpy()
        assert np.max(np.abs(s - sv[c][:len(s)])) <= 1e-12 * sv[c][0], c
    assert abs(eps64 - eps_ref) <= 1e-12 * eps_ref
    # 3×BF16: same kept set, ε within the perturbation's Frobenius norm
    assert {c: int(v.numel()) for c, v in Skb.items()} == got
    pert = np.linalg.norm(heff["bf16x3"] - heff["y64"])
    assert abs(epsbf - eps64) <= pert * (1 + 1e-6) + 1e-12 * eps64, (epsbf, eps64, pert)
