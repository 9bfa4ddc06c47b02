"""GPU: entry points a Julia binding calls directly, each checked through the C ABI.

  * hipGraph capture (tkmi_graph_begin / _end / _launch) of a permute → grouped GEMM →
    permute chain, replayed several times with eager work in between; every replay equals
    the eager chain bit for bit (integer-valued data: exact in FP64);
  * tkmi_gemm_plan_autotune: the configuration it keeps gives the same bits as the same
    configuration forced with tkmi_gemm_plan_set_tile, and both equal numpy (integer data);
  * planarcontract! / planartrace! (bosonic: src/planar/planaroperations.jl:66-111,150-190)
    against the dense embedding;
  * the transform refuses overlapping source and destination vectors, but runs add!/scale!
    in place (the elementwise plan), before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from tkmi.spaces import HomSpace, ProductSpace, U1Space

from helpers import cfg1_spaces, oracle_legs, kind_of
from oracle import abelian as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _int_randn(W, rng, T, lo=-3, hi=4):
    return T.TensorMap.from_numpy(rng.integers(lo, hi, W.dim()).astype(np.float64), W)


def test_graph_capture_through_c_abi(T):
    """permute(x) → mul!(L, ·) → permute back, captured on a side stream with the C ABI's
    own capture calls and replayed four times; x changes eagerly between replays."""
    from tkmi import _lib
    lib = _lib.load()
    V, P = cfg1_spaces()
    WX = HomSpace(ProductSpace(V, P), ProductSpace(V))
    WL = HomSpace(ProductSpace(V), ProductSpace(V))
    rng = np.random.default_rng(5)
    x = _int_randn(WX, rng, T)
    L = _int_randn(WL, rng, T)
    p1, p2 = (0,), (1, 2)          # x as V ← (P ⊗ V)′: GEMM layout for L·x
    Wx2 = WX.permute(p1, p2)
    Wy2 = HomSpace(WL.codomain, Wx2.domain)
    q1, q2 = (0, 1), (2,)          # back to V ⊗ P ← V
    Wy = Wy2.permute(q1, q2)
    x2 = T.TensorMap.empty(Wx2)
    y2 = T.TensorMap.empty(Wy2)
    y = T.TensorMap.empty(Wy)
    pin = [T.treebraider(Wx2, WX, p1, p2), T.treebraider(Wy, Wy2, q1, q2), T.gemm_plan(y2, L, x2)]

    def chain(stream):
        _lib.check(lib.tkmi_transform_exec(pin[0].handle, _lib.TKMI_F64, x.data.data_ptr(), x2.data.data_ptr(),
                                           1.0, 0.0, stream), "permute x")
        _lib.check(lib.tkmi_gemm_exec(pin[2].handle, _lib.PREC_F64, 1.0, 0.0, L.data.data_ptr(),
                                      x2.data.data_ptr(), y2.data.data_ptr(), stream), "mul!")
        _lib.check(lib.tkmi_transform_exec(pin[1].handle, _lib.TKMI_F64, y2.data.data_ptr(), y.data.data_ptr(),
                                           0.5, 0.0, stream), "permute y")

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = C.c_void_p()
    with torch.cuda.stream(side):
        s = side.cuda_stream
        _lib.check(lib.tkmi_graph_begin(s), "graph begin")
        chain(s)
        _lib.check(lib.tkmi_graph_end(s, C.byref(g)), "graph end")
    torch.cuda.current_stream().wait_stream(side)
    try:
        for rep in range(4):
            x.data.copy_(torch.from_numpy(rng.integers(-3, 4, WX.dim()).astype(np.float64)))
            y.data.fill_(np.nan)
            torch.cuda.synchronize()
            _lib.check(lib.tkmi_graph_launch(g, side.cuda_stream), "graph launch")
            torch.cuda.synchronize()
            got = y.numpy().copy()
            chain(torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            eager = y.numpy()
            np.testing.assert_array_equal(got, eager)
            # and the eager chain is L·x in the dense embedding
            kind = kind_of(WX)
            dx = O.to_dense(x.numpy(), *oracle_legs(WX), kind)
            dL = O.to_dense(L.numpy(), *oracle_legs(WL), kind)
            dy = 0.5 * np.einsum("ab,bpc->apc", dL, dx)
            np.testing.assert_array_equal(O.to_dense(eager, *oracle_legs(Wy), kind), dy)
    finally:
        lib.tkmi_graph_destroy(g)


def test_graph_begin_needs_a_stream(T):
    from tkmi import _lib
    lib = _lib.load()
    assert lib.tkmi_graph_begin(None) == 1  # TKMI_ERR_ARGUMENT


def test_gemm_autotune_keeps_a_configuration_that_is_exact(T):
    """The autotuner times every tile configuration on the plan's own operands; the one it
    keeps reproduces the forced configuration and numpy bit for bit (integer data)."""
    from test_gemm_gpu import _make_groups, _ref
    from tkmi import _lib
    L = _lib.load()
    rng = np.random.default_rng(9)
    shapes = [(290, 272, 24), (130, 70, 33), (65, 63, 17), (1414, 300, 290), (8, 8, 8)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 0, 1], [0, 0, 1, 0, 1])
    A = rng.integers(-4, 5, na).astype(np.float64)
    B = rng.integers(-4, 5, nb).astype(np.float64)
    Cm = rng.integers(-4, 5, nc).astype(np.float64)
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [_lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = _lib.i32(arr[:, 5]), _lib.i32(arr[:, 8])
    h = C.c_void_p()
    _lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                       cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
               "plan")
    try:
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        dC = torch.from_numpy(Cm).cuda()
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(L.tkmi_gemm_plan_autotune(h, _lib.PREC_F64, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), 1, s),
                   "autotune")
        cfg, nt = C.c_int32(), C.c_int64()
        _lib.check(L.tkmi_gemm_plan_query(h, _lib.PREC_F64, C.byref(cfg), C.byref(nt)), "query")
        assert 0 <= cfg.value < 13 and nt.value > 0
        ref = _ref(tab, A, B, Cm, 2.0, -1.0)
        dC.copy_(torch.from_numpy(Cm))
        _lib.check(L.tkmi_gemm_exec(h, _lib.PREC_F64, 2.0, -1.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s),
                   "exec tuned")
        tuned = dC.cpu().numpy()
        np.testing.assert_array_equal(tuned, ref)
        _lib.check(L.tkmi_gemm_plan_set_tile(h, _lib.PREC_F64, 1000 + cfg.value), "force")
        dC.copy_(torch.from_numpy(Cm))
        _lib.check(L.tkmi_gemm_exec(h, _lib.PREC_F64, 2.0, -1.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s),
                   "exec forced")
        np.testing.assert_array_equal(dC.cpu().numpy(), tuned)
    finally:
        L.tkmi_gemm_plan_destroy(h)


def test_planarcontract_dense(T):
    """planarcontract!(C, A, pA, B, pB, pAB) for U1 (bosonic: planar = braided order),
    C[a p; b] = Σ_y L[a; y]·A[y p; b], against the dense embedding."""
    V, P = cfg1_spaces()
    WA = HomSpace(ProductSpace(V, P), ProductSpace(V))
    WL = HomSpace(ProductSpace(V), ProductSpace(V))
    rng = np.random.default_rng(21)
    A = T.TensorMap.randn(WA, rng)
    L = T.TensorMap.randn(WL, rng)
    Cs = T.TensorMap.randn(WA, rng)
    c0 = Cs.numpy()
    T.planarcontract_(Cs, L, ((0,), (1,)), A, ((0,), (1, 2)), ((0, 1), (2,)), 0.7, 0.3)
    kind = kind_of(WA)
    dA = O.to_dense(A.numpy(), *oracle_legs(WA), kind)
    dL = O.to_dense(L.numpy(), *oracle_legs(WL), kind)
    ref = 0.7 * np.einsum("ay,ypb->apb", dL, dA) + 0.3 * O.to_dense(c0, *oracle_legs(WA), kind)
    got = O.to_dense(Cs.numpy(), *oracle_legs(WA), kind)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_planartrace_dense(T):
    """planartrace!(C, A, p, q): the trace of the adjacent pair (legs 1, 2) of
    A: V ⊗ W ⊗ W′ ← V, against the dense index sum."""
    V = U1Space((0, 3), (1, 2), (-1, 2))
    Wl = U1Space((0, 2), (1, 1), (2, 1))
    WA = HomSpace(ProductSpace(V, Wl, Wl.dual()), ProductSpace(V))
    rng = np.random.default_rng(22)
    A = T.TensorMap.randn(WA, rng)
    p, q = ((0,), (3,)), ((1,), (2,))
    Cs = T.TensorMap.randn(WA.select(*p), rng)
    c0 = Cs.numpy()
    T.planartrace_(Cs, A, p, q, 1.5, -0.5)
    kind = kind_of(WA)
    dA = O.to_dense(A.numpy(), *oracle_legs(WA), kind)
    ref = 1.5 * np.einsum("awwb->ab", dA) - 0.5 * O.to_dense(c0, *oracle_legs(Cs.space), kind)
    got = O.to_dense(Cs.numpy(), *oracle_legs(Cs.space), kind)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_overlapping_vectors_are_refused(T):
    """A permute whose destination overlaps its source would race between tiles: the
    call fails with TKMI_ERR_ARGUMENT and nothing is launched; add!/scale! in place and
    α = 0 (no reads) are allowed."""
    V = U1Space((0, 4), (1, 3), (-1, 3))
    W = HomSpace(ProductSpace(V, V), ProductSpace())
    rng = np.random.default_rng(23)
    A = T.TensorMap.randn(W, rng)
    a0 = A.numpy()
    with pytest.raises(ValueError, match="overlap"):
        T.permute_(A, A, (1, 0), ())
    # a view into the middle of the same storage overlaps too
    big = torch.from_numpy(np.concatenate([a0, a0])).cuda()
    X = T.TensorMap(big[:W.dim()], W)
    Y = T.TensorMap(big[W.dim() // 2: W.dim() // 2 + W.dim()], W)
    with pytest.raises(ValueError, match="overlap"):
        T.permute_(Y, X, (1, 0), ())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(A.numpy(), a0)        # untouched
    T.scale_(A, 0.5)
    T.add_(A, A, 1.0, 1.0)
    np.testing.assert_array_equal(A.numpy(), 0.5 * a0 + 0.5 * a0)
    T.permute_(A, A, (1, 0), (), 0.0, 2.0)              # α = 0: nothing read
    np.testing.assert_array_equal(A.numpy(), 2.0 * a0)


def test_calib_mfma_f64():
    """tkmi_calib_mfma_f64: the per-wave sum of its outputs is the closed form
    16·iters·Σ_k (Σ_i a_{i+16k})(Σ_j b_{j+16k}) over the wave's lanes l (a = 1 + t,
    b = 1 − t, t the thread index): four accumulators, four MFMAs each per round, and
    every MFMA adds A·B (the sum over the output fragment does not depend on its lane
    layout).  Bad arguments are rejected."""
    from tkmi import _lib
    L = _lib.load()
    blocks, iters = 3, 5
    out = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(L.tkmi_calib_mfma_f64(blocks, iters, out.data_ptr(), s), "calib")
    got = out.cpu().numpy().reshape(blocks, 4, 64).sum(axis=2)
    for w in range(4):
        t = 64 * w + np.arange(64)
        a, b = 1.0 + t, 1.0 - t
        want = 16 * iters * sum(a[16 * k:16 * k + 16].sum() * b[16 * k:16 * k + 16].sum() for k in range(4))
        np.testing.assert_array_equal(got[:, w], np.full(blocks, want))
    assert L.tkmi_calib_mfma_f64(0, iters, out.data_ptr(), s) != 0
    assert L.tkmi_calib_mfma_f64(blocks, 0, out.data_ptr(), s) != 0
    assert L.tkmi_calib_mfma_f64(blocks, iters, None, s) != 0
