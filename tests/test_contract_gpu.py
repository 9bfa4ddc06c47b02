"""GPU parity: contractions (tensorcontract! → permutes + grouped GEMM) vs the
oracle's dense embedding (U1: contraction of the dense arrays is exact in
exact arithmetic; FP64 tolerance 1e-13 relative Frobenius)."""
import numpy as np
import pytest
import torch

from tkmi.spaces import HomSpace, ProductSpace, U1Space

from oracle import abelian as O
from helpers import cfg1_spaces, cfg2_spaces, oracle_legs, random_u1_space, dense_heff

pytestmark = pytest.mark.gpu

TOL = 1e-13


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def dense(t):
    cod, dom = oracle_legs(t.space)
    return O.to_dense(t.numpy(), cod, dom, O.U1Kind)


def undense(d, W):
    cod, dom = oracle_legs(W)
    return O.from_dense(d, cod, dom, O.U1Kind)


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_cfg1_contraction(T):
    """L'[a;b] := L[x;y] * A[y p;b] * conj(A[x p;a])."""
    V, P = cfg1_spaces()
    rng = np.random.default_rng(0)
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    L = T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(V)), rng)
    T1 = T.contract_labels(L, "xy", A, "ypb", "xpb", 2)
    Lp = T.contract_labels(T1, "xpb", A, "xpa", "ab", 1, conjB=True)
    assert Lp.space == HomSpace(ProductSpace(V), ProductSpace(V))
    ref = np.einsum("xy,ypb,xpa->ab", dense(L), dense(A), dense(A))
    assert relerr(Lp.numpy(), undense(ref, Lp.space)) < TOL


@pytest.mark.parametrize("seed", range(8))
def test_random_pairwise(T, seed):
    rng = np.random.default_rng(500 + seed)
    Vs = [random_u1_space(rng, maxdeg=6) for _ in range(4)]
    a, b, c, d = Vs
    # A[i j; k] with j,k contracted against B[k' j'; l]
    A = T.TensorMap.randn(HomSpace(ProductSpace(a, b), ProductSpace(c)), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(c, b.dual()), ProductSpace(d)), rng)
    C = T.contract_labels(A, "ijk", B, "kjl", "il", 1)
    ref = np.einsum("ijk,kjl->il", dense(A), dense(B))
    if np.linalg.norm(ref) == 0:
        assert np.linalg.norm(C.numpy()) == 0
        return
    assert relerr(C.numpy(), undense(ref, C.space)) < TOL
    # alpha / beta accumulation into an existing C
    C2 = T.TensorMap.randn(C.space, rng)
    c0 = C2.numpy()
    T.contract_labels(A, "ijk", B, "kjl", "il", 1, out=C2, alpha=0.5, beta=2.0)
    assert relerr(C2.numpy(), 2.0 * c0 + 0.5 * C.numpy()) < TOL


def test_heff_chain_vs_dense(T):
    """Two-site H_eff chain at reduced D against a dense einsum."""
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=96, qmax=4)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(1))
    y = H.apply()
    ref = dense_heff(dense(H.L), dense(H.x), dense(H.M1), dense(H.M2), dense(H.R))
    assert relerr(y.numpy(), undense(ref, y.space)) < TOL


def test_heff_graph_replays(T):
    """The captured chain replays to the eager result, repeatedly, with eager
    work in between."""
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=256, qmax=6)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(2))
    eager = H.apply().numpy().copy()
    H.capture()
    for i in range(3):
        y = H.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.numpy(), eager)
        H.apply()  # eager work between replays
        torch.cuda.synchronize()


def test_graph_gathers_read_the_mpo_at_every_replay(T):
    """The captured chain gathers the MPO coefficients of its two small-operand steps at
    every replay: changing M1 and M2 in place after the capture changes the replayed result
    exactly as it changes the eager one (nothing of M is frozen into the graph)."""
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=256, qmax=6)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(3))
    H.capture()
    for scale in (2.0, -0.5):
        H.M1.data.mul_(scale)
        H.M2.data.add_(1.0)
        y = H.replay()
        torch.cuda.synchronize()
        got = y.numpy().copy()
        eager = H.apply().numpy()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got, eager)


def test_graph_survives_cache_eviction(T):
    """A captured graph pins the plans it launches: emptying every plan cache (and
    collecting the evicted plans) leaves the replay intact."""
    import gc
    from tkmi import treetransformers as TT
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=256, qmax=6)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(6))
    eager = H.apply().numpy().copy()
    H.capture()
    assert H.graph_pins  # the GEMM, transform and fused plans of the chain
    TT.empty_globalcaches()
    H.outs = [None] * len(H.outs)  # drop the eager intermediates as well
    gc.collect()
    torch.cuda.synchronize()
    for _ in range(2):
        y = H.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.numpy(), eager)


def test_full_contraction_to_a_scalar(T):
    """@tensor C = A[i j k] * B[i j k]-style full contraction (the reference's MPO benchmark
    ends in one, benchmark/comparisons/tensorkit_timers.jl:6-17): the result lives in the
    scalar space (one 1×1 block of the unit sector) and equals the dense contraction; its
    GEMM is a single 1×1 group with a long K, computed as a dot product."""
    rng = np.random.default_rng(7)
    a = U1Space((0, 40), (1, 30), (-1, 30), (2, 11))
    b = U1Space((0, 3), (1, 2), (-1, 2))
    A = T.TensorMap.randn(HomSpace(ProductSpace(a, b, a.dual()), ProductSpace()), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(a.dual(), b.dual(), a), ProductSpace()), rng)
    C = T.contract_labels(A, "ijk", B, "ijk", "", 0)
    assert C.space.dim() == 1 and C.numout == 0 and C.numin == 0
    ref = float(np.einsum("ijk,ijk->", dense(A), dense(B)))
    assert abs(float(C.numpy()[0]) - ref) <= 1e-12 * np.linalg.norm(A.numpy()) * np.linalg.norm(B.numpy())
    # α/β into an existing scalar, and conj of the second operand (real data: the same)
    C2 = T.TensorMap.randn(C.space, rng)
    c0 = float(C2.numpy()[0])
    T.contract_labels(A, "ijk", B, "ijk", "", 0, conjB=False, out=C2, alpha=2.0, beta=-1.0)
    assert abs(float(C2.numpy()[0]) - (2.0 * ref - c0)) <= 1e-12 * (abs(ref) + abs(c0) + 1)
