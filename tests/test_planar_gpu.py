"""GPU: the planar operations of sectors without a bosonic braiding
(src/planar/planaroperations.jl:29-190), through the device transform and grouped GEMM.

1. ℙ spaces (the braiding-free trivial sector): the reference's own planaradd!,
   planartrace! and planarcontract! cases (test/tensors/planar.jl:59-99, with their
   spaces and index tuples, 0-based here), against numpy on the dense arrays — the
   reference checks them against tensoradd!/tensortrace!/tensorcontract! on ℂ spaces.
2. FermionParity ⊠ U1: every cyclic transpose of a rank-4 tensor against the oracle's
   planar transpose (coefficient 1 for every bend and fold of these sectors: the dense
   permutation of axes), bit for bit, also as transpose! with non-dyadic α, β; it differs
   from the braided permute by the Koszul signs.
3. FermionParity ⊠ U1 planartrace! and planarcontract! (reorder_indices, the planar
   transposes, one mul!) against the dense index sums: without crossings no braiding
   sign enters, so the planar contraction is the plain contraction of the dense arrays,
   unlike contract!, which braids.
"""
import numpy as np
import pytest
import torch

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, PlanarSpace

from helpers import oracle_legs, kind_of
from oracle import abelian as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def P(n, dual=False):
    return PlanarSpace(n, dual)


def dense_P(t):
    """A ℙ tensor's data is its one block: the dense array in column-major order."""
    W = t.space
    dims = [V.dim() for V in W.codomain] + [V.dim() for V in W.domain]
    return t.numpy().reshape(dims, order="F")


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_planar_trivial_planaradd(T):
    A = T.TensorMap.randn(HomSpace(ProductSpace(P(2), P(3)), ProductSpace(P(6), P(5), P(4))),
                          np.random.default_rng(1))
    WC = HomSpace(ProductSpace(P(5, True), P(6, True)), ProductSpace(P(4), P(3, True), P(2, True)))
    C = T.TensorMap.randn(WC, np.random.default_rng(2))
    c0 = dense_P(C)
    p = ((3, 2), (4, 1, 0))
    alpha, beta = 0.3, 0.7
    T.planaradd_(C, A, p, alpha, beta)
    ref = beta * c0 + alpha * dense_P(A).transpose(3, 2, 4, 1, 0)
    np.testing.assert_array_equal(dense_P(C), ref)   # one coefficient-1 move: bit for bit
    # braiding is refused for ℙ
    with pytest.raises(ValueError):
        T.permute(A, (1, 0), (2, 3, 4), copy=True)


def test_planar_trivial_planartrace(T):
    rng = np.random.default_rng(3)
    A = T.TensorMap.randn(HomSpace(ProductSpace(P(2), P(3)), ProductSpace(P(2), P(5), P(4))), rng)
    C = T.TensorMap.randn(HomSpace(ProductSpace(P(5, True), P(3)), ProductSpace(P(4))), rng)
    c0 = dense_P(C)
    T.planartrace_(C, A, ((3, 1), (4,)), ((0,), (2,)), 1.0, 1.0)
    ref = c0 + np.einsum("ijikl->kjl", dense_P(A))
    assert relerr(dense_P(C), ref) <= 1e-14


def test_planar_trivial_planarcontract(T):
    rng = np.random.default_rng(4)
    A = T.TensorMap.randn(HomSpace(ProductSpace(P(2), P(3)), ProductSpace(P(2), P(5), P(4))), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(P(2), P(4)), ProductSpace(P(4), P(3))), rng)
    C = T.TensorMap.randn(HomSpace(ProductSpace(P(5, True), P(2, True), P(2)), ProductSpace(P(2, True), P(4))),
                          rng)
    c0 = dense_P(C)
    T.planarcontract_(C, A, ((0, 2, 3), (4, 1)), B, ((1, 3), (0, 2)), ((2, 1, 0), (3, 4)), 1.0, 1.0)
    ab = np.tensordot(dense_P(A), dense_P(B), axes=([4, 1], [1, 3])).transpose(2, 1, 0, 3, 4)
    assert relerr(dense_P(C), c0 + ab) <= 1e-14


# ---------------------------------------------------------------------------
# FermionParity ⊠ U1
# ---------------------------------------------------------------------------
def F(pairs, dual=False):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs.items()}, dual)


X = [
    {(0, 0): 3, (1, 0): 2, (0, 1): 2, (1, 1): 3, (0, -1): 1, (1, -1): 2},
    {(0, 0): 1, (1, 1): 2, (1, -1): 1},
    {(0, 0): 2, (1, 0): 1, (0, 1): 1, (1, -1): 2},
    {(1, 0): 2, (0, 1): 3, (1, 1): 1, (0, -1): 2},
    {(0, 0): 2, (1, 1): 1, (0, 1): 1},
]


def dense_F(t):
    W = t.space
    return O.to_dense(t.numpy(), *oracle_legs(W), kind_of(W))


def _cyclic_perms(n1, n2):
    n = n1 + n2
    lin = list(range(n1)) + list(range(n1 + n2 - 1, n1 - 1, -1))
    out = []
    for r in range(n):
        rot = lin[r:] + lin[:r]
        for k in range(n + 1):
            out.append((tuple(rot[:k]), tuple(reversed(rot[k:]))))
    return out


def test_fermionic_transposes_bit_exact(T):
    WA = HomSpace(ProductSpace(F(X[0]), F(X[1])), ProductSpace(F(X[2], True), F(X[3])))
    rng = np.random.default_rng(5)
    A = T.TensorMap.randn(WA, rng)
    a = A.numpy()
    legs = oracle_legs(WA)
    kind = kind_of(WA)
    differs = 0
    for p1, p2 in _cyclic_perms(2, 2):
        got = T.transpose(A, p1, p2)
        ref, _, _ = O.transpose(a, *legs, p1, p2, kind)
        np.testing.assert_array_equal(got.numpy(), ref, err_msg=str((p1, p2)))
        # transpose! with non-dyadic scalars: round(β·C) + round(α·A′), bit for bit
        C = T.TensorMap.randn(got.space, rng)
        c0 = C.numpy()
        T.transpose_(C, A, p1, p2, 0.3, 0.7)
        np.testing.assert_array_equal(C.numpy(), 0.7 * c0 + 0.3 * ref)
        # the braided permute of the same legs carries Koszul signs
        if not np.array_equal(T.permute(A, p1, p2, copy=True).numpy(), ref):
            differs += 1
    assert differs > 0
    # repartition is planar too: a transpose with the cyclic order unchanged
    R = T.repartition(A, 3)
    ref, _, _ = O.transpose(a, *legs, (0, 1, 3), (2,), kind)
    np.testing.assert_array_equal(R.numpy(), ref)


def test_fermionic_planartrace(T):
    V, Wl = F(X[0]), F(X[4])
    rng = np.random.default_rng(6)
    # neighbouring legs in the codomain: Σ_w A[a w w b]
    WA = HomSpace(ProductSpace(V, Wl, Wl.dual()), ProductSpace(V))
    A = T.TensorMap.randn(WA, rng)
    C = T.TensorMap.randn(WA.select((0,), (3,)), rng)
    c0 = dense_F(C)
    T.planartrace_(C, A, ((0,), (3,)), ((1,), (2,)), 1.5, -0.5)
    ref = 1.5 * np.einsum("awwb->ab", dense_F(A)) - 0.5 * c0
    assert relerr(dense_F(C), ref) <= 1e-14
    # the first codomain leg with the domain leg (neighbours through the cyclic order):
    # Σ_a A2[a x; a]
    WA2 = HomSpace(ProductSpace(V, Wl), ProductSpace(V))
    A2 = T.TensorMap.randn(WA2, rng)
    C2 = T.TensorMap.zeros(WA2.select((1,), ()))
    T.planartrace_(C2, A2, ((1,), ()), ((0,), (2,)), 1.0, 0.0)
    ref2 = np.einsum("axa->x", dense_F(A2))
    assert relerr(dense_F(C2), ref2) <= 1e-14


def test_fermionic_planarcontract(T):
    """The reference example's index tuples on fermionic spaces: A: X0⊗X1 ← X2⊗X3⊗X4,
    B: Y0⊗X4 ← Y2⊗X1 (the contracted pairs A4–B1, A1–B3), C = the open legs in pAB's
    order.  Also a plain mul! layout (no transposes) and a contraction of every leg of A."""
    rng = np.random.default_rng(7)
    X0, X1, X2, X3, X4 = (F(x) for x in X)
    Y0, Y2 = F(X[3], True), F(X[0])
    WA = HomSpace(ProductSpace(X0, X1), ProductSpace(X2, X3, X4))
    WB = HomSpace(ProductSpace(Y0, X4), ProductSpace(Y2, X1))
    A = T.TensorMap.randn(WA, rng)
    B = T.TensorMap.randn(WB, rng)
    WC = HomSpace(ProductSpace(WA[3], WA[2], WA[0]), ProductSpace(WB[0].dual(), WB[2].dual()))
    C = T.TensorMap.randn(WC, rng)
    c0 = dense_F(C)
    pA, pB, pAB = ((0, 2, 3), (4, 1)), ((1, 3), (0, 2)), ((2, 1, 0), (3, 4))
    T.planarcontract_(C, A, pA, B, pB, pAB, 0.7, 0.3)
    ab = np.tensordot(dense_F(A), dense_F(B), axes=([4, 1], [1, 3])).transpose(2, 1, 0, 3, 4)
    assert relerr(dense_F(C), 0.3 * c0 + 0.7 * ab) <= 1e-13
    # contract! braids the same legs: Koszul signs and twists make it a different map
    Cb = T.TensorMap.zeros(WC)
    T.contract_(Cb, A, pA, B, pB, pAB, 1.0, 0.0)
    assert relerr(dense_F(Cb), ab) > 1e-3

    # mul! layout: C[a b; e] = A[a b; c d]·B[c d; e]
    WA3 = HomSpace(ProductSpace(X0, X1), ProductSpace(X2, X3))
    WB3 = HomSpace(ProductSpace(X2, X3), ProductSpace(X4))
    A3, B3 = T.TensorMap.randn(WA3, rng), T.TensorMap.randn(WB3, rng)
    C3 = T.TensorMap.zeros(HomSpace(ProductSpace(X0, X1), ProductSpace(X4)))
    T.planarcontract_(C3, A3, ((0, 1), (2, 3)), B3, ((0, 1), (2,)), ((0, 1), (2,)))
    ref3 = np.tensordot(dense_F(A3), dense_F(B3), axes=([2, 3], [0, 1]))
    assert relerr(dense_F(C3), ref3) <= 1e-13

    # every leg of A contracted: A4: X0 ⊗ X1 ← X2, B4: X0′ ⊗ X2 ⊗ X1′ ← X4 (planar pairing
    # A0–B0, A2–B1, A1–B2), C4 = B4's open leg (its unit-sector part)
    WA4 = HomSpace(ProductSpace(X0, X1), ProductSpace(X2))
    WB4 = HomSpace(ProductSpace(X0.dual(), X2, X1.dual()), ProductSpace(X4))
    A4, B4 = T.TensorMap.randn(WA4, rng), T.TensorMap.randn(WB4, rng)
    C4 = T.TensorMap.zeros(HomSpace(ProductSpace(), ProductSpace(X4)))
    T.planarcontract_(C4, A4, ((), (0, 2, 1)), B4, ((0, 1, 2), (3,)), ((), (0,)))
    ref4 = np.einsum("abc,acbd->d", dense_F(A4), dense_F(B4))
    assert relerr(dense_F(C4).reshape(-1), ref4) <= 1e-13
