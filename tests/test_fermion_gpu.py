"""GPU: fermionic (fZ2⊠U1) path, two properties.

1. The device path does not depend on the order of product sectors.  The order of
   `FermionParity ⊠ U1Irrep` sectors comes from TensorKitSectors, which is not part of the
   reference's sources, so the host mirror assumes it (compared from the last factor).
   Here the same tensors are built under that order and under the opposite one (compared
   from the first factor): every sub-block then sits at another offset of the flat vector,
   yet the permutes give the same sub-blocks bit for bit, keyed by their sector labels,
   and the contractions the same sub-blocks to rounding.  So the kernels take whatever
   layout the host's order fixes; the order itself stays the one assumption.

2. The differentiation workload's cost f(A) = real(tr(permute(A)·contract(A, conj(A), M))),
   with its gradient from the hand-written adjoint kernels (the permute's braid pullback
   included) against an exact directional derivative.
"""
import math

import numpy as np
import pytest
import torch

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace

from helpers import oracle_legs, kind_of
from oracle import abelian as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _as_strided(x, sz, st, off):
    return np.lib.stride_tricks.as_strided(x[off:], shape=tuple(sz), strides=tuple(s * x.itemsize for s in st))


def labelled(t):
    """Sub-blocks of a tensor keyed by the uncoupled sectors of their tree pair."""
    x = t.numpy()
    out = {}
    for (f1, f2), (sz, st, off) in t.space.degeneracystructure()[2].items():
        out[(f1.uncoupled, f2.uncoupled)] = _as_strided(x, sz, st, off).copy()
    return out


def from_labelled(T, W, blocks):
    total, _, subs = W.degeneracystructure()
    x = np.zeros(total)
    for (f1, f2), (sz, st, off) in subs.items():
        _as_strided(x, sz, st, off)[...] = blocks[(f1.uncoupled, f2.uncoupled)]
    return T.TensorMap.from_numpy(x, W)


_LEGS = [
    {(0, 0): 3, (1, 0): 2, (0, 1): 2, (1, 1): 3, (0, -1): 1, (1, -1): 2, (0, 2): 2},
    {(0, 0): 1, (1, 1): 1, (1, -1): 1},
    {(0, 0): 2, (1, 0): 1, (0, 1): 1, (1, -1): 2, (1, 2): 1},
    {(1, 0): 2, (0, 1): 3, (1, 1): 1, (0, -1): 2},
]


def _space(i, dual=False):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in _LEGS[i].items()},
                       dual=dual)


def _build(T):
    """A: V0 ⊗ V1 ← V2′ ⊗ V3 and B: V2′ ⊗ V3 ← V0, under the current sector order."""
    WA = HomSpace(ProductSpace(_space(0), _space(1)), ProductSpace(_space(2, True), _space(3)))
    WB = HomSpace(ProductSpace(_space(2, True), _space(3)), ProductSpace(_space(0)))
    return WA, WB


def _run(T, A, B):
    out = {}
    for p1, p2 in [((1, 0), (3, 2)), ((2,), (0, 3, 1)), ((3, 1, 0), (2,)), ((0, 2), (1, 3))]:
        out[("permute", p1, p2)] = labelled(T.permute(A, p1, p2, copy=True))
    # C[a b; e] = A[a b; c d]·B[c d; e] (mul! over coupled sectors) and
    # E[b d; x y] = A[a b; c d]·conj(A[a x; c y]), whose contracted dual leg brings a twist
    out["mul"] = labelled(T.contract_labels(A, "abcd", B, "cde", "abe", 2))
    out["contract"] = labelled(T.contract_labels(A, "abcd", A, "axcy", "bdxy", 2, conjB=True))
    return out


def test_product_sector_order_does_not_reach_the_kernels(T):
    import sys
    from tkmi import treetransformers as TT
    FT = sys.modules["tkmi.fusiontrees"]   # the package re-exports a function of that name
    WA, WB = _build(T)
    order_last = [V.sectors() for V in list(WA.codomain) + list(WA.domain)]
    rng = np.random.default_rng(41)
    A = T.TensorMap.randn(WA, rng)
    B = T.TensorMap.randn(WB, rng)
    LA, LB = labelled(A), labelled(B)
    ref = _run(T, A, B)

    def reset():
        TT.empty_globalcaches()
        FT._fusiontrees_cached.cache_clear()

    orig = ProductSector._order
    try:
        ProductSector._order = property(lambda self: tuple(s._order for s in self.sectors))
        reset()
        WA2, WB2 = _build(T)
        order_first = [V.sectors() for V in list(WA2.codomain) + list(WA2.domain)]
        assert order_first != order_last          # the layouts really differ
        A2 = from_labelled(T, WA2, LA)
        B2 = from_labelled(T, WB2, LB)
        assert not np.array_equal(A2.numpy(), A.numpy())
        got = _run(T, A2, B2)
    finally:
        ProductSector._order = orig
        reset()
    for key, blocks in ref.items():
        assert set(got[key]) == set(blocks), key
        for lab, blk in blocks.items():
            if key[0] == "permute":
                np.testing.assert_array_equal(got[key][lab], blk, err_msg=str((key, lab)))
            else:
                # a different order of the contracted index changes the summation order
                assert np.linalg.norm(got[key][lab] - blk) <= 1e-13 * max(1.0, np.linalg.norm(blk)), (key, lab)


# ---------------------------------------------------------------------------
# the differentiation workload's cost
# ---------------------------------------------------------------------------
def fspace(total, qmax=4, width2=16.0):
    secs = [(p, q) for q in range(-qmax, qmax + 1) for p in (0, 1)]
    w = np.array([math.exp(-q * q / width2) for p, q in secs])
    d = np.maximum(1, np.round(total * w / w.sum())).astype(int)
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): int(n) for (p, q), n in zip(secs, d)})


def physical():
    return GradedSpace({ProductSector(FermionParity(0), U1Irrep(0)): 1,
                        ProductSector(FermionParity(1), U1Irrep(1)): 1,
                        ProductSector(FermionParity(1), U1Irrep(-1)): 1})


def identity_like(T, W):
    x = np.zeros(W.dim())
    for c, (d1, d2), off in W.degeneracystructure()[1]:
        blk = np.zeros((d1, d2))
        np.fill_diagonal(blk, 1.0)
        x[off:off + d1 * d2] = blk.reshape(-1, order="F")
    return T.TensorMap.from_numpy(x, W)


PERM = ((1, 0), (3, 2))   # permute(A): swaps the virtual and first physical leg, braids the domain


def cost(T, A, M, Id, with_grad):
    """f(A) = tr(Ap·X), Ap = permute(A, PERM), X = contract(A, conj(A), M):
         AM[x; y]       = A[x s t; y]·M[; s t]
         X[Y T; S y]    = AM[x; y]·conj(A[x S T; Y])
         R[s x; S y]    = Ap[s x; Y T]·X[Y T; S y]      (Ap's domain meets X's codomain)
       f = tr(R) = inner(R, Id).  Every leg pairs with its dual only for the two-site A
       (V ⊗ P ⊗ P ← V) and M a form on P ⊗ P (∅ ← P ⊗ P): the cost is a cubic polynomial
       in A (real data: real(f) = f)."""
    from tkmi import chainrules as CR
    Ap, pbAp = CR.rrule_permute(A, *PERM)
    AM, pbAM = CR.rrule_contract_labels(A, "xsty", M, "st", "xy", 1)
    X, pbX = CR.rrule_contract_labels(AM, "xy", A, "xSTY", "YTSy", 2, conjB=True)
    R, pbR = CR.rrule_contract_labels(Ap, "sxYT", X, "YTSy", "sxSy", 2)
    f = CR.inner(R, Id)
    if not with_grad:
        return f, None
    dAp, dX = pbR(Id)
    dAM, dA2 = pbX(dX)
    dA3, _ = pbAM(dAM)
    dA1 = pbAp(dAp)
    return f, CR.add(CR.add(dA1, dA2), dA3)


def test_cfg5_cost_gradient(T):
    from tkmi import chainrules as CR
    V = fspace(1024)   # D ≈ 1024 over (p, q), p ∈ {0,1}, q ∈ -4:4
    P = physical()
    rng = np.random.default_rng(4)
    WA = HomSpace(ProductSpace(V, P, P), ProductSpace(V))
    A = T.TensorMap.randn(WA, rng)
    WM = HomSpace(ProductSpace(), ProductSpace(P, P))
    # M Hermitian (real symmetric) as a form on P ⊗ P: M[s t] = M[t s] entry by entry
    m0 = T.TensorMap.randn(WM, rng).numpy()
    Md = O.to_dense(m0, *oracle_legs(WM), kind_of(WM))
    Md = 0.5 * (Md + Md.transpose(1, 0))
    M = T.TensorMap.from_numpy(O.from_dense(Md, *oracle_legs(WM), kind_of(WM)), WM)
    # the braid of permute(A) is bit-exact against the oracle's Koszul signs
    Ap = T.permute(A, *PERM, copy=True)
    ref, _, _ = O.permute(A.numpy(), *oracle_legs(WA), *PERM, kind_of(WA))
    np.testing.assert_array_equal(Ap.numpy(), ref)
    # R's space is P ⊗ V ← P ⊗ V (square blocks): tr(R) = inner(R, Id)
    Id = identity_like(T, HomSpace(ProductSpace(P, V), ProductSpace(P, V)))
    f, g = cost(T, A, M, Id, True)
    assert abs(f) > 0
    # f is cubic in A: the 5-point stencil [f(-2h) - 8f(-h) + 8f(h) - f(2h)] / 12h is its
    # exact directional derivative; a wrong braid sign, twist or transposed operand in any
    # pullback shows as an O(1) relative error of ⟨g, D⟩
    h = 1e-2
    for k in range(32):
        D = T.TensorMap.randn(A.space, rng)
        fs = [cost(T, CR.add(A, D, s * h, 1.0), M, Id, False)[0] for s in (-2, -1, 1, 2)]
        fd = (fs[0] - 8 * fs[1] + 8 * fs[2] - fs[3]) / (12 * h)
        an = CR.inner(g, D)
        scale = math.sqrt(CR.inner(g, g) * CR.inner(D, D))
        assert abs(fd - an) <= 1e-10 * scale, (k, fd, an, f, scale)
