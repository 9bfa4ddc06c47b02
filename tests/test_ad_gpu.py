"""GPU: reverse-mode rules (adjoint permute and contraction kernels) checked
against finite differences along random directions, with
fermionic (fZ2⊠U1) braiding signs and twists in play (the differentiation
workload: V graded by (p, q), p ∈ {0,1}, q ∈ -4:4, Gaussian degeneracies)."""
import math

import numpy as np
import pytest
import torch

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


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
    """Tensor with identity blocks (for tr(R) = inner(R, Id))."""
    x = np.zeros(W.dim())
    for c, (d1, d2), off in W.degeneracystructure()[1]:
        blk = np.zeros((d1, d2))
        np.fill_diagonal(blk, 1.0)
        x[off:off + d1 * d2] = blk.reshape(-1, order="F")
    return T.TensorMap.from_numpy(x, W)


def cost(T, A, M, Id, with_grad):
    from tkmi import chainrules as CR
    Tt, pbT = CR.rrule_contract_labels(A, "xsz", A, "zty", "xsty", 3)
    MT, pbM = CR.rrule_contract_labels(M, "STst", Tt, "xsty", "xSTy", 3)
    R, pbR = CR.rrule_contract_labels(Tt, "xSTY", MT, "xSTy", "Yy", 1, conjA=True)
    f = CR.inner(R, Id)
    if not with_grad:
        return f, None
    dT1, dMT = pbR(Id)
    _, dT2 = pbM(dMT)
    dT = CR.add(dT1, dT2)
    dA1, dA2 = pbT(dT)
    return f, CR.add(dA1, dA2)


def test_fermionic_cost_gradient_fd(T):
    from tkmi import chainrules as CR
    V = fspace(1024)  # D ≈ 1024 over (p, q), p ∈ {0,1}, q ∈ -4:4
    P = physical()
    rng = np.random.default_rng(4)
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    WM = HomSpace(ProductSpace(P, P), ProductSpace(P, P))
    M0 = T.TensorMap.randn(WM, rng)
    # Hermitian part: M = (M0 + M0†)/2
    M = T.TensorMap.empty(WM)
    T.tensoradd_(M, M0, ((2, 3), (0, 1)), True, 0.5, 0.0)
    T.add_(M, M0, 0.5, 1.0)
    Id = identity_like(T, HomSpace(ProductSpace(V), ProductSpace(V)))
    f, g = cost(T, A, M, Id, True)
    # The cost is a quartic polynomial in A (R = T†·M·T with T = A·A), so along
    # any direction the 5-point stencil [f(-2h) - 8f(-h) + 8f(h) - f(2h)] / 12h
    # is exact: no truncation error, and h = 1e-2 keeps the cancellation
    # roundoff (≈ ε·|f|/h) far below the tolerance.  A wrong sign, twist or
    # transposed operand in a pullback shows as an O(1) relative error of ⟨g, D⟩.
    h = 1e-2
    for k in range(32):
        D = T.TensorMap.randn(A.space, rng)
        fs = [cost(T, CR.add(A, D, s * h, 1.0), M, Id, False)[0] for s in (-2, -1, 1, 2)]
        fd = (fs[0] - 8 * fs[1] + 8 * fs[2] - fs[3]) / (12 * h)
        an = CR.inner(g, D)
        # scale: the Cauchy-Schwarz bound ‖g‖·‖D‖ of |⟨g, D⟩|
        scale = math.sqrt(CR.inner(g, g) * CR.inner(D, D))
        assert abs(fd - an) <= 1e-10 * scale, (k, fd, an, f, scale)


def test_permute_pullback_is_adjoint(T):
    """⟨Δ, permute(v)⟩ = ⟨pullback(Δ), v⟩ for fermionic and SU(2) tensors."""
    from tkmi import chainrules as CR
    from tkmi.spaces import SU2Space
    rng = np.random.default_rng(9)
    V = fspace(24)
    P = physical()
    for W, (p1, p2) in [(HomSpace(ProductSpace(V, P), ProductSpace(V, P)), ((3, 0), (2, 1))),
                        (HomSpace(ProductSpace(SU2Space((0, 2), (0.5, 2), (1, 1)), SU2Space((0.5, 1), (1, 2))),
                                  ProductSpace(SU2Space((0, 1), (0.5, 2)))), ((1,), (2, 0)))]:
        v = T.TensorMap.randn(W, rng)
        y, pb = CR.rrule_permute(v, p1, p2)
        d = T.TensorMap.randn(y.space, rng)
        lhs = CR.inner(d, y)
        rhs = CR.inner(pb(d), v)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_contract_pullback_fd_u1(T):
    """Pullback of one contraction with conj operands vs finite differences."""
    from tkmi import chainrules as CR
    from helpers import cfg1_spaces
    V, P = cfg1_spaces()
    rng = np.random.default_rng(5)
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    C, pb = CR.rrule_contract_labels(A, "xpa", B, "xpb", "ab", 1, conjA=True)
    dC = T.TensorMap.randn(C.space, rng)
    dA, dB = pb(dC)
    # directional derivative along (DA, DB)
    DA = T.TensorMap.randn(A.space, rng)
    DB = T.TensorMap.randn(B.space, rng)
    h = 1e-6
    Cp = T.contract_labels(CR.add(A, DA, h), "xpa", CR.add(B, DB, h), "xpb", "ab", 1, conjA=True)
    Cm = T.contract_labels(CR.add(A, DA, -h), "xpa", CR.add(B, DB, -h), "xpb", "ab", 1, conjA=True)
    fd = (CR.inner(dC, Cp) - CR.inner(dC, Cm)) / (2 * h)
    an = CR.inner(dA, DA) + CR.inner(dB, DB)
    assert abs(fd - an) <= 1e-7 * abs(an)
