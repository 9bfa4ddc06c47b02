"""GPU: partial trace (trace_permute! / TO.tensortrace! / tr) through the device
transform, against the reference's own trace tests (test/tensors/contractions.jl:17-69):
  * "Trace: test via conversion" (bosonic): the traced tensor equals the same index sum
    over the dense embedding (the oracle's U1 placement and SU(2) Clebsch-Gordan);
  * "Partial trace: self-consistency": t[c d b; c d a] == (t[e d c; e b a])[a c; b c],
    here with fermionic (fZ2⊠U1) sectors, where twists and signs enter;
  * "Full trace: self-consistency": tr after the twists == t[a b b a] == (t[a c c b])[a a]."""
import math

import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space, SpaceMismatch, IndexError_

from helpers import oracle_legs, kind_of
from oracle import abelian as O
from oracle import su2 as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def fspace(pairs):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})


def _su2_legs(W):
    return ([({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.codomain],
            [({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.domain])


def test_u1_trace_via_conversion(T):
    rng = np.random.default_rng(0)
    V1 = U1Space((0, 3), (1, 2), (-1, 2))
    V2 = U1Space((0, 2), (1, 3), (2, 1))
    V3 = U1Space((0, 2), (-1, 1))
    W = HomSpace(ProductSpace(V1, V2.dual(), V3, V2, V1.dual(), V3.dual()), ProductSpace())
    A = T.TensorMap.randn(W, rng)
    p, q = ((5, 2), ()), ((0, 1), (4, 3))
    C = T.TensorMap.randn(W.select(*p), rng)
    c0 = C.numpy()
    T.trace_permute_(C, A, p, q, 0.5, -1.5)
    kind = kind_of(W)
    ref = -1.5 * O.to_dense(c0, *oracle_legs(C.space), kind) + \
        0.5 * np.einsum("cdbdca->ab", O.to_dense(A.numpy(), *oracle_legs(W), kind))
    np.testing.assert_allclose(O.to_dense(C.numpy(), *oracle_legs(C.space), kind), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("case", [0, 1])
def test_su2_trace_via_conversion(T, case):
    rng = np.random.default_rng(10 + case)
    V1 = SU2Space((0, 2), (0.5, 1))
    V2 = SU2Space((0.5, 1), (1, 2))
    V3 = SU2Space((0, 1), (1, 1))
    if case == 0:
        W = HomSpace(ProductSpace(V1, V2.dual(), V3, V2, V1.dual(), V3.dual()), ProductSpace())
        p, q, spec = ((5, 2), ()), ((0, 1), (4, 3)), "cdbdca->ab"
    else:
        W = HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V1, V2, V3))
        p, q, spec = ((5,), (2,)), ((0, 1), (3, 4)), "cdbcda->ab"
    A = T.TensorMap.randn(W, rng)
    B = T.tensortrace(A, p, q)
    ref = np.einsum(spec, S.to_dense(A.numpy(), *_su2_legs(W)))
    got = S.to_dense(B.numpy(), *_su2_legs(B.space))
    assert np.linalg.norm(got - ref) <= 1e-13 * max(1.0, np.linalg.norm(ref))


def test_fermionic_partial_trace_self_consistency(T):
    rng = np.random.default_rng(3)
    V1 = fspace([((0, 0), 2), ((1, 1), 1), ((1, -1), 2)])
    V2 = fspace([((0, 0), 1), ((1, 1), 2), ((0, 2), 1)])
    V3 = fspace([((0, 0), 2), ((1, -1), 1)])
    t = T.TensorMap.randn(HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V1, V2, V3)), rng)
    t2 = T.tensortrace(t, ((5,), (2,)), ((0, 1), (3, 4)))          # t2[a; b] := t[c d b; c d a]
    t4 = T.tensortrace(t, ((5, 4), (2, 1)), ((0,), (3,)))          # t4[a b; c d] := t[e d c; e b a]
    t5 = T.tensortrace(t4, ((0,), (2,)), ((1,), (3,)))             # t5[a; b] := t4[a c; b c]
    assert t2.space == t5.space
    x2, x5 = t2.numpy(), t5.numpy()
    assert np.linalg.norm(x2 - x5) <= 1e-13 * np.linalg.norm(x2)
    assert np.linalg.norm(x2) > 0


def test_fermionic_full_trace_self_consistency(T):
    rng = np.random.default_rng(4)
    V1 = fspace([((0, 0), 2), ((1, 1), 1), ((1, -1), 1)])
    V2 = fspace([((0, 0), 1), ((1, 1), 2), ((0, 2), 1)])
    t = T.TensorMap.randn(HomSpace(ProductSpace(V1, V2.dual(), V2, V1.dual()), ProductSpace()), rng)
    t2 = T.permute(t, (0, 1), (3, 2), copy=True)
    if not V1.isdual:
        T.twist_(t2, (0,))
    if V2.isdual:
        T.twist_(t2, (1,))
    ss = T.tr(t2)
    s2 = T.tensortrace(t, ((), ()), ((0, 1), (3, 2))).numpy()[0]          # t[a, b, b, a]
    t3 = T.tensortrace(t, ((0, 3), ()), ((1,), (2,)))                      # t3[a, b] := t[a, c, c, b]
    s3 = T.tensortrace(t3, ((), ()), ((0,), (1,))).numpy()[0]              # t3[a, a]
    assert math.isclose(ss, s2, rel_tol=1e-12, abs_tol=1e-12)
    assert math.isclose(ss, s3, rel_tol=1e-12, abs_tol=1e-12)


def test_tr_is_weighted_block_trace(T):
    rng = np.random.default_rng(5)
    for V in (U1Space((0, 3), (1, 2), (-1, 4)), SU2Space((0, 2), (0.5, 3), (1, 1))):
        t = T.TensorMap.randn(HomSpace(ProductSpace(V, V), ProductSpace(V, V)), rng)
        ref = sum(c.dim() * np.trace(b) for c, b in t.blocks_host().items())
        assert math.isclose(T.tr(t), ref, rel_tol=1e-13, abs_tol=1e-13)


def test_trace_errors(T):
    rng = np.random.default_rng(6)
    V = U1Space((0, 2), (1, 1))
    Vo = U1Space((0, 1), (1, 1))
    t = T.TensorMap.randn(HomSpace(ProductSpace(V, Vo), ProductSpace(V, Vo)), rng)
    C = T.TensorMap.zeros(HomSpace(ProductSpace(Vo), ProductSpace(Vo)))
    with pytest.raises(IndexError_):
        T.trace_permute_(C, t, ((1,), (3,)), ((0,), ()))
    with pytest.raises(IndexError_):
        T.trace_permute_(C, t, ((1,), (3,)), ((0,), (3,)))          # index 3 used twice
    with pytest.raises(SpaceMismatch):
        T.trace_permute_(C, t, ((0,), (2,)), ((1,), (3,)))          # tdst space is Vo←Vo, select gives V←V
    with pytest.raises(SpaceMismatch):
        T.tr(T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(Vo)), rng))
