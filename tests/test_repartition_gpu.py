"""GPU: repartition / repartition! and the deprecated add_permute! / add_braid! /
add_transpose! wrappers (src/tensors/indexmanipulations.jl:433-484, 542-553), tested as
the reference does (test/tensors/indexmanipulations.jl:55-98):
- conversion: convert(Array, repartition(t, k)) ≈ permutedims(convert(Array, t),
  (1…k, reverse(k+1…5))) — here with the oracle's independent dense embeddings (U(1)
  placement, SU(2) Clebsch–Gordan);
- norm and inner-product invariance, for fermionic sectors as well."""
import warnings

import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space

from oracle import abelian as O
from oracle import su2 as S
from helpers import oracle_legs, kind_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _su2_legs(W):
    return ([({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.codomain],
            [({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.domain])


def _dense(W, x):
    if W.unit.fusion_style == "unique":
        return O.to_dense(x, *oracle_legs(W), kind_of(W))
    return S.to_dense(x, *_su2_legs(W))


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_repartition_via_conversion(T, kind):
    rng = np.random.default_rng(31)
    if kind == "u1":
        Vs = [U1Space((0, 2), (1, 1)), U1Space((0, 1), (-1, 2)), U1Space((1, 1), (-1, 1)),
              U1Space((0, 1), (2, 1)), U1Space((0, 2), (-1, 1))]
    else:
        Vs = [SU2Space((0, 1), (0.5, 1)), SU2Space((0.5, 1)), SU2Space((0, 1), (1, 1)), SU2Space((0.5, 2)),
              SU2Space((0, 1), (0.5, 1))]
    W = HomSpace(ProductSpace(*Vs), ProductSpace())
    t = T.TensorMap.randn(W, rng)
    a = _dense(W, t.numpy())
    for k in range(6):
        t3 = T.repartition(t, k)
        assert t3.numout == k and t3.numin == 5 - k
        a3 = _dense(t3.space, t3.numpy())
        ref = np.transpose(a, tuple(range(k)) + tuple(reversed(range(k, 5))))
        assert np.allclose(a3, ref, rtol=1e-13, atol=1e-13), k
        # repartition! into an existing tensor, and back to all-codomain: the identity
        t4 = T.TensorMap.empty(t3.space)
        T.repartition_(t4, t, 1.0, 0.0)
        assert np.array_equal(t4.numpy(), t3.numpy())
        back = T.TensorMap.empty(W)
        T.repartition_(back, t3, 1.0, 0.0)
        assert np.allclose(back.numpy(), t.numpy(), rtol=1e-14, atol=1e-14)


def test_repartition_fermionic_inner_product_invariance(T):
    from tkmi import chainrules as CR
    f = lambda pairs: GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})
    Vs = [f([((0, 0), 2), ((1, 1), 1)]), f([((1, -1), 2), ((0, 0), 1)]), f([((0, 1), 1), ((1, 0), 2)]),
          f([((0, 0), 1), ((1, 1), 1)]), f([((1, 0), 1), ((0, -1), 2)])]
    W = HomSpace(ProductSpace(*Vs), ProductSpace())
    rng = np.random.default_rng(32)
    t, t2 = T.TensorMap.randn(W, rng), T.TensorMap.randn(W, rng)
    for k in range(6):
        r, r2 = T.repartition(t, k), T.repartition(t2, k)
        assert abs(CR.inner(r, r) - CR.inner(t, t)) <= 1e-12 * CR.inner(t, t)
        assert abs(CR.inner(r2, r) - CR.inner(t2, t)) <= 1e-12 * abs(CR.inner(t2, t)) + 1e-12
        back = T.repartition(r, 5)
        assert np.allclose(back.numpy(), t.numpy(), rtol=1e-14, atol=1e-14)


def test_repartition_shares_data_without_copy(T):
    rng = np.random.default_rng(33)
    V = U1Space((0, 2), (1, 1))
    t = T.TensorMap.randn(HomSpace(ProductSpace(V, V), ProductSpace(V)), rng)
    assert T.repartition(t, 2) is t
    assert T.repartition(t, 2, copy=True) is not t
    with pytest.raises(ValueError):
        T.repartition(t, 2, 2)


def test_deprecated_add_wrappers(T):
    rng = np.random.default_rng(34)
    V = U1Space((0, 2), (1, 1), (-1, 1))
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, V), ProductSpace(V)), rng)
    p = ((2, 0), (1,))
    C0 = T.TensorMap.randn(A.space.permute(*p), rng)
    ref = C0.copy()
    T.permute_(ref, A, p[0], p[1], 0.5, -1.0)
    for fn, args in ((T.add_permute_, (p, 0.5, -1.0)), (T.add_braid_, (p, (0, 1, 2), 0.5, -1.0))):
        C = C0.copy()
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            fn(C, A, *args)
        assert any(issubclass(x.category, DeprecationWarning) for x in w)
        assert np.array_equal(C.numpy(), ref.numpy())
    pt = ((1,), (0, 2))  # a cyclic transpose of (0, 1 | 2)
    Ct0 = T.TensorMap.randn(A.space.permute(*pt), rng)
    Ct, reft = Ct0.copy(), Ct0.copy()
    T.transpose_(reft, A, pt[0], pt[1], 1.0, 0.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        T.add_transpose_(Ct, A, pt, 1.0, 0.0)
    assert any(issubclass(x.category, DeprecationWarning) for x in w)
    assert np.array_equal(Ct.numpy(), reft.numpy())


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_planaradd_is_transpose(T, kind):
    """planaradd! (src/planar/planaroperations.jl:29-36) is transpose!: same result for a
    cyclic transpose, α and β included; a non-cyclic index order is rejected."""
    rng = np.random.default_rng(35)
    if kind == "u1":
        V = U1Space((0, 2), (1, 1), (-1, 1))
    else:
        V = SU2Space((0, 1), (0.5, 1), (1, 1))
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, V), ProductSpace(V, V)), rng)
    p = ((1, 3), (0, 2))  # the cyclic order (0 1 | 3 2) rotated by one
    C0 = T.TensorMap.randn(A.space.permute(*p), rng)
    ref, C = C0.copy(), C0.copy()
    T.transpose_(ref, A, p[0], p[1], 0.5, -1.0)
    T.planaradd_(C, A, p, 0.5, -1.0)
    assert np.array_equal(C.numpy(), ref.numpy())
    with pytest.raises(Exception):
        T.planaradd_(T.TensorMap.empty(A.space.permute((0, 2), (1, 3))), A, ((0, 2), (1, 3)))
