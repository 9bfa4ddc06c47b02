"""GPU: flip and twist on the device, tested as the reference does
(test/tensors/indexmanipulations.jl:100-131):
- flipping inverse: t == flip(flip(t, i), i; inv) == flip(flip(t, i; inv), i);
- flip via contraction: flipping both legs of a contracted pair leaves a contraction
  unchanged, and flipping open legs equals flipping the result — which fixes the
  flip coefficients (a wrong Frobenius–Schur phase or twist changes the result for the
  fermionic and SU(2) cases)."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _fermion(pairs):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})


def _five(kind):
    if kind == "u1":
        return [U1Space((0, 2), (1, 1), (-1, 1)), U1Space((0, 1), (1, 2)), U1Space((1, 1), (-1, 2)),
                U1Space((0, 2), (1, 1)), U1Space((0, 1), (-1, 1), (2, 1))]
    if kind == "fermion":
        return [_fermion([((0, 0), 2), ((1, 1), 1)]), _fermion([((1, -1), 2), ((0, 0), 1)]),
                _fermion([((0, 1), 1), ((1, 0), 2)]), _fermion([((0, 0), 1), ((1, 1), 1), ((0, 1), 1), ((1, 0), 1)]),
                _fermion([((1, 0), 1), ((0, -1), 2), ((0, 0), 1), ((1, 1), 1)])]
    return [SU2Space((0, 1), (0.5, 1)), SU2Space((0.5, 1), (1, 1)), SU2Space((0, 1), (1, 1)),
            SU2Space((0.5, 2)), SU2Space((0, 1), (0.5, 1), (1, 1))]


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_flip_inverse(T, kind):
    V1, V2, V3, V4, V5 = _five(kind)
    rng = np.random.default_rng(41)
    t = T.TensorMap.randn(HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V4, V5).dual()), rng)
    for i in range(5):
        a = T.flip(T.flip(t, i), i, inv=True)
        b = T.flip(T.flip(t, i, inv=True), i)
        assert a.space == t.space and b.space == t.space
        assert np.array_equal(a.numpy(), t.numpy()) and np.array_equal(b.numpy(), t.numpy())


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_flip_via_contraction(T, kind):
    V1, V2, V3, V4, V5 = _five(kind)
    rng = np.random.default_rng(42)
    t1 = T.TensorMap.randn(HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V4)), rng)
    t2 = T.TensorMap.randn(HomSpace(ProductSpace(V2.dual(), V5), ProductSpace(V4.dual(), V1)), rng)

    def con(a, b):  # ta[a, b] := t1[x, y, a, z] * t2[y, b, z, x]
        return T.contract_labels(a, "xyaz", b, "ybzx", "ab", 1)

    assert t1.dim() > 0 and t2.dim() > 0
    ta = con(t1, t2)
    ref = ta.numpy()
    scale = np.linalg.norm(ref)
    for i1, i2 in (((0,), (3,)), ((1, 3), (0, 2))):
        tb = con(T.flip(t1, i1), T.flip(t2, i2))
        assert tb.space == ta.space
        assert np.linalg.norm(tb.numpy() - ref) <= 1e-12 * scale, (i1, i2)
    tb = con(T.flip(t1, (0, 2)), T.flip(t2, (1, 3)))
    fa = T.flip(ta, (0, 1))
    assert tb.space == fa.space
    assert np.linalg.norm(tb.numpy() - fa.numpy()) <= 1e-12 * scale


def test_twist_checks_and_shortcuts(T):
    rng = np.random.default_rng(43)
    F = _fermion([((0, 0), 2), ((1, 1), 1), ((1, -1), 1)])
    t = T.TensorMap.randn(HomSpace(ProductSpace(F, F), ProductSpace(F)), rng)
    with pytest.raises(ValueError):
        T.twist_(t, (3,))
    assert T.twist(t, ()) is t                              # nothing to twist
    tw = T.twist(t, (0, 2))
    assert tw is not t
    back = T.twist(tw, (0, 2), inv=True)
    assert np.array_equal(back.numpy(), t.numpy())          # θ = ±1: its own inverse
    V = U1Space((0, 2), (1, 1))
    u = T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(V)), rng)
    assert T.has_shared_twist(u, (0, 1)) and T.twist(u, (0, 1)) is u   # bosonic: identity
    assert T.twist(u, (0,), copy=True) is not u
    assert not T.has_shared_twist(t, (0,))


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_flip_is_unitary(T, kind):
    """⟨flip(t, I), s⟩ = ⟨t, flip(s, I; inv = !inv)⟩ for the inner product Σ_c dim(c)·⟨b₁, b₂⟩:
    the adjoint of flip is the opposite flip, which is what its pullback applies."""
    from tkmi import chainrules as CR
    V1, V2, V3, V4, V5 = _five(kind)
    rng = np.random.default_rng(44)
    W = HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V4, V5).dual())
    t = T.TensorMap.randn(W, rng)
    for I, inv in (((1,), False), ((0, 3), True), ((2, 4), False)):
        s = T.TensorMap.randn(W.flip(I), rng)
        lhs = CR.inner(T.flip(t, I, inv), s)
        rhs = CR.inner(t, T.flip(s, I, not inv))
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0), (I, inv)


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_flip_rules(T, kind):
    """ChainRules rrule, Mooncake rrule!!/frule!! and Enzyme augmented_primal/reverse/forward
    of flip (ext/TensorKitChainRulesCoreExt/linalg.jl:99-103,
    ext/TensorKitEnzymeExt/indexmanipulations.jl:119-170): the pullback accumulates the
    opposite flip of the cotangent, the forward tangent is the flip of the tangent."""
    from tkmi import chainrules as CR, inplace_rules as R
    V1, V2, V3, V4, V5 = _five(kind)
    rng = np.random.default_rng(45)
    W = HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V4, V5).dual())
    t = T.TensorMap.randn(W, rng)
    I = (0, 3)
    dy = T.TensorMap.randn(W.flip(I), rng)
    want = T.flip(dy, I, True).numpy()
    y, pb = CR.rrule_flip(t, I)
    assert np.array_equal(y.numpy(), T.flip(t, I).numpy())
    assert np.array_equal(pb(dy).numpy(), want)
    # Mooncake: accumulates into the existing tangent
    dt0 = T.TensorMap.randn(W, rng)
    dt = dt0.copy()
    y2, pb2 = R.rrule(T.flip, R.Duplicated(t, dt), I)
    assert pb2(dy) == (None, None)
    assert np.allclose(dt.numpy(), dt0.numpy() + want, rtol=1e-14, atol=1e-14)
    # Enzyme: the caller accumulates into the shadow the augmented primal hands back
    dt = dt0.copy()
    y3, shadow = R.augmented_primal(T.flip, R.Duplicated(t, dt), I, False)
    assert not shadow.numpy().any() and shadow.space == y3.space
    T.add_(shadow, dy, 1.0, 1.0)
    R.reverse(T.flip, shadow, R.Duplicated(t, dt), I, False)
    assert np.allclose(dt.numpy(), dt0.numpy() + want, rtol=1e-14, atol=1e-14)
    # forward mode
    y4, dy4 = R.forward(T.flip, R.Duplicated(t, dt0), I)
    assert np.array_equal(dy4.numpy(), T.flip(dt0, I).numpy())
    y5, dy5 = R.frule(T.flip, R.Const(t), I, True)
    assert np.array_equal(y5.numpy(), T.flip(t, I, True).numpy()) and not dy5.numpy().any()
