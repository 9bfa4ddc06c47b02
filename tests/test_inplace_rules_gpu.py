"""GPU: the mutating-AD rules (Mooncake `rrule!!`/`frule!!`, Enzyme
`augmented_primal`/`reverse`/`forward`; ext/TensorKitMooncakeExt/*.jl,
ext/TensorKitEnzymeExt/*.jl) over the device kernels.

Every primal here is a polynomial of degree ≤ 3 in (C, A, B, α, β) (C′ = β·C + α·f(A, B)),
so the 5-point stencil [F(-2h) - 8F(-h) + 8F(h) - F(2h)]/12h is its exact directional
derivative up to roundoff.  That pins the forward rules; the reverse rules are then pinned
by duality with TensorKit's inner product Σ_c dim(c)·⟨x_c, y_c⟩ (the one Mooncake's
tangent `_dot` uses, ext/TensorKitMooncakeExt/tangent.jl:152):
    ⟨w, F′(x)·v⟩ = ⟨β·w, v_C⟩ + Σ_X ⟨X̄, v_X⟩ + ᾱ·v_α + β̄·v_β,
with the tangents X̄ accumulated onto random prefills (the += convention) and the
overwritten output restored by the Mooncake pullback.  A wrong sign, twist, weight or
transposed operand in a pullback shows as an O(1) error."""
import math

import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, SU2Space, U1Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _fermion(pairs):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})


def _spaces(kind):
    if kind == "u1":
        return U1Space((0, 5), (1, 4), (-1, 3), (2, 2)), U1Space((0.5, 1), (-0.5, 1))
    if kind == "fermion":
        return (_fermion([((0, 0), 4), ((1, 1), 3), ((1, -1), 3), ((0, 2), 2)]),
                _fermion([((0, 0), 1), ((1, 1), 1), ((1, -1), 1)]))
    return SU2Space((0, 3), (0.5, 3), (1, 2)), SU2Space((0.5, 1), (1, 1))


def _rand(T, W, rng):
    return T.TensorMap.randn(W, rng)


def _check(T, f, args, tpos, spos, rng, tol=1e-10, enzyme=False):
    """Forward rule vs the 5-point stencil, reverse rule vs the forward rule (duality)."""
    from tkmi import chainrules as CR
    from tkmi import inplace_rules as IR
    vt = {i: _rand(T, args[i].space, rng) for i in tpos}
    vs = {i: float(rng.standard_normal()) for i in spos}

    def F(h):
        a = list(args)
        a[0] = args[0].copy()
        for i in tpos:
            a[i] = CR.add(args[i], vt[i], h, 1.0)
        for i in spos:
            a[i] = args[i] + h * vs[i]
        f(*a)
        return a[0]

    h = 1e-2
    fs = [F(s * h) for s in (-2, -1, 1, 2)]
    fd = CR.add(CR.add(fs[0], fs[1], -8.0, 1.0), CR.add(fs[3], fs[2], 8.0, -1.0), 1.0, 1.0)
    T.scale_(fd, 1.0 / (12 * h))
    F0 = F(0.0)
    nfd = math.sqrt(CR.inner(fd, fd))
    assert nfd > 1e-3

    # forward rule
    fargs = list(args)
    fargs[0] = IR.Duplicated(args[0].copy(), vt[0].copy() if 0 in vt else None)
    for i in tpos:
        if i != 0:
            fargs[i] = IR.Duplicated(args[i], vt[i])
    for i in spos:
        fargs[i] = IR.Duplicated(args[i], vs[i])
    fwd = IR.forward if enzyme else IR.frule
    out, dout = fwd(f, *fargs)
    err = CR.add(dout, fd, -1.0, 1.0)
    assert math.sqrt(CR.inner(err, err)) <= tol * nfd
    e0 = CR.add(out, F0, -1.0, 1.0)
    assert math.sqrt(CR.inner(e0, e0)) <= 1e-13 * math.sqrt(CR.inner(F0, F0))

    # reverse rule
    w = _rand(T, args[0].space, rng)
    pre = {i: _rand(T, args[i].space, rng) for i in tpos if i != 0}
    rargs = list(args)
    C_before = args[0].copy()
    rargs[0] = IR.Duplicated(args[0].copy(), w.copy())
    for i in pre:
        rargs[i] = IR.Duplicated(args[i], pre[i].copy())
    for i in spos:
        rargs[i] = IR.Active(args[i])
    if enzyme:
        out, tape = IR.augmented_primal(f, *rargs)
        grads = IR.reverse(f, tape, *rargs)
    else:
        out, pb = IR.rrule(f, *rargs)
        e1 = CR.add(out, F0, -1.0, 1.0)
        assert math.sqrt(CR.inner(e1, e1)) <= 1e-13 * math.sqrt(CR.inner(F0, F0))
        grads = pb()
        # the pullback restored the overwritten output
        e2 = CR.add(rargs[0].val, C_before, -1.0, 1.0)
        assert math.sqrt(CR.inner(e2, e2)) == 0.0
    lhs = CR.inner(w, fd)
    rhs = CR.inner(rargs[0].dval, vt[0])
    for i in pre:
        d = CR.add(rargs[i].dval, pre[i], -1.0, 1.0)
        rhs += CR.inner(d, vt[i])
    for i in spos:
        rhs += grads[i] * vs[i]
    scale = math.sqrt(CR.inner(w, w)) * nfd  # Cauchy-Schwarz bound of |⟨w, F′v⟩|
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs, scale)
    return grads


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_permute_rule(T, kind):
    rng = np.random.default_rng(1)
    V, P = _spaces(kind)
    A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(V, P)), rng)
    p1, p2 = (3, 0), (2, 1)
    C = _rand(T, A.space.permute(p1, p2), rng)
    _check(T, T.permute_, [C, A, p1, p2, 0.7, -0.4], (0, 1), (4, 5), rng)


def test_braid_rule_fermion(T):
    rng = np.random.default_rng(2)
    V, P = _spaces("fermion")
    A = _rand(T, HomSpace(ProductSpace(V, P, V), ProductSpace(P)), rng)
    p1, p2, levels = (2, 1), (3, 0), (3, 1, 0, 2)
    C = _rand(T, A.space.permute(p1, p2), rng)
    _check(T, T.braid_, [C, A, p1, p2, levels, 1.3, 0.5], (0, 1), (5, 6), rng)


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_transpose_rule(T, kind):
    rng = np.random.default_rng(3)
    V, P = _spaces(kind)
    A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(P, V)), rng)
    p1, p2 = (1, 3), (0, 2)  # a cyclic shift of the planar order (0, 1, 3, 2)
    C = _rand(T, A.space.permute(p1, p2), rng)
    _check(T, T.transpose_, [C, A, p1, p2, -0.9, 0.25], (0, 1), (4, 5), rng)


def test_twist_rule_fermion(T):
    from tkmi import chainrules as CR
    from tkmi import inplace_rules as IR
    rng = np.random.default_rng(4)
    V, P = _spaces("fermion")
    t0 = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    dt0 = _rand(T, t0.space, rng)
    # forward: twist is linear, so the tangent is the twisted tangent
    t, dt = t0.copy(), dt0.copy()
    out, dout = IR.frule(T.twist_, IR.Duplicated(t, dt), (1, 2))
    ref = dt0.copy()
    T.twist_(ref, (1, 2))
    e = CR.add(dout, ref, -1.0, 1.0)
    assert CR.inner(e, e) == 0.0
    # reverse: ⟨w, twist(v)⟩ = ⟨twist(w), v⟩ and the primal is restored
    t = t0.copy()
    w = _rand(T, t0.space, rng)
    arg = IR.Duplicated(t, w.copy())
    out, pb = IR.rrule(T.twist_, arg, (1, 2))
    pb()
    e = CR.add(arg.val, t0, -1.0, 1.0)
    assert CR.inner(e, e) == 0.0
    lhs = CR.inner(w, ref)
    assert abs(lhs - CR.inner(arg.dval, dt0)) <= 1e-12 * abs(lhs)


def _contract_case(T, kind, rng):
    V, P = _spaces(kind)
    W = _spaces("u1")[1] if kind == "u1" else P
    if kind == "u1":
        # an H_eff MPO step: small B, so blas_contract! takes the fused small-operand path
        Wm = U1Space((0, 2), (1, 1), (-1, 1))
        A = _rand(T, HomSpace(ProductSpace(V, Wm.dual()), ProductSpace(P, P, V)), rng)
        B = _rand(T, HomSpace(ProductSpace(Wm, P), ProductSpace(P, Wm)), rng)
        pA, pB, pAB = ((0, 3, 4), (1, 2)), ((0, 1), (2, 3)), ((0, 3), (1, 2, 4))
    else:
        # A's domain leg (V) contracts with B's first codomain leg (V'), which is dual: with
        # fermions the contraction twists an operand
        A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(V.dual())), rng)
        B = _rand(T, HomSpace(ProductSpace(V.dual(), W), ProductSpace(V)), rng)
        pA, pB, pAB = ((0, 1), (2,)), ((0,), (1, 2)), ((0, 2), (1, 3))
    from tkmi.tensors import tensorcontract_space
    VC = tensorcontract_space(A, pA, False, B, pB, False, pAB)
    C = _rand(T, VC, rng)
    return C, A, pA, B, pB, pAB


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
@pytest.mark.parametrize("front", ["mooncake", "enzyme"])
def test_blas_contract_rule(T, kind, front):
    from tkmi import fusedcontract as FC
    rng = np.random.default_rng(5)
    C, A, pA, B, pB, pAB = _contract_case(T, kind, rng)
    if kind == "u1":
        assert FC.eligible(C, A, pA, B, pB)
    _check(T, T.blas_contract_, [C, A, pA, B, pB, pAB, 0.8, -0.6], (0, 1, 3), (6, 7), rng,
           enzyme=(front == "enzyme"))


@pytest.mark.parametrize("kind", ["u1", "fermion", "su2"])
def test_trace_permute_rule(T, kind):
    rng = np.random.default_rng(6)
    V, P = _spaces(kind)
    A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(P, V)), rng)
    p, q = ((0,), (3,)), ((1,), (2,))  # trace the P legs, keep V ← V
    C = _rand(T, A.space.select(*p), rng)
    _check(T, T.trace_permute_, [C, A, p, q, 1.1, 0.3], (0, 1), (4, 5), rng)


@pytest.mark.parametrize("kind", ["u1", "su2"])
@pytest.mark.parametrize("front", ["mooncake", "enzyme"])
def test_mul_rule(T, kind, front):
    rng = np.random.default_rng(7)
    V, P = _spaces(kind)
    A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(V)), rng)
    B = _rand(T, HomSpace(ProductSpace(V), ProductSpace(V, P)), rng)
    C = _rand(T, A.space.compose(B.space), rng)
    _check(T, T.mul_, [C, A, B, -0.5, 0.9], (0, 1, 2), (3, 4), rng, enzyme=(front == "enzyme"))


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_tr_rule(T, kind):
    from tkmi import chainrules as CR
    from tkmi import inplace_rules as IR
    rng = np.random.default_rng(8)
    V, P = _spaces(kind)
    A = _rand(T, HomSpace(ProductSpace(V, P), ProductSpace(V, P)), rng)
    dA = _rand(T, A.space, rng)
    val, dval = IR.frule(T.tr, IR.Duplicated(A, dA))
    assert abs(val - T.tr(A)) <= 1e-12 * abs(val)
    assert abs(dval - T.tr(dA)) <= 1e-12 * abs(dval)  # tr is linear
    pre = _rand(T, A.space, rng)
    arg = IR.Duplicated(A, pre.copy())
    _, pb = IR.rrule(T.tr, arg)
    pb(2.5)
    d = CR.add(arg.dval, pre, -1.0, 1.0)
    # ⟨Δ, v⟩ = 2.5·tr(v) for every v
    v = _rand(T, A.space, rng)
    assert abs(CR.inner(d, v) - 2.5 * T.tr(v)) <= 1e-12 * max(1.0, abs(2.5 * T.tr(v)))
