"""GPU parity: grouped abelian permute (K1) through the C ABI vs the oracle.
U1 and fermionic permutes are bit-exact (pure data movement / ±1 scaling)."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, ProductSpace, HomSpace, U1Space

from oracle import abelian as O
from helpers import cfg1_spaces, cfg2_spaces, oracle_legs, kind_of, random_u1_space, random_perm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _ref_permute(W, x, p1, p2, alpha=1.0):
    cod, dom = oracle_legs(W)
    out, _, _ = O.permute(x, cod, dom, p1, p2, kind_of(W), alpha)
    return out


def test_cfg1_permute(T):
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    rng = np.random.default_rng(0)
    A = T.TensorMap.randn(W, rng)
    x = A.numpy()
    B = T.permute(A, (2, 0), (1,))
    np.testing.assert_array_equal(B.numpy(), _ref_permute(W, x, (2, 0), (1,)))


# α, β include values that are not powers of two: the reference's tensoradd! rounds
# β·C and (α·coeff)·A separately (Julia does not fuse a*b + c), which a fused
# multiply-add in the kernel would not reproduce (the CPU table tests show these draws
# tell the two apart)
_ALPHAS = [1.0, -2.0, 0.5, 0.3, -1.7]
_BETAS = [0.0, 1.0, -0.25, 0.7, 1.0 / 3.0]


def _expected(ref64, c0, beta, dtype):
    """β·C + ref in FP64 with separately rounded terms; FP32 data is promoted by the
    Float64 scalars and rounded once when stored."""
    c = c0.astype(np.float64)
    out = beta * c + ref64 if beta != 0.0 else ref64
    return out.astype(dtype)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("seed", range(20))
def test_random_u1_permutes(T, seed, dtype):
    import torch
    rng = np.random.default_rng(1000 + seed)
    n1, n2 = int(rng.integers(1, 4)), int(rng.integers(0, 4))
    W = HomSpace(ProductSpace(*[random_u1_space(rng, maxdeg=7) for _ in range(n1)]),
                 ProductSpace(*[random_u1_space(rng, maxdeg=7) for _ in range(n2)]))
    if W.dim() == 0:
        pytest.skip("empty")
    p1, p2 = random_perm(rng, n1, n2)
    tdt = getattr(torch, dtype)
    A = T.TensorMap.randn(W, rng, dtype=tdt)
    x = A.numpy().astype(np.float64)
    alpha = _ALPHAS[seed % 5]
    beta = _BETAS[(seed // 5 + seed) % 5]
    C = T.TensorMap.randn(W.permute(p1, p2), rng, dtype=tdt)
    c0 = C.numpy()
    T.permute_(C, A, p1, p2, alpha, beta)
    ref = _expected(_ref_permute(W, x, p1, p2, alpha), c0, beta, np.dtype(dtype))
    np.testing.assert_array_equal(C.numpy(), ref)


def test_cfg2_permutes(T):
    """All repartitions of the two-site tensor V⊗P⊗P←V used by the H_eff step."""
    V, P, _ = cfg2_spaces()
    W = HomSpace(ProductSpace(V, P, P), ProductSpace(V))
    rng = np.random.default_rng(1)
    A = T.TensorMap.randn(W, rng)
    x = A.numpy()
    for p1, p2 in [((0,), (1, 2, 3)), ((0, 1), (2, 3)), ((3, 2, 1, 0), ()), ((1, 2), (0, 3)), ((3,), (0, 1, 2))]:
        B = T.permute(A, p1, p2)
        np.testing.assert_array_equal(B.numpy(), _ref_permute(W, x, p1, p2))


def test_permute_roundtrip_and_f32(T):
    V, P, _ = cfg2_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V, P))
    rng = np.random.default_rng(2)
    import torch
    A = T.TensorMap.randn(W, rng, dtype=torch.float32)
    B = T.permute(A, (3, 1), (0, 2))
    Cb = T.permute(B, (2, 1), (3, 0))
    np.testing.assert_array_equal(Cb.numpy(), A.numpy())


def test_adjoint_source(T):
    """tensoradd!(C, A, p, conjA=true): permute of the lazy adjoint."""
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    rng = np.random.default_rng(4)
    A = T.TensorMap.randn(W, rng)
    # C = permute(conj(A), pA): A′ = A' with pA mapped by adjointtensorindices
    pA = ((2, 0), (1,))
    q1, q2 = T.adjointtensorindices(A, pA[0]), T.adjointtensorindices(A, pA[1])
    Wa = W.adjoint()
    C = T.TensorMap.empty(Wa.permute(q1, q2))
    T.tensoradd_(C, A, pA, True)
    # oracle: the adjoint of real data is permutedims with the domain legs first
    cod, dom = oracle_legs(W)
    dense = O.to_dense(A.numpy(), cod, dom, O.U1Kind).transpose(2, 0, 1)
    acod, adom = oracle_legs(Wa)
    xa = O.from_dense(np.ascontiguousarray(dense), acod, adom, O.U1Kind)
    ref, _, _ = O.permute(xa, acod, adom, q1, q2, O.U1Kind)
    np.testing.assert_array_equal(C.numpy(), ref)


def _fz2u1_space(rng, dual):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): int(rng.integers(1, 5))
                        for p in (0, 1) for q in (-1, 0, 1) if rng.random() < 0.8}, dual=dual)


@pytest.mark.parametrize("seed", range(6))
def test_fermionic_permute(T, seed):
    rng = np.random.default_rng(300 + seed)
    n1, n2 = int(rng.integers(1, 3)), int(rng.integers(1, 3))
    W = HomSpace(ProductSpace(*[_fz2u1_space(rng, bool(rng.integers(0, 2))) for _ in range(n1)]),
                 ProductSpace(*[_fz2u1_space(rng, bool(rng.integers(0, 2))) for _ in range(n2)]))
    if W.dim() == 0:
        pytest.skip("empty")
    p1, p2 = random_perm(rng, n1, n2)
    A = T.TensorMap.randn(W, rng)
    B = T.permute(A, p1, p2, copy=True)
    np.testing.assert_array_equal(B.numpy(), _ref_permute(W, A.numpy(), p1, p2))
    # with scalars: C = β·C + α·permute(A), signs ±1 folded into α·coeff
    alpha, beta = ((0.3, 0.7), (-1.7, 1.0 / 3.0), (2.0, 1.0))[seed % 3]
    C = T.TensorMap.randn(B.space, rng)
    c0 = C.numpy()
    T.permute_(C, A, p1, p2, alpha, beta)
    ref = _expected(_ref_permute(W, A.numpy(), p1, p2, alpha), c0, beta, np.float64)
    np.testing.assert_array_equal(C.numpy(), ref)


def test_scale_and_add(T):
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    rng = np.random.default_rng(5)
    A = T.TensorMap.randn(W, rng)
    B = T.TensorMap.randn(W, rng)
    a, b = A.numpy(), B.numpy()
    T.add_(B, A, 2.0, -1.0)
    np.testing.assert_array_equal(B.numpy(), -1.0 * b + 2.0 * a)
    Z = T.TensorMap.zeros(W)
    assert not Z.numpy().any()


def test_space_mismatch_raises(T):
    from tkmi.spaces import SpaceMismatch
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    A = T.TensorMap.randn(W, np.random.default_rng(0))
    C = T.TensorMap.empty(W)
    with pytest.raises(SpaceMismatch):
        T.permute_(C, A, (2, 0), (1,))
    with pytest.raises(IndexError):
        T.permute_(C, A, (0, 0), (1,))


# ---------------------------------------------------------------------------
# generic (SU(2)) transform: K2 vs the oracle's C loop over the same tables and
# vs the dense CG embedding
# ---------------------------------------------------------------------------
def _su2_legs(W):
    return ([({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.codomain],
            [({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.domain])


@pytest.mark.parametrize("k", range(5))
def test_su2_permute_dense(T, k):
    import itertools
    from tkmi.spaces import SU2Space
    from oracle import su2 as S
    rng = np.random.default_rng(40 + k)
    V = SU2Space((0, 2), (0.5, 1), (1, 2))
    U = SU2Space((0.5, 2), (1.5, 1))
    W = HomSpace(ProductSpace(V, U.dual()), ProductSpace(U, V))
    A = T.TensorMap.randn(W, rng)
    dA = S.to_dense(A.numpy(), *_su2_legs(W))
    for p in list(itertools.permutations(range(4)))[::3]:
        p1, p2 = p[:k], p[k:]
        B = T.permute(A, p1, p2, copy=True)
        dB = S.to_dense(B.numpy(), *_su2_legs(B.space))
        ref = dA.transpose(p)
        assert np.linalg.norm(dB - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("div", [16, 4])
def test_su2_permute_vs_c_loop_cfg3_scaled(T, div):
    """cfg-3 leg space with degeneracies ÷16 (dims 2-6) and ÷4 (dims 8-24, 40.5 M
    elements): K2 against the oracle's table loop.  The last two permutations have
    plans whose transpose tiles take slices along a third dim (÷16: 1190 tiles against
    5268 without slices; ÷4: 20 320 against 84 288), with partial slices at the ends of
    the 6- and 24-wide dims; the first permutation has none."""
    import ctypes
    import torch
    from tkmi.spaces import SU2Space
    from tkmi import _lib, chainrules as CR
    from oracle import kernels as OK
    V = SU2Space(*((j, n // div) for j, n in ((0, 64), (0.5, 96), (1, 96), (1.5, 64), (2, 32))))
    W = HomSpace(ProductSpace(V, V), ProductSpace(V, V))
    rng = np.random.default_rng(2)
    A = T.TensorMap.randn(W, rng)
    x = A.numpy()
    A32 = T.TensorMap.from_numpy(x, W, torch.float32)
    x32 = A32.numpy().astype(np.float64)
    sliced_bound = {16: 1190, 4: 20320}[div]
    for n, (p1, p2) in enumerate([((0, 2), (1, 3)), ((3, 1), (2, 0)), ((1, 0, 3), (2,))]):
        B = T.permute(A, p1, p2)
        plan = T.treebraider(B.space, W, p1, p2)
        y = np.zeros(B.space.dim())
        OK.generic_permute(plan.table, p1 + p2, x, y)
        assert np.linalg.norm(B.numpy() - y) <= 1e-14 * np.linalg.norm(y)
        if n > 0:
            # these plans take slices along a third dim (4.4x fewer tiles than without)
            nt = ctypes.c_int64()
            _lib.check(_lib.load().tkmi_transform_plan_info(plan.handle, ctypes.byref(nt), None, None), "info")
            assert nt.value <= sliced_bound, (div, p1, p2, nt.value)
        if True:
            # (the first permutation keeps its short innermost leg innermost: vector-transpose
            # tiles, forward and adjoint)
            # the adjoint exec of the same plan: ⟨d, permute(A)⟩ = ⟨pullback(d), A⟩
            yb, pb = CR.rrule_permute(A, p1, p2)
            d = T.TensorMap.randn(yb.space, rng)
            lhs, rhs = CR.inner(d, yb), CR.inner(pb(d), A)
            assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
            # FP32 through the same plan (sums in FP64, one rounding at the store)
            B32 = T.permute(A32, p1, p2)
            y32 = np.zeros(B.space.dim())
            OK.generic_permute(plan.table, p1 + p2, x32, y32)
            assert np.linalg.norm(B32.numpy() - y32) <= 1e-7 * np.linalg.norm(y32)
        # round trip
        inv = [list(p1 + p2).index(i) for i in range(4)]
        C = T.permute(B, tuple(inv[:2]), tuple(inv[2:]))
        assert np.linalg.norm(C.numpy() - x) <= 1e-13 * np.linalg.norm(x)


def test_su2_many_trees_dense(T):
    """Fusion blocks with up to 15 trees (rank-6 tensor of spin-1 legs): the
    kernel deals destination trees out in chunks of 8 and consumes source trees
    in chunks, summing in registers.  Checked against the dense CG embedding,
    and the pullback against ⟨Δ, permute(v)⟩ = ⟨pullback(Δ), v⟩."""
    from tkmi.spaces import SU2Space
    from tkmi import chainrules as CR
    from oracle import su2 as S
    rng = np.random.default_rng(77)
    V = SU2Space((1, 2))
    U = SU2Space((0.5, 1), (1, 1))
    W = HomSpace(ProductSpace(V, U, V), ProductSpace(V, V, U.dual()))
    A = T.TensorMap.randn(W, rng)
    dA = S.to_dense(A.numpy(), *_su2_legs(W))
    for p1, p2 in [((5, 1, 3), (0, 4, 2)), ((2, 0), (5, 1, 3, 4)), ((0, 1, 2, 3), (4, 5))]:
        plan = T.treebraider(W.permute(p1, p2), W, p1, p2)
        B = T.permute(A, p1, p2, copy=True)
        dB = S.to_dense(B.numpy(), *_su2_legs(B.space))
        ref = dA.transpose(p1 + p2)
        assert np.linalg.norm(dB - ref) <= 1e-13 * np.linalg.norm(ref)
        y, pb = CR.rrule_permute(A, p1, p2)
        d = T.TensorMap.randn(y.space, rng)
        lhs = CR.inner(d, y)
        rhs = CR.inner(pb(d), A)
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    assert max(max(r[1].shape) for r in plan.table) > 8
