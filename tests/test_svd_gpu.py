"""GPU: svd_compact / svd_trunc (FP64 block SVD + device truncation selection).

Checks follow the reference's own SVD tests (test/factorizations/svd.jl:197-295):
U S Vᴴ reproduces t, ‖t − Ũ S̃ Ṽᴴ‖ = ε, and the kept set / ε agree with a
direct restatement of truncrank ∩ trunctol over LAPACK (numpy) singular values."""
import numpy as np
import pytest

from tkmi.spaces import HomSpace, ProductSpace, U1Space, SU2Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def host_blocks(t):
    return t.blocks_host()


def reference_truncation(sv, howmany, atol):
    """sv: {c: (σ descending, weight)} → kept counts, ε (sequential scan of the
    globally sorted values, src/factorizations/truncation.jl:171-203,212-221,315-322)."""
    allv = [(s, w, c, i) for c, (vals, w) in sv.items() for i, s in enumerate(vals)]
    order = sorted(range(len(allv)), key=lambda j: (-allv[j][0], j))
    keep = [False] * len(allv)
    tot = 0
    for j in order:
        tot += allv[j][1]
        if tot > howmany:
            break
        keep[j] = True
    for j, (s, w, c, i) in enumerate(allv):
        if s < atol:
            keep[j] = False
    counts = {c: 0 for c in sv}
    err2 = 0.0
    for k, (s, w, c, i) in zip(keep, allv):
        if k:
            counts[c] += 1
        else:
            err2 += w * s * s
    return counts, np.sqrt(err2)


ALGS = ["qr_iteration", "divide_and_conquer", "jacobi", "block_jacobi"]
EIGH_ALGS = ["qr_iteration", "divide_and_conquer", "jacobi"]


def _svd_checks(T, t, howmany, atol, alg=None):
    from tkmi.factorizations import svd_compact, svd_trunc
    U, (S, offs), Vt = svd_compact(t, alg)
    Sh = S.cpu().numpy()
    Ab = host_blocks(t)
    Ub, Vb = host_blocks(U), host_blocks(Vt)
    sv = {}
    for c, A in Ab.items():
        o, r = offs[c]
        s_ref = np.linalg.svd(A, compute_uv=False)
        np.testing.assert_allclose(Sh[o:o + r], s_ref, rtol=1e-12, atol=1e-12 * s_ref[0])
        rec = Ub[c] @ np.diag(Sh[o:o + r]) @ Vb[c]
        assert np.linalg.norm(rec - A) <= 1e-12 * np.linalg.norm(A)
        # isometric factors (isisometric(U), isisometric(Vᴴ'))
        assert np.linalg.norm(Ub[c].T @ Ub[c] - np.eye(r)) <= 1e-12 * max(1, r)
        assert np.linalg.norm(Vb[c] @ Vb[c].T - np.eye(r)) <= 1e-12 * max(1, r)
        sv[c] = (s_ref, c.dim())
    Ut, Sk, Vtt, eps = svd_trunc(t, howmany, atol, alg)
    counts, eps_ref = reference_truncation(sv, howmany, atol)
    assert {c: len(v) for c, v in Sk.items()} == {c: n for c, n in counts.items() if n}
    assert abs(eps - eps_ref) <= 1e-10 * max(eps_ref, 1e-300) + 1e-14
    # ‖t − Ũ S̃ Ṽᴴ‖ = ε  (with the dim(c) weight of the Frobenius norm)
    Utb, Vtb = host_blocks(Ut), host_blocks(Vtt)
    d2 = 0.0
    for c, A in Ab.items():
        if c in Sk:
            R = A - Utb[c] @ np.diag(Sk[c].cpu().numpy()) @ Vtb[c]
        else:
            R = A
        d2 += c.dim() * np.sum(R * R)
    tnorm = np.sqrt(sum(c.dim() * np.sum(A * A) for c, A in Ab.items()))
    assert abs(np.sqrt(d2) - eps) <= 1e-9 * eps + 1e-13 * tnorm


@pytest.mark.parametrize("alg", ALGS)
def test_svd_trunc_u1(T, alg):
    rng = np.random.default_rng(0)
    V = U1Space((0, 40), (1, 30), (-1, 30), (2, 12), (-2, 12))
    P = U1Space((0.5, 1), (-0.5, 1))
    t = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(P.dual(), V)), rng)
    _svd_checks(T, t, howmany=60, atol=1e-10, alg=alg)
    _svd_checks(T, t, howmany=10 ** 6, atol=0.0, alg=alg)


@pytest.mark.parametrize("alg", ALGS)
def test_svd_trunc_su2_weights(T, alg):
    rng = np.random.default_rng(1)
    V = SU2Space((0, 6), (0.5, 5), (1, 4), (1.5, 2))
    t = T.TensorMap.randn(HomSpace(ProductSpace(V, V), ProductSpace(V)), rng)
    _svd_checks(T, t, howmany=40, atol=1e-10, alg=alg)


@pytest.mark.parametrize("alg", ALGS)
def test_svd_rank_deficient_tol(T, alg):
    """trunctol discards (near-)zero singular values of a rank-deficient block."""
    rng = np.random.default_rng(2)
    V = U1Space((0, 12), (1, 6))
    W = HomSpace(ProductSpace(V), ProductSpace(V))
    x = np.zeros(W.dim())
    for c, (d1, d2), off in W.degeneracystructure()[1]:
        a = rng.standard_normal((d1, 2)) @ rng.standard_normal((2, d2))
        x[off:off + d1 * d2] = a.reshape(-1, order="F")
    t = T.TensorMap.from_numpy(x, W)
    from tkmi.factorizations import svd_trunc
    Ut, Sk, Vtt, eps = svd_trunc(t, 100, 1e-10, alg)
    assert {c: len(v) for c, v in Sk.items()} == {c: 2 for c in Sk}
    assert eps < 1e-12


def test_svd_unknown_algorithm(T):
    from tkmi.factorizations import svd_compact
    V = U1Space((0, 3), (1, 2))
    t = T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(V)), np.random.default_rng(3))
    with pytest.raises(ValueError):
        svd_compact(t, "gesvd_magic")


# ---------------------------------------------------------------------------
# eigh_full / eigh_trunc
# ---------------------------------------------------------------------------
def _sym_tensor(T, W, rng):
    x = np.zeros(W.dim())
    for c, (d1, d2), off in W.degeneracystructure()[1]:
        a = rng.standard_normal((d1, d1))
        x[off:off + d1 * d1] = ((a + a.T) / 2).reshape(-1, order="F")
    return T.TensorMap.from_numpy(x, W)


def reference_eigh_truncation(vals, howmany, atol):
    """vals: {c: (λ ascending, weight)} → {c: kept mask}, ε: the weighted count over
    values sorted by descending |λ| (ties by position), then |λ| ≥ atol
    (truncrank(howmany; by=abs) ∩ trunctol(; atol), truncation.jl:171-203)."""
    allv = [(abs(l), w, c, i) for c, (ls, w) in vals.items() for i, l in enumerate(ls)]
    order = sorted(range(len(allv)), key=lambda j: (-allv[j][0], j))
    keep = [False] * len(allv)
    tot = 0
    for j in order:
        tot += allv[j][1]
        if tot > howmany:
            break
        keep[j] = True
    masks = {c: np.zeros(len(ls), dtype=bool) for c, (ls, w) in vals.items()}
    err2 = 0.0
    for k, (a, w, c, i) in zip(keep, allv):
        k = k and a >= atol
        masks[c][i] = k
        if not k:
            err2 += w * a * a
    return masks, np.sqrt(err2)


@pytest.mark.parametrize("alg", EIGH_ALGS)
def test_eigh_full_and_trunc(T, alg):
    from tkmi.factorizations import eigh_full, eigh_trunc
    rng = np.random.default_rng(30)
    for V, howmany in ((U1Space((0, 9), (1, 7), (-1, 6), (2, 3)), 40), (SU2Space((0, 4), (0.5, 3), (1, 2)), 60)):
        W = HomSpace(ProductSpace(V, V), ProductSpace(V, V))
        t = _sym_tensor(T, W, rng)
        (lam, offs), Vm = eigh_full(t, alg)
        Ab, Vb, lh = t.blocks_host(), Vm.blocks_host(), lam.cpu().numpy()
        vals = {}
        for c, A in Ab.items():
            o, r = offs[c]
            ref = np.linalg.eigvalsh(A)
            np.testing.assert_allclose(lh[o:o + r], ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))
            assert np.linalg.norm(Vb[c] @ np.diag(lh[o:o + r]) @ Vb[c].T - A) <= 1e-12 * np.linalg.norm(A)
            assert np.linalg.norm(Vb[c].T @ Vb[c] - np.eye(r)) <= 1e-12 * max(1, r)
            vals[c] = (ref, c.dim())
        Dk, Vt, eps = eigh_trunc(t, howmany, 1e-3, alg)
        masks, eps_ref = reference_eigh_truncation(vals, howmany, 1e-3)
        assert {c: len(v) for c, v in Dk.items()} == {c: int(m.sum()) for c, m in masks.items() if m.any()}
        assert abs(eps - eps_ref) <= 1e-10 * max(eps_ref, 1e-300) + 1e-13
        Vtb = Vt.blocks_host()
        d2 = 0.0
        for c, A in Ab.items():
            R = A.copy()
            if c in Dk:
                dk = Dk[c].cpu().numpy()
                np.testing.assert_allclose(dk, vals[c][0][masks[c]], rtol=1e-12, atol=1e-12)
                R = A - Vtb[c] @ np.diag(dk) @ Vtb[c].T
            d2 += c.dim() * np.sum(R * R)
        tnorm = np.sqrt(sum(c.dim() * np.sum(A * A) for c, A in Ab.items()))
        assert abs(np.sqrt(d2) - eps) <= 1e-9 * eps + 1e-12 * tnorm
