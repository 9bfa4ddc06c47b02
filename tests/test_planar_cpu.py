"""CPU: the planar tree moves (folds, cyclic rotations, transpose, planar traces) and the
index bookkeeping of planarcontract! (`reorder_indices`), host side only.

* transpose((f₁, f₂), p) followed by the inverse cyclic permutation returns the pair with
  coefficient 1 (the reference's `U * U′ == 1`, test/symmetries/doubletree.jl:153-155);
  one clockwise rotation undoes one anticlockwise rotation;
* with bosonic braiding the planar coefficients are the braided ones (`U″ ≈ U`,
  doubletree.jl:157-160): transpose_pair equals braid_pair for U1 trees;
* for FermionParity ⊠ U1 every bend, fold, rotation and planar trace coefficient is 1
  (F = A = B = 1, χ = 1, dimensions 1), while braid_pair gives Koszul signs: the planar
  transpose of a fermionic pair differs from its permute exactly by those signs;
* reorder_indices puts A's and B's legs in cyclic orders that transpose accepts, keeps
  the pairing of contracted legs and puts C's legs in pAB's order, on the reference's own
  planarcontract example (test/tensors/planar.jl:88-99) and on contractions with no open
  leg on one side.
"""
import itertools

import numpy as np
import pytest

from tkmi import fusiontrees as _ft  # noqa: F401  (module import for the planar moves)
from tkmi.sectors import U1Irrep, FermionParity, ProductSector, PlanarTrivial
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, PlanarSpace
import sys

FT = sys.modules["tkmi.fusiontrees"]


def _fspace(pairs, dual=False):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs}, dual)


def _cyclic_perms(n1, n2):
    """Every (p₁, p₂) with the linearised order a rotation of (0..n1-1, reverse(domain))."""
    n = n1 + n2
    lin = list(range(n1)) + list(range(n1 + n2 - 1, n1 - 1, -1))
    out = []
    for r in range(n):
        rot = lin[r:] + lin[:r]
        for k in range(n + 1):
            out.append((tuple(rot[:k]), tuple(reversed(rot[k:]))))
    return out


def _inverse(p1, p2, n1):
    """(q₁, q₂) that maps transpose(t, (p₁, p₂)) back to t."""
    p = list(p1) + list(p2)
    inv = [p.index(i) for i in range(len(p))]
    return tuple(inv[:n1]), tuple(inv[n1:])


def _pairs(W):
    return list(W.fusiontrees())


def _spaces():
    fV = _fspace([((0, 0), 2), ((1, 1), 1), ((1, -1), 2), ((0, 2), 1)])
    fP = _fspace([((0, 0), 1), ((1, 1), 1)])
    uV = U1Space((0, 2), (1, 1), (-1, 2), (2, 1))
    uP = U1Space((0, 1), (1, 1), (-1, 1))
    return {
        "fermion": HomSpace(ProductSpace(fV, fP.dual()), ProductSpace(fP, fV)),
        "u1": HomSpace(ProductSpace(uV, uP), ProductSpace(uP.dual(), uV)),
    }


@pytest.mark.parametrize("case", ["fermion", "u1"])
def test_transpose_then_inverse_is_identity(case):
    W = _spaces()[case]
    n1, n2 = W.numout, W.numin
    for p1, p2 in _cyclic_perms(n1, n2):
        q1, q2 = _inverse(p1, p2, n1)
        for f1, f2 in _pairs(W):
            (g1, g2), c = FT.transpose_pair(f1, f2, p1, p2)
            (h1, h2), c2 = FT.transpose_pair(g1, g2, q1, q2)
            assert (h1, h2) == (f1, f2), (p1, p2)
            assert abs(c * c2 - 1) < 1e-14


@pytest.mark.parametrize("case", ["fermion", "u1"])
def test_rotations_undo_each_other(case):
    W = _spaces()[case]
    for f1, f2 in _pairs(W):
        (g1, g2), c = FT.cycleclockwise_pair(f1, f2)
        (h1, h2), c2 = FT.cycleanticlockwise_pair(g1, g2)
        assert (h1, h2) == (f1, f2)
        assert abs(c * c2 - 1) < 1e-14


def test_bosonic_transpose_is_permute():
    W = _spaces()["u1"]
    for p1, p2 in _cyclic_perms(W.numout, W.numin):
        for f1, f2 in _pairs(W):
            assert FT.transpose_pair(f1, f2, p1, p2) == FT.braid_pair(f1, f2, p1, p2)


def test_fermionic_transpose_has_no_braiding_signs():
    """Coefficient 1 for every fermionic transpose; the braided permute of the same pair
    carries the Koszul sign of its crossings, which is −1 for some pairs."""
    W = _spaces()["fermion"]
    signs = set()
    for p1, p2 in _cyclic_perms(W.numout, W.numin):
        for f1, f2 in _pairs(W):
            g, c = FT.transpose_pair(f1, f2, p1, p2)
            gb, cb = FT.braid_pair(f1, f2, p1, p2)
            assert g == gb and c == 1
            signs.add(cb)
    assert signs == {1, -1}


def test_fermionic_planar_trace_coefficients():
    """Planar traces of neighbouring dual legs: coefficient 1 for the pairs whose traced
    legs carry dual sectors, none for the others."""
    fV = _fspace([((0, 0), 2), ((1, 1), 1), ((1, -1), 1)])
    W = HomSpace(ProductSpace(fV, fV, fV.dual()), ProductSpace(fV))
    # trace legs 1 and 2 (neighbours), keep (0,), (3,)
    n = 0
    for f1, f2 in _pairs(W):
        out = FT.planar_trace_pair(f1, f2, (0,), (3,), (1,), (2,))
        if f1.uncoupled[1] == f1.uncoupled[2].dual():
            assert list(out.values()) == [1], (f1, f2)
            (g1, g2), = out.keys()
            assert g1.uncoupled == (f1.uncoupled[0],) and g2.uncoupled == f2.uncoupled
            n += 1
        else:
            assert out == {}
    assert n > 0


def test_planar_trivial_sector():
    """ℙ spaces: one braiding-free sector; the planar moves run with coefficient 1 and a
    braid is refused."""
    V = PlanarSpace(3)
    W = HomSpace(ProductSpace(V, V.dual()), ProductSpace(V))
    pairs = _pairs(W)
    assert len(pairs) == 1
    f1, f2 = pairs[0]
    for p1, p2 in _cyclic_perms(2, 1):
        assert FT.transpose_pair(f1, f2, p1, p2)[1] == 1
    with pytest.raises(ValueError):
        PlanarTrivial().Rsymbol(PlanarTrivial(), PlanarTrivial())


def _check_reordered(codA, domA, codB, domB, pA, pB, pAB):
    from tkmi.tensors import reorder_indices
    (oA, cA), (cB, oB) = pA, pB
    oA2, cA2, oB2, cB2 = reorder_indices(codA, domA, codB, domB, oA, cA, oB, cB, *pAB)
    # cyclic for the planar transposes of A into (oA2 ← cA2) and of B into (cB2 ← oB2)
    nA1, nA2 = len(codA), len(domA)
    nB1, nB2 = len(codB), len(domB)

    def cyclic(p1, p2, n1, n2):
        lin = FT.linearizepermutation(p1, p2, n1, n2)
        n = len(lin)
        return all(lin[(i + 1) % n] == (lin[i] + 1) % n for i in range(n))

    assert cyclic(oA2, cA2, nA1, nA2) and cyclic(cB2, oB2, nB1, nB2)
    assert sorted(zip(cA2, cB2)) == sorted(zip(cA, cB))
    ab = list(oA) + list(oB)
    assert list(oA2) + list(oB2) == [ab[i] for i in list(pAB[0]) + list(pAB[1])]
    return oA2, cA2, oB2, cB2


def test_reorder_indices_reference_example():
    # A: ℂ²⊗ℂ³ ← ℂ²⊗ℂ⁵⊗ℂ⁴, B: ℂ²⊗ℂ⁴ ← ℂ⁴⊗ℂ³, pA = ((1,3,4),(5,2)), pB = ((2,4),(1,3)),
    # pAB = ((3,2,1),(4,5)), 0-based
    got = _check_reordered((0, 1), (2, 3, 4), (0, 1), (2, 3), ((0, 2, 3), (4, 1)), ((1, 3), (0, 2)),
                           ((2, 1, 0), (3, 4)))
    assert got[0] == (3, 2, 0)


def test_reorder_indices_no_open_legs_on_one_side():
    # every leg of A contracted: C has B's open leg only; the planar pairing runs A's
    # legs against B's in opposite cyclic directions
    got = _check_reordered((0, 1), (2,), (0, 1, 2), (3,), ((), (0, 2, 1)), ((0, 1, 2), (3,)), ((), (0,)))
    assert got[1] == (0, 2, 1) and got[3] == (0, 1, 2)
    # every leg of B contracted
    _check_reordered((0,), (1, 2), (0, 1), (), ((0,), (1, 2)), ((0, 1), ()), ((0,), ()))
    # a pairing that would cross is refused
    from tkmi.tensors import reorder_indices
    with pytest.raises(ValueError):
        reorder_indices((0, 1), (2,), (0, 1, 2), (3,), (), (0, 1, 2), (3,), (0, 1, 2), (), (0,))
    # not planar: the open legs of A are not adjacent in its cyclic order
    with pytest.raises(ValueError):
        reorder_indices((0, 1), (2, 3), (0, 1), (), (0, 2), (1, 3), (), (0, 1), (0, 1), ())


# ---------------------------------------------------------------------------
# the host tables of the planar transposes and traces, applied with numpy
# ---------------------------------------------------------------------------
def _apply_abelian(tab, perm, x, ydim, n):
    y = np.zeros(ydim)
    for c, d, s in zip(tab["coeffs"], tab["dst"], tab["src"]):
        dv = np.lib.stride_tricks.as_strided(y[d[2 * n]:], shape=tuple(d[:n]), strides=tuple(8 * v for v in d[n:2 * n]))
        sv = np.lib.stride_tricks.as_strided(x[s[2 * n]:], shape=tuple(s[:n]), strides=tuple(8 * v for v in s[n:2 * n]))
        dv[...] = c * sv.transpose(perm)
    return y


def _apply_trace(rows, x, ydim):
    y = np.zeros(ydim)
    for coeffs, szd, dsts, slices in rows:
        (d,) = dsts
        dv = np.lib.stride_tricks.as_strided(y[d[-1]:], shape=tuple(szd), strides=tuple(8 * v for v in d[:-1]))
        for c, sl in zip(coeffs.reshape(-1), slices):
            dv += c * np.lib.stride_tricks.as_strided(x[sl[-1]:], shape=tuple(szd),
                                                      strides=tuple(8 * v for v in sl[:-1]))
    return y


def test_fermionic_transpose_tables_match_the_oracle():
    """The transposer's table (planar coefficients, permute's data movement), applied
    with numpy, equals the oracle's planar transpose for every cyclic permutation."""
    from helpers import oracle_legs, kind_of
    from oracle import abelian as O
    from tkmi import treetransformers as TT
    W = _spaces()["fermion"]
    x = np.random.default_rng(8).standard_normal(W.dim())
    for p1, p2 in _cyclic_perms(W.numout, W.numin):
        Wd = W.permute(p1, p2)
        tab = TT.abelian_table(Wd, W, p1, p2, move=FT.transpose_pair)
        got = _apply_abelian(tab, list(p1) + list(p2), x, Wd.dim(), W.numind)
        ref, _, _ = O.transpose(x, *oracle_legs(W), p1, p2, kind_of(W))
        np.testing.assert_array_equal(got, ref)


def test_planar_trace_tables_match_dense():
    """Planar trace tables (fermionic and ℙ) applied with numpy against the dense sums."""
    from helpers import oracle_legs, kind_of
    from oracle import abelian as O
    from tkmi import treetransformers as TT
    rng = np.random.default_rng(9)
    fV = _fspace([((0, 0), 2), ((1, 1), 1), ((1, -1), 2)])
    fW = _fspace([((0, 0), 2), ((1, 1), 1), ((0, 1), 1)])
    cases = [(HomSpace(ProductSpace(fV, fW, fW.dual()), ProductSpace(fV)), ((0,), (3,)), ((1,), (2,)), "awwb->ab"),
             (HomSpace(ProductSpace(fV, fW), ProductSpace(fV)), ((1,), ()), ((0,), (2,)), "axa->x")]
    for W, (p1, p2), (q1, q2), spec in cases:
        x = rng.standard_normal(W.dim())
        Wd = W.select(p1, p2)
        rows = TT.trace_table(Wd, W, p1, p2, q1, q2, planar=True)
        got = _apply_trace(rows, x, Wd.dim())
        kind = kind_of(W)
        ref = np.einsum(spec, O.to_dense(x, *oracle_legs(W), kind))
        np.testing.assert_allclose(O.to_dense(got, *oracle_legs(Wd), kind), ref, rtol=0, atol=1e-13)
    # ℙ: the reference's planartrace case, A: ℙ²⊗ℙ³ ← ℙ²⊗ℙ⁵⊗ℙ⁴, p = ((4,2),(5,)), q = ((1,),(3,))
    W = HomSpace(ProductSpace(PlanarSpace(2), PlanarSpace(3)),
                 ProductSpace(PlanarSpace(2), PlanarSpace(5), PlanarSpace(4)))
    x = rng.standard_normal(W.dim())
    Wd = W.select((3, 1), (4,))
    got = _apply_trace(TT.trace_table(Wd, W, (3, 1), (4,), (0,), (2,), planar=True), x, Wd.dim())
    ref = np.einsum("ijikl->kjl", x.reshape((2, 3, 2, 5, 4), order="F"))
    np.testing.assert_allclose(got.reshape((5, 3, 4), order="F"), ref, rtol=0, atol=1e-13)
