"""CPU: SU(2) symbols and the generic (recoupling) transform tables.

Pins, in order:
  1. the oracle's Clebsch-Gordan closed form (orthonormality, known values);
  2. the host mirror's F-symbols (unitarity, pentagon) and B-symbol /
     Frobenius-Schur phase against the CG-based Z isomorphism
     (docs/src/man/fusiontrees.md:280-289);
  3. SU(2) block layout, host mirror vs oracle;
  4. the generic tree-transformer tables, executed by the oracle's C loop,
     satisfy the reference's dense check
     convert(Array, permute(t, p)) ≈ permutedims(convert(Array, t), p)
     (test/tensors/indexmanipulations.jl:76-98) for every permutation.
"""
import itertools
import math

import numpy as np
import pytest

from tkmi.sectors import SU2Irrep
from tkmi.spaces import HomSpace, ProductSpace, SU2Space
from tkmi.treetransformers import generic_table

from oracle import su2 as S
from oracle import kernels as OK


def leg(V):
    return ({k.two_j: n for k, n in V._dims.items()}, V.isdual)


def legs(W):
    return [leg(V) for V in W.codomain], [leg(V) for V in W.domain]


def test_cg_orthonormality_and_values():
    # <1/2 1/2, 1/2 -1/2 | 0 0> = 1/sqrt(2), <1/2 -1/2, 1/2 1/2 | 0 0> = -1/sqrt(2)
    assert math.isclose(S.clebschgordan(1, 1, 1, -1, 0, 0), 1 / math.sqrt(2), rel_tol=1e-15)
    assert math.isclose(S.clebschgordan(1, -1, 1, 1, 0, 0), -1 / math.sqrt(2), rel_tol=1e-15)
    # <1 1, 1/2 -1/2 | 3/2 1/2> = sqrt(1/3)
    assert math.isclose(S.clebschgordan(2, 2, 1, -1, 3, 1), math.sqrt(1 / 3), rel_tol=1e-15)
    for ta, tb in [(1, 1), (2, 1), (2, 2), (3, 4)]:
        M = np.concatenate([S.cg_tensor(ta, tb, tc).reshape((ta + 1) * (tb + 1), tc + 1)
                            for tc in range(abs(ta - tb), ta + tb + 1, 2)], axis=1)
        np.testing.assert_allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-14)


def _spins(maxtwo):
    return [SU2Irrep(two_j=t) for t in range(maxtwo + 1)]


def test_F_unitarity_and_pentagon():
    js = _spins(3)
    for a, b, c, d in itertools.product(js, repeat=4):
        es = [e for e in a.otimes(b) if d in e.otimes(c)]
        fs = [f for f in b.otimes(c) if d in a.otimes(f)]
        if not es:
            continue
        F = np.array([[a.Fsymbol(b, c, d, e, f) for f in fs] for e in es])
        np.testing.assert_allclose(F @ F.T, np.eye(len(es)), atol=1e-13)
    # pentagon: F^{fcd}_{e;gl} F^{abl}_{e;fk} = Σ_h F^{abc}_{g;fh} F^{ahd}_{e;gk} F^{bcd}_{k;hl}
    js = _spins(2)
    n = 0
    for a, b, c, d in itertools.product(js, repeat=4):
        for f in a.otimes(b):
            for g in f.otimes(c):
                for e in g.otimes(d):
                    for l in c.otimes(d):
                        for k in b.otimes(l):
                            lhs = f.Fsymbol(c, d, e, g, l) * a.Fsymbol(b, l, e, f, k)
                            rhs = sum(a.Fsymbol(b, c, g, f, h) * a.Fsymbol(h, d, e, g, k) * b.Fsymbol(c, d, k, h, l)
                                      for h in b.otimes(c))
                            assert abs(lhs - rhs) < 1e-12, (a, b, c, d, e, f, g, k, l)
                            n += 1
    assert n > 1000


def test_fs_phase_and_Z_transpose():
    for t in range(5):
        Zm = S.Z(t)
        fs = SU2Irrep(two_j=t).frobenius_schur_phase()
        np.testing.assert_allclose(Zm.T, fs * Zm, atol=1e-15)


def test_su2_layout_vs_oracle():
    V = SU2Space((0, 2), (0.5, 1), (1, 2))
    Wd = SU2Space((0.5, 2), (1.5, 1))
    for W in [HomSpace(ProductSpace(V, V), ProductSpace(V)),
              HomSpace(ProductSpace(V, Wd.dual()), ProductSpace(V, Wd)),
              HomSpace(ProductSpace(V, V, V), ProductSpace()),
              HomSpace(ProductSpace(), ProductSpace(V, Wd))]:
        total, blocks, subs = W.degeneracystructure()
        ot, ob, osubs = S.layout(*legs(W))
        assert total == ot
        ours = list(subs.items())
        assert len(ours) == len(osubs)
        for ((f1, f2), (sz, st, off)), ((t1, t2, c), osz, ost, ooff) in zip(ours, osubs):
            assert tuple(s.two_j for s in f1.uncoupled) == t1[0]
            assert tuple(s.two_j for s in f1.innerlines) == t1[1]
            assert tuple(s.two_j for s in f2.uncoupled) == t2[0]
            assert f1.coupled.two_j == c
            assert sz == osz and off == ooff


def _dense_check(W, p1, p2, rng):
    Wd = W.permute(p1, p2)
    rows = generic_table(Wd, W, p1, p2)
    x = rng.standard_normal(W.dim())
    y = np.zeros(Wd.dim())
    OK.generic_permute(rows, tuple(p1) + tuple(p2), x, y)
    A = S.to_dense(x, *legs(W))
    B = S.to_dense(y, *legs(Wd))
    ref = A.transpose(tuple(p1) + tuple(p2))
    err = np.linalg.norm(B - ref) / np.linalg.norm(ref)
    assert err < 1e-13, (p1, p2, err)


def test_generic_tables_dense_all_perms_rank3():
    rng = np.random.default_rng(0)
    V = SU2Space((0, 1), (0.5, 2), (1, 1))
    U = SU2Space((0.5, 1), (1, 2))
    W = HomSpace(ProductSpace(V, U), ProductSpace(V.dual()))
    for p in itertools.permutations(range(3)):
        for k in range(4):
            _dense_check(W, p[:k], p[k:], rng)


@pytest.mark.parametrize("k", range(5))
def test_generic_tables_dense_rank4(k):
    rng = np.random.default_rng(10 + k)
    V = SU2Space((0, 1), (0.5, 1), (1, 1))
    U = SU2Space((0.5, 1), (1.5, 1))
    W = HomSpace(ProductSpace(V, U.dual()), ProductSpace(U, V))
    for p in itertools.permutations(range(4)):
        _dense_check(W, p[:k], p[k:], rng)
