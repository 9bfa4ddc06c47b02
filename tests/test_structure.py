"""CPU: block/sub-block layout of the host mirror vs the oracle restatement,
plus invariants the reference's own tests use
(test/symmetries/spaces.jl:343-392, test/symmetries/singletree.jl:20-26)."""
import itertools
import math

import numpy as np
import pytest

from tkmi.sectors import U1Irrep, SU2Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, ProductSpace, HomSpace, U1Space, SU2Space

from oracle import abelian as O
from helpers import (cfg1_spaces, cfg2_spaces, cfg2_degeneracies, oracle_legs, kind_of, random_u1_space,
                     random_perm)


def _compare_layout(W):
    total, blocks, subs = W.degeneracystructure()
    cod, dom = oracle_legs(W)
    kind = kind_of(W)
    ototal, oblocks, osubs = O.layout(cod, dom, kind)
    assert total == ototal
    assert [(b[1], b[2]) for b in blocks] == [((b[1], b[2]), b[3]) for b in oblocks]
    # sub-blocks in the same order with identical (sizes, strides, offset)
    ours = list(subs.values())
    theirs = list(osubs.values())
    assert len(ours) == len(theirs)
    for (sz, st, off), (osz, ost, ooff) in zip(ours, theirs):
        assert sz == osz and off == ooff
        assert all(a == b for a, b, n in zip(st, ost, sz) if n > 1)


def test_u1_sector_order_documented():
    # isless: |q| first, positive before negative (docs/src/man/sectors.md:262-264)
    qs = [U1Irrep(q) for q in (0, 1, -1, 2, -2, 0.5, -0.5)]
    assert [s.charge for s in sorted(qs)] == [0, 0.5, -0.5, 1, -1, 2, -2]
    # the oracle's independent key agrees
    raw = sorted([0, 2, -2, 4, -4, 1, -1], key=O.U1Kind.key)
    assert raw == [0, 1, -1, 2, -2, 4, -4]


def test_dual_space_sector_order():
    V = U1Space((0, 2), (1, 3), (-1, 1))
    assert [s.charge for s in V.sectors()] == [0, 1, -1]
    assert [s.charge for s in V.dual().sectors()] == [0, -1, 1]


def test_cfg1_layout():
    V, P = cfg1_spaces()
    assert V.dim() == 56 and sorted(c.charge for c in V.sectors()) == [-3, -2, -1, 0, 1, 2, 3]
    A = HomSpace(ProductSpace(V, P), ProductSpace(V))
    total, blocks, subs = A.degeneracystructure()
    assert total == 1728 and len(blocks) == 7 and len(subs) == 19
    L = HomSpace(ProductSpace(V), ProductSpace(V))
    assert L.dim() == 624
    _compare_layout(A)
    _compare_layout(L)


def test_cfg2_sizes():
    deg = cfg2_degeneracies(mode="round")
    assert sum(deg.values()) == 2052
    assert [deg[q] for q in range(0, 9)] == [290, 272, 226, 165, 107, 61, 31, 14, 5]
    assert sum(cfg2_degeneracies(mode="lr").values()) == 2048
    V, P, W = cfg2_spaces()
    x = HomSpace(ProductSpace(V, P, P), ProductSpace(V))
    assert len(x.blocksectors()) == 17
    assert len(x.fusiontrees()) == 66
    assert x.dim() == 1659624
    _compare_layout(x)


@pytest.mark.parametrize("seed", range(12))
def test_random_u1_layouts(seed):
    rng = np.random.default_rng(seed)
    n1 = int(rng.integers(0, 4))
    n2 = int(rng.integers(0, 4))
    if n1 + n2 == 0:
        n1 = 1
    cod = [random_u1_space(rng) for _ in range(n1)]
    dom = [random_u1_space(rng) for _ in range(n2)]
    W = HomSpace(ProductSpace(*cod), ProductSpace(*dom))
    _compare_layout(W)
    # permuted spaces as well
    for _ in range(3):
        p1, p2 = random_perm(rng, n1, n2)
        _compare_layout(W.permute(p1, p2))


def test_fz2u1_layout():
    rng = np.random.default_rng(7)
    mk = lambda: GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): int(rng.integers(1, 3))
                              for p in (0, 1) for q in (-1, 0, 1)}, dual=bool(rng.integers(0, 2)))
    W = HomSpace(ProductSpace(mk(), mk()), ProductSpace(mk()))
    _compare_layout(W)
    _compare_layout(W.permute((2, 0), (1,)))


def test_blockdim_identities():
    # Σ_c blockdim(P, c) dim(c) = dim(P) (test/symmetries/spaces.jl:343-392)
    V = SU2Space((0, 2), (0.5, 1), (1, 3))
    P = ProductSpace(V, V, V.dual())
    u = SU2Irrep(0)
    tot = sum(P.blockdim(c) * c.dim() for c in P.blocksectors(u))
    assert tot == V.dim() ** 3
    V1, P1 = cfg1_spaces()
    PP = ProductSpace(V1, P1, V1.dual())
    assert sum(PP.blockdim(c) for c in PP.blocksectors(U1Irrep(0))) == V1.dim() ** 2 * 3


def test_su2_tree_counts():
    # number of splitting trees = multiplicity of c in a⊗b⊗c... (test/symmetries/singletree.jl:20-26)
    from tkmi.fusiontrees import fusiontrees
    h = SU2Irrep(0.5)
    # (1/2)^⊗4 -> 0 has multiplicity 2, -> 1 has 3, -> 2 has 1
    for c, n in ((0, 2), (1, 3), (2, 1)):
        assert len(fusiontrees([(h,)] * 4, SU2Irrep(c), (False,) * 4)) == n
    one = SU2Irrep(1)
    # 1⊗1⊗1 -> 1 multiplicity 3, -> 0: 1
    assert len(fusiontrees([(one,)] * 3, SU2Irrep(1), (False,) * 3)) == 3
    assert len(fusiontrees([(one,)] * 3, SU2Irrep(0), (False,) * 3)) == 1


def test_cfg3_su2_block_counts():
    V = SU2Space((0, 64), (0.5, 96), (1, 96), (1.5, 64), (2, 32))
    assert V.dim() == 64 + 96 * 2 + 96 * 3 + 64 * 4 + 32 * 5
    W = HomSpace(ProductSpace(V, V), ProductSpace(V, V))
    assert len(W.fusionblocks()) == 269
    assert len(W.fusiontrees()) == 427
    assert W.dim() == 10355736576
