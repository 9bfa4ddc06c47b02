"""CPU: flip of fusion-tree pairs and spaces (`braiding_manipulations.jl:388-436`,
`homspace.jl:252-261`, `gradedspace.jl:116-122`), with the properties the reference
states: flip followed by flip(inv=true) restores the tree pair with coefficient 1, and
only four flips of the same leg compose to the identity."""
import itertools

import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space
from tkmi.treetransformers import flip_pair


def _fermion(pairs):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})


def _spaces():
    V = U1Space((0, 1), (1, 1), (-1, 1))
    yield "u1", HomSpace(ProductSpace(V, V.dual(), V), ProductSpace(V, V.dual()))
    F = _fermion([((0, 0), 1), ((1, 1), 1), ((1, -1), 1)])
    yield "fermion", HomSpace(ProductSpace(F, F.dual(), F), ProductSpace(F.dual(), F))
    S = SU2Space((0.5, 1), (1, 1))
    yield "su2", HomSpace(ProductSpace(S, S.dual(), S), ProductSpace(S, S.dual()))


@pytest.mark.parametrize("name,W", list(_spaces()), ids=[n for n, _ in _spaces()])
def test_flip_pair_inverse_and_order_four(name, W):
    n = W.numind
    pairs = list(W.fusiontrees())
    assert pairs
    for (f1, f2) in pairs:
        for i in range(n):
            g, c1 = flip_pair(f1, f2, (i,), False)
            h, c2 = flip_pair(*g, (i,), True)
            assert h == (f1, f2) and c1 * c2 == 1
            h, c2 = flip_pair(f1, f2, (i,), True)
            k, c3 = flip_pair(*h, (i,), False)
            assert k == (f1, f2) and c2 * c3 == 1
            # four flips of one leg: the identity
            pair, c = (f1, f2), 1.0
            for _ in range(4):
                pair, ci = flip_pair(*pair, (i,), False)
                c *= ci
            assert pair == (f1, f2) and c == 1
        # several legs at once = one after the other
        for inds in itertools.combinations(range(n), 2):
            g, c = flip_pair(f1, f2, inds, False)
            g1, ca = flip_pair(f1, f2, inds[:1], False)
            g2, cb = flip_pair(*g1, inds[1:], False)
            assert g == g2 and c == ca * cb


@pytest.mark.parametrize("name,W", list(_spaces()), ids=[n for n, _ in _spaces()])
def test_flip_space(name, W):
    for i in range(W.numind):
        W1 = W.flip((i,))
        assert W1.dim() == W.dim()
        assert W1[i].isdual != W[i].isdual
        assert W1.flip((i,)) == W
        for j in range(W.numind):
            if j != i:
                assert W1[j] == W[j]
