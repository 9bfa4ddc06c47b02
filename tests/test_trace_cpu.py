"""CPU: partial-trace tables (trace_permute!, src/tensors/tensoroperations.jl:203-300).

The tables are executed with numpy strided views and compared with the reference's
"Trace: test via conversion" identity (test/tensors/contractions.jl:62-69):
    @tensor t2[a, b] := t[c, d, b, d, c, a]  ≈  the same sum over convert(Array, t),
using the oracle's independent dense embeddings (U1 placement, SU(2) Clebsch-Gordan).
Also pins `split` of fusion trees against its definition
(src/fusiontrees/basic_manipulations.jl:30-55)."""
import numpy as np
import pytest

from tkmi.fusiontrees import split_tree, FusionTree
from tkmi.sectors import U1Irrep
from tkmi.spaces import HomSpace, ProductSpace, U1Space, SU2Space
from tkmi.treetransformers import trace_table

from helpers import oracle_legs, kind_of
from oracle import abelian as O
from oracle import su2 as S


def apply_rows(rows, x, y, alpha, beta):
    out = beta * y
    for U, szd, sod, slices in rows:
        so = sod[0]
        dst = np.lib.stride_tricks.as_strided(out[so[-1]:], shape=szd, strides=tuple(8 * s for s in so[:-1]))
        acc = np.zeros(szd)
        for c, sl in zip(U[0], slices):
            acc = acc + c * np.lib.stride_tricks.as_strided(x[sl[-1]:], shape=szd,
                                                             strides=tuple(8 * s for s in sl[:-1]))
        dst += alpha * acc
    return out


def test_split_tree_definition():
    a, b, c, d = U1Irrep(1), U1Irrep(-2), U1Irrep(3), U1Irrep(0)
    f = FusionTree((a, b, c, d), U1Irrep(2), (False, True, False, True))
    for M in range(0, 5):
        f1, f2 = split_tree(f, M)
        assert f1.uncoupled == f.uncoupled[:M] and f1.isdual == f.isdual[:M]
        inner = (a,) + f.innerlines + (f.coupled,)
        assert f1.coupled == (U1Irrep(0) if M == 0 else inner[M - 1])
        assert f2.uncoupled == (f1.coupled,) + f.uncoupled[M:]
        assert f2.isdual == (False,) + f.isdual[M:]
        assert f2.coupled == f.coupled
        # join(split(f, M)) == f: f's inner lines are recovered from the two halves
        if M == 0:
            rebuilt = f2.innerlines[1:]
        elif M == len(f.uncoupled):
            rebuilt = f1.innerlines
        else:
            rebuilt = f1.innerlines + ((f1.coupled,) if M >= 2 else ()) + f2.innerlines
        assert rebuilt == f.innerlines


def _u1_rank6(rng):
    V1 = U1Space((0, 2), (1, 1), (-1, 2))
    V2 = U1Space((0, 1), (1, 2), (2, 1))
    V3 = U1Space((0, 1), (-1, 1))
    return HomSpace(ProductSpace(V1, V2.dual(), V3, V2, V1.dual(), V3.dual()), ProductSpace())


def test_u1_trace_via_conversion():
    rng = np.random.default_rng(0)
    W = _u1_rank6(rng)
    x = rng.standard_normal(W.dim())
    p, q = ((5, 2), ()), ((0, 1), (4, 3))
    Wd = W.select(*p)
    rows = trace_table(Wd, W, *p, *q)
    y0 = rng.standard_normal(Wd.dim())
    y = apply_rows(rows, x, y0, 0.5, -1.5)
    kind = kind_of(W)
    dense = O.to_dense(x, *oracle_legs(W), kind)
    ref = -1.5 * O.to_dense(y0, *oracle_legs(Wd), kind) + 0.5 * np.einsum("cdbdca->ab", dense)
    np.testing.assert_allclose(O.to_dense(y, *oracle_legs(Wd), kind), ref, rtol=1e-13, atol=1e-13)


def test_u1_partial_trace_to_map():
    """t[a; b] := t[c d b; c d a] on V1⊗V2⊗V3 ← V1⊗V2⊗V3 (codomain-domain pairs)."""
    rng = np.random.default_rng(1)
    V1, V2, V3 = U1Space((0, 2), (1, 1)), U1Space((0, 1), (-1, 2)), U1Space((0, 2), (1, 1), (-1, 1))
    W = HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V1, V2, V3))
    x = rng.standard_normal(W.dim())
    p, q = ((5,), (2,)), ((0, 1), (3, 4))
    Wd = W.select(*p)
    rows = trace_table(Wd, W, *p, *q)
    y = apply_rows(rows, x, np.zeros(Wd.dim()), 1.0, 0.0)
    kind = kind_of(W)
    dense = O.to_dense(x, *oracle_legs(W), kind)
    np.testing.assert_allclose(O.to_dense(y, *oracle_legs(Wd), kind), np.einsum("cdbcda->ab", dense),
                               rtol=1e-13, atol=1e-13)


def _su2_legs(W):
    return ([({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.codomain],
            [({k.two_j: n for k, n in V._dims.items()}, V.isdual) for V in W.domain])


@pytest.mark.parametrize("case", [0, 1])
def test_su2_trace_via_conversion(case):
    rng = np.random.default_rng(10 + case)
    V1 = SU2Space((0, 1), (0.5, 1))
    V2 = SU2Space((0.5, 1), (1, 1))
    V3 = SU2Space((0, 1), (1, 1))
    if case == 0:
        W = HomSpace(ProductSpace(V1, V2.dual(), V3, V2, V1.dual(), V3.dual()), ProductSpace())
        p, q, spec = ((5, 2), ()), ((0, 1), (4, 3)), "cdbdca->ab"
    else:
        W = HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V1, V2, V3))
        p, q, spec = ((5,), (2,)), ((0, 1), (3, 4)), "cdbcda->ab"
    x = rng.standard_normal(W.dim())
    Wd = W.select(*p)
    rows = trace_table(Wd, W, *p, *q)
    y = apply_rows(rows, x, np.zeros(Wd.dim()), 1.0, 0.0)
    dense = S.to_dense(x, *_su2_legs(W))
    ref = np.einsum(spec, dense)
    got = S.to_dense(y, *_su2_legs(Wd))
    assert np.linalg.norm(got - ref) <= 1e-13 * max(1.0, np.linalg.norm(ref))
