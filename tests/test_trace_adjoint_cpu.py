"""CPU: the pullback table of a partial trace (`trace_permute_pullback_ΔA!`,
src/pullbacks/tensoroperations.jl:50-64).

The forward trace tables are pinned by dense embeddings (test_trace_cpu.py).  The
pullback must be their adjoint for TensorKit's inner product ⟨x, y⟩ = Σ_c dim(c)·⟨x_c, y_c⟩
(the inner product the reference's AD rules and tests use): ⟨ΔC, T·x⟩ = ⟨T‡·ΔC, x⟩
for random x, ΔC.  SU(2) cases where the traced coupled sector differs from the kept one
make the dim(c) weights matter; a table with the weights dropped fails the identity."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space
from tkmi.treetransformers import trace_table, trace_adjoint_table

from test_trace_cpu import apply_rows


def apply_adjoint_rows(rows, dC, dA, alpha):
    out = dA.copy()
    for U, szd, dsts, srcs in rows:
        sv = [np.lib.stride_tricks.as_strided(dC[s[-1]:], shape=szd, strides=tuple(8 * v for v in s[:-1]))
              for s in srcs]
        for r, d in enumerate(dsts):
            dv = np.lib.stride_tricks.as_strided(out[d[-1]:], shape=szd, strides=tuple(8 * v for v in d[:-1]))
            dv += alpha * sum((U[r, k] * sv[k] for k in range(len(srcs))), np.zeros(szd))
    return out


def winner(W, x, y):
    tot = 0.0
    for c, (d1, d2), off in W.degeneracystructure()[1]:
        tot += c.dim() * float(np.dot(x[off:off + d1 * d2], y[off:off + d1 * d2]))
    return tot


def _fermion(pairs):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})


def _cases():
    V1, V2, V3 = U1Space((0, 2), (1, 1), (-1, 2)), U1Space((0, 1), (1, 2), (2, 1)), U1Space((0, 1), (-1, 1))
    yield "u1_rank6", HomSpace(ProductSpace(V1, V2.dual(), V3, V2, V1.dual(), V3.dual()), ProductSpace()), \
        ((5, 2), ()), ((0, 1), (4, 3))
    V1, V2, V3 = U1Space((0, 2), (1, 1)), U1Space((0, 1), (-1, 2)), U1Space((0, 2), (1, 1), (-1, 1))
    yield "u1_map", HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V1, V2, V3)), ((5,), (2,)), ((0, 1), (3, 4))
    F1 = _fermion([((0, 0), 2), ((1, 1), 1), ((1, -1), 2)])
    F2 = _fermion([((0, 0), 1), ((1, 1), 2)])
    yield "fermion_map", HomSpace(ProductSpace(F1, F2, F1), ProductSpace(F1, F2)), ((0, 2), (3,)), ((1,), (4,))
    yield "fermion_dual", HomSpace(ProductSpace(F1, F2.dual(), F2), ProductSpace(F1)), ((0,), (3,)), ((1,), (2,))
    S1, S2, S3 = SU2Space((0, 1), (0.5, 1)), SU2Space((0.5, 1), (1, 1)), SU2Space((0, 1), (1, 1))
    yield "su2_rank6", HomSpace(ProductSpace(S1, S2.dual(), S3, S2, S1.dual(), S3.dual()), ProductSpace()), \
        ((5, 2), ()), ((0, 1), (4, 3))
    yield "su2_map", HomSpace(ProductSpace(S1, S2, S3), ProductSpace(S1, S2, S3)), ((5,), (2,)), ((0, 1), (3, 4))
    S4 = SU2Space((0.5, 2), (1.5, 1))
    yield "su2_pair", HomSpace(ProductSpace(S4, S2), ProductSpace(S4, S2)), ((0,), (2,)), ((1,), (3,))


@pytest.mark.parametrize("name,W,p,q", list(_cases()), ids=[c[0] for c in _cases()])
def test_trace_pullback_is_weighted_adjoint(name, W, p, q):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    Wd = W.select(*p)
    fwd = trace_table(Wd, W, *p, *q)
    adj = trace_adjoint_table(Wd, W, *p, *q)
    x = rng.standard_normal(W.dim())
    dC = rng.standard_normal(Wd.dim())
    dA0 = rng.standard_normal(W.dim())
    Tx = apply_rows(fwd, x, np.zeros(Wd.dim()), 1.0, 0.0)
    dA = apply_adjoint_rows(adj, dC, dA0, 0.75)
    lhs = 0.75 * winner(Wd, dC, Tx)
    rhs = winner(W, dA - dA0, x)          # accumulation: only T‡ΔC was added
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)
    assert abs(lhs) > 1e-6
    # every destination slice of the pullback lies in exactly one row
    seen = set()
    for U, szd, dsts, srcs in adj:
        for d in dsts:
            assert d[-1] not in seen
            seen.add(d[-1])
    if name in ("su2_map", "su2_pair"):
        # planted mistake: the plain transpose of the forward table (no dim(c) weights);
        # these two cases trace between different coupled sectors, so the weights matter
        dAb = dA0.copy()
        for U, szd, sod, slices in fwd:
            so = sod[0]
            src = np.lib.stride_tricks.as_strided(dC[so[-1]:], shape=szd, strides=tuple(8 * v for v in so[:-1]))
            for c, sl in zip(U[0], slices):
                dv = np.lib.stride_tricks.as_strided(dAb[sl[-1]:], shape=szd, strides=tuple(8 * v for v in sl[:-1]))
                dv += 0.75 * c * src
        assert abs(lhs - winner(W, dAb - dA0, x)) > 1e-6 * abs(lhs)
