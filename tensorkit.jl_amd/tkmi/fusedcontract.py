"""Contraction with a small operand as one transform.

`blas_contract!` (`src/tensors/tensoroperations.jl:392-462`) runs C = β·C + α·P_C(P_A(A)·P_B(B))
as two operand permutes, a `mul!` per coupled sector and (when C is not in GEMM layout) a
result permute.  When B is tiny and every sector's GEMM is skinny (the MPO steps of an
effective Hamiltonian: A ~ 10⁷ elements, B ~ 10¹), the permutes and the skinny GEMM are
three bandwidth-bound passes over the large operands.  For abelian sectors the three
linear maps compose, sub-block by sub-block, into one generic transform:

  destination slice  = a column of a sub-block of the GEMM-layout product C′ (its codomain
                       legs, one index tuple of its domain legs fixed), placed in C
                       through the result permute;
  source slice       = a column of an A' = P_A(A) sub-block, i.e. a strided slice of the
                       A sub-block it comes from (the contracted legs fixed);
  U[dst, src]        = coeff_A · coeff_B · coeff_C · B[pos]: the coefficients of the three
                       permutes times the element of B that the GEMM multiplies in.

U depends on B's data, so the plan keeps (coeff, position) pairs and the device gathers
U from B at every call (tkmi_transform_exec_gather).  A is read once and C written once.
The composition only chains the operand-permute tables and the block layouts that the
unfused path uses; it introduces no new fusion-tree coefficients.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import _lib
from . import treetransformers as TT
from .sectors import UNIQUE

MAX_B_ELEMENTS = 4096  # "small operand"
SKINNY_N, SKINNY_K = 32, 64  # every sector's GEMM must be skinny for the fusion to pay


def _pair_map(Wdst, Wsrc, p1, p2, dst_twist):
    """Destination tree pair of an abelian permute → (coeff, src sizes, src strides,
    src offset); the permutation maps destination dim d to source dim (p1+p2)[d]."""
    tab = TT.abelian_table(Wdst, Wsrc, tuple(p1), tuple(p2), None, tuple(dst_twist))
    n = Wsrc.numind
    out = {}
    for (spair, dpair), c, srow in zip(tab["pairs"], tab["coeffs"], tab["src"]):
        out[dpair] = (float(c), tuple(srow[:n]), tuple(srow[n:2 * n]), int(srow[2 * n]))
    return out


def eligible(C, A, pA, B, pB):
    """The fused form applies to abelian sectors, plain (non-adjoint) operands of one real
    element type (the transform reads A, B and C with C's, and its coefficients, which
    hold B's elements, are real), a tiny B and skinny GEMMs in every coupled sector.  Mixed element types fall through to the unfused path,
    which raises for them."""
    if C.space.unit.fusion_style != UNIQUE:
        return False
    if any(type(t).__name__ != "TensorMap" for t in (A, B, C)):
        return False
    if not (A.dtype == B.dtype == C.dtype) or A.dtype.is_complex:
        return False
    if B.space.dim() > MAX_B_ELEMENTS or len(pA[0]) == 0:
        return False
    WB2 = B.space.permute(*pB)
    for c, (k, n), off in WB2.degeneracystructure()[1]:
        if n > SKINNY_N or k > SKINNY_K:
            return False
    return True


def _pair_map_from_src(Wdst, Wsrc, p1, p2):
    """Source tree pair of an abelian permute → (coeff, dst sizes, dst strides, dst
    offset); destination dim d comes from source dim (p1+p2)[d]."""
    tab = TT.abelian_table(Wdst, Wsrc, tuple(p1), tuple(p2))
    n = Wdst.numind
    out = {}
    for (spair, dpair), c, drow in zip(tab["pairs"], tab["coeffs"], tab["dst"]):
        out[spair] = (float(c), tuple(drow[:n]), tuple(drow[n:2 * n]), int(drow[2 * n]))
    return out


def fused_table(C, A, pA, B, pB, pC, twistA_legs, twistB_legs):
    """Rows (U coefficient table, destination sizes, destination slices, source slices)
    of the fused transform, plus the gather table (coeff, position in B) of every U
    entry in plan order (each block's R×C matrix column-major).  pC: the permutation
    taking the GEMM-layout product C′ = A′·B′ to C (tensoradd!(C, C′, pC))."""
    WA2 = A.space.permute(*pA)
    WB2 = B.space.permute(*pB)
    WC2 = WA2.compose(WB2)  # the product in GEMM layout
    amap = _pair_map(WA2, A.space, pA[0], pA[1], twistA_legs)
    bmap = _pair_map(WB2, B.space, pB[0], pB[1], twistB_legs)
    cmap = _pair_map_from_src(C.space, WC2, pC[0], pC[1])
    permA, permB = tuple(pA[0]) + tuple(pA[1]), tuple(pB[0]) + tuple(pB[1])
    permC = tuple(pC[0]) + tuple(pC[1])
    invC = [0] * len(permC)
    for d, e in enumerate(permC):
        invC[e] = d  # C′ dim e is C dim invC[e]
    n1 = WC2.numout  # codomain legs of C′ = codomain legs of A′ (the slice dims)
    subA, subC2 = WA2.subblockstructure(), WC2.subblockstructure()
    WC = WC2

    bands = {}  # (coupled, f1) → C pairs / A' pairs
    for pair in WC.fusiontrees():
        bands.setdefault((pair[0].coupled, pair[0]), {"C": [], "A": []})["C"].append(pair)
    for pair in WA2.fusiontrees():
        key = (pair[0].coupled, pair[0])
        if key in bands:
            bands[key]["A"].append(pair)
    # B' sub-block covering a (row, col) of block c: index by (c, g) and (c, f2)
    b_by_tree = {}
    for pair in WB2.fusiontrees():
        b_by_tree[(pair[0].coupled, pair[0], pair[1])] = pair

    rows, gcoef, gidx = [], [], []
    for (c, f1), band in bands.items():
        dst = []  # (C′ pair, q multi-index, slice of C, coefficient of the result permute)
        for cpair in band["C"]:
            sz = subC2[cpair][0]
            cc, szc, stc, offc = cmap[cpair]
            dom = sz[n1:]
            for q in itertools.product(*[range(d) for d in dom]):
                o = offc + sum(qi * stc[invC[n1 + i]] for i, qi in enumerate(q))
                dst.append((cpair, q, [stc[invC[e]] for e in range(n1)] + [o], cc))
        src = []  # (A' pair, kappa multi-index, slice)
        for apair in band["A"]:
            ca, szs, sts, offs = amap[apair]
            szA = subA[apair][0]
            dom = szA[n1:]
            for kap in itertools.product(*[range(d) for d in dom]):
                o = offs + sum(ki * sts[permA[n1 + i]] for i, ki in enumerate(kap))
                src.append((apair, kap, [sts[permA[d]] for d in range(n1)] + [o]))
        if not dst:
            continue
        R, Cn = len(dst), len(src)
        # U column-major: entry (j, i) at j + i·R
        coef = np.zeros(R * Cn)
        idx = np.full(R * Cn, -1, dtype=np.int64)
        for i, (apair, kap, _) in enumerate(src):
            ca = amap[apair][0]
            g = apair[1]
            for j, (cpair, q, _, cc) in enumerate(dst):
                f2 = cpair[1]
                bpair = b_by_tree.get((c, g, f2))
                if bpair is None:
                    continue
                cb, szb, stb, offb = bmap[bpair]
                mi = tuple(kap) + tuple(q)  # B′ multi-index: codomain (contracted) then domain
                pos = offb + sum(m * stb[permB[d]] for d, m in enumerate(mi))
                coef[j + i * R] = ca * cb * cc
                idx[j + i * R] = pos
        szd = tuple(subC2[band["C"][0]][0][:n1])
        rows.append((R, Cn, szd, [d[2] for d in dst], [s_[2] for s_ in src]))
        gcoef.append(coef)
        gidx.append(idx)
    return rows, (np.concatenate(gcoef + [np.zeros(0)]), np.concatenate(gidx + [np.zeros(0, dtype=np.int64)]))


def pack_fused(rows):
    return {
        "rows": np.array([r[0] for r in rows], dtype=np.int64),
        "cols": np.array([r[1] for r in rows], dtype=np.int64),
        "U": np.zeros(sum(r[0] * r[1] for r in rows)),
        "sz_dst": np.array([r[2] for r in rows], dtype=np.int64).reshape(-1),
        "sz_src": np.array([r[2] for r in rows], dtype=np.int64).reshape(-1),
        "so_dst": np.array([x for r in rows for x in r[3]], dtype=np.int64).reshape(-1),
        "so_src": np.array([x for r in rows for x in r[4]], dtype=np.int64).reshape(-1),
    }


_FUSED_CACHE = TT._LRU(TT.DEFAULT_GLOBALCACHE_SIZE)


def fused_plan(C, A, pA, B, pB, pC, twistA_legs, twistB_legs):
    key = (C.space, A.space, B.space, tuple(map(tuple, pA)), tuple(map(tuple, pB)), tuple(map(tuple, pC)),
           tuple(twistA_legs), tuple(twistB_legs))
    plan = _FUSED_CACHE.get(key)
    if plan is None:
        rows, (coef, idx) = fused_table(C, A, pA, B, pB, pC, twistA_legs, twistB_legs)
        n = len(pA[0])  # slice rank: the codomain legs of A′ (= those of the GEMM-layout product)
        plan = TT.plan_from_packed("fused", tuple(range(n)), pack_fused(rows), C.space.dim(), A.space.dim(),
                                   table=rows)
        ca, cp = _lib.f64(coef)
        ia, ip = _lib.i64(idx)
        _lib.check(_lib.load().tkmi_transform_plan_set_gather(plan.handle, len(coef), cp, ip),
                   "tkmi_transform_plan_set_gather")
        plan.gather = (coef, idx)
        plan.flops = float(sum(2.0 * m * n_ * k for _, (m, n_, k) in _gemm_shapes(C, A, pA, B, pB)))
        _FUSED_CACHE.put(key, plan)
    return plan


def _gemm_shapes(C, A, pA, B, pB):
    """(coupled, (m, n, k)) of the per-sector GEMMs the fused plan replaces."""
    LA = {c: (d1, d2) for c, (d1, d2), off in A.space.permute(*pA).degeneracystructure()[1]}
    LB = {c: (d1, d2) for c, (d1, d2), off in B.space.permute(*pB).degeneracystructure()[1]}
    return [(c, (LA[c][0], LB[c][1], LA[c][1])) for c in LA if c in LB]
