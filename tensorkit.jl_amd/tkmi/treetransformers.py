"""Tree transformers: the tables a permute/braid kernel consumes, built and
cached on the host and uploaded to the device once.

Mirrors `src/tensors/treetransformers.jl`:
- `AbelianTreeTransformer` (`:10-42`): one (coeff, dst structure, src structure)
  entry per source tree pair, in the source's tree order.
- `GenericTreeTransformer` (`:44-120`): per fusion block of the source, the
  recoupling matrix U (dst trees × src trees), the common sub-block size and
  the (strides, offset) of every tree; sorted by descending
  `length(U)*prod(sz)` (`:186-205`).
- `treebraider`/`treetransposer` LRU caches (`:156-180`, `src/auxiliary/caches.jl:135-173`);
  the transposer's coefficients come from the planar moves (bends and folds), with the
  same data movement as a permute.

The serialized table is handed to `tkmi_abelian_plan_create` /
`tkmi_generic_plan_create`, which return an immutable device plan.
"""
from __future__ import annotations

import itertools
from collections import OrderedDict
from contextlib import contextmanager

import numpy as np

from . import _lib
from .fusiontrees import (FusionTree, braid_pair, braid_block, permute_block, split_tree, transpose_pair,
                          planar_trace_pair)
from .sectors import UNIQUE
from .spaces import clear_space_caches

DEFAULT_GLOBALCACHE_SIZE = 10 ** 4  # src/auxiliary/caches.jl:19


class TransformPlan:
    """Owner of one device transform plan (C handle) plus its host table."""

    def __init__(self, handle, kind, nentries, elements, table):
        self.handle = handle
        self._destroy = _lib.load().tkmi_transform_plan_destroy
        self.kind = kind
        self.nentries = nentries
        self.elements = elements
        self.table = table

    def __del__(self):
        # the destroy function is bound at creation: module globals may already be
        # torn down when the interpreter exits
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self.handle = None


# Open pinning() scopes: each is a list that keeps every plan handed out by a cache
# inside the scope alive (see pinning).
_PIN_STACK = []


def _pin(plan):
    for pins in _PIN_STACK:
        pins.append(plan)


@contextmanager
def pinning():
    """Keep every device plan looked up or created inside the scope alive for as long as
    the yielded list lives.  A captured HIP graph holds only the device addresses of the
    plans it launches (tables, split-K workspaces, split panels); a plan evicted from its
    cache and collected afterwards would be freed under the graph."""
    pins = []
    _PIN_STACK.append(pins)
    try:
        yield pins
    finally:
        # by identity: two scopes' lists can hold the same plans and compare equal
        del _PIN_STACK[next(i for i, x in enumerate(_PIN_STACK) if x is pins)]


class _LRU:
    """Bounded plan cache (least recently used evicted first); plans it hands out inside
    a pinning() scope are pinned there."""

    def __init__(self, maxsize):
        self.maxsize = maxsize
        self.d = OrderedDict()
        self.hits = 0
        self.misses = 0
        _ALL_CACHES.append(self)

    def get(self, key):
        v = self.d.get(key)
        if v is not None:
            self.d.move_to_end(key)
            self.hits += 1
            _pin(v)
        else:
            self.misses += 1
        return v

    def put(self, key, v):
        _pin(v)
        self.d[key] = v
        self.d.move_to_end(key)
        while len(self.d) > self.maxsize:
            self.d.popitem(last=False)

    def clear(self):
        self.d.clear()


_ALL_CACHES = []
_TRANSFORMER_CACHE = _LRU(DEFAULT_GLOBALCACHE_SIZE)


def empty_globalcaches():
    """`empty_globalcaches!()` (`src/auxiliary/caches.jl:2-5`): the transformer cache,
    this package's other plan caches (grouped GEMMs, fused contractions) and the shared
    space structure."""
    for c in _ALL_CACHES:
        c.clear()
    clear_space_caches()


def global_cache_info():
    c = _TRANSFORMER_CACHE
    return {"size": len(c.d), "maxsize": c.maxsize, "hits": c.hits, "misses": c.misses}


def _perm0(p1, p2):
    return tuple(p1) + tuple(p2)


def twist_factor(tree_pair, legs, numout):
    """θ = Π_{i∈legs} twist(uncoupled sector at leg i) for a tree pair
    (`twist!`, `src/tensors/indexmanipulations.jl:62-78`)."""
    f1, f2 = tree_pair
    th = 1
    for i in legs:
        s = f1.uncoupled[i] if i < numout else f2.uncoupled[i - numout]
        th *= s.twist()
    return th


def abelian_table(Vdst, Vsrc, p1, p2, src_structure=None, dst_twist=(), move=braid_pair):
    """Rows of `AbelianTreeTransformer(braid, p, Vdst, Vsrc)`.

    `src_structure` overrides the source's sub-block structure (used for the
    lazy adjoint of a tensor); `dst_twist` lists destination legs whose twist
    is folded into the coefficient (the fermionic `twist!` after the copy in
    `blas_contract!`, `src/tensors/tensoroperations.jl:425-441`); `move` is the tree
    pair map (braid_pair for braid/permute, transpose_pair for the planar transpose)."""
    fts_src = src_structure if src_structure is not None else Vsrc.subblockstructure()
    fts_dst = Vdst.subblockstructure()
    n = Vsrc.numind
    coeffs, dsts, srcs, pairs = [], [], [], []
    for pair in Vsrc.fusiontrees():
        f1, f2 = pair
        (g1, g2), coeff = move(f1, f2, p1, p2)
        if dst_twist:
            coeff *= twist_factor((g1, g2), dst_twist, Vdst.numout)
        sd = fts_dst[(g1, g2)]
        ss = fts_src[pair]
        coeffs.append(float(coeff))
        dsts.append(list(sd[0]) + list(sd[1]) + [sd[2]])
        srcs.append(list(ss[0]) + list(ss[1]) + [ss[2]])
        pairs.append((pair, (g1, g2)))
    return {
        "coeffs": np.array(coeffs, dtype=np.float64),
        "dst": np.array(dsts, dtype=np.int64).reshape(len(dsts), 2 * n + 1),
        "src": np.array(srcs, dtype=np.int64).reshape(len(srcs), 2 * n + 1),
        "pairs": pairs,
    }


def generic_table(Vdst, Vsrc, p1, p2, levels1=None, levels2=None, src_structure=None, dst_twist=()):
    """Rows of `GenericTreeTransformer(braid, p, Vdst, Vsrc)` (one per fusion
    block), sorted by descending weight like `sort!(transformer)`."""
    fts_src = src_structure if src_structure is not None else Vsrc.subblockstructure()
    fts_dst = Vdst.subblockstructure()
    n1, n2 = Vsrc.numout, Vsrc.numin
    if levels1 is None:
        levels1 = tuple(range(n1))
    if levels2 is None:
        levels2 = tuple(range(n1, n1 + n2))
    unit = Vsrc.unit
    rows = []
    for fs_src in Vsrc.fusionblocks():
        fs_dst, U = braid_block(fs_src, p1, p2, levels1, levels2, unit)
        if dst_twist:
            th = np.array([twist_factor(t, dst_twist, Vdst.numout) for t in fs_dst.trees], dtype=np.float64)
            U = th[:, None] * U
        sz_src = fts_src[fs_src.trees[0]][0]
        sz_dst = fts_dst[fs_dst.trees[0]][0]
        so_src = [list(fts_src[t][1]) + [fts_src[t][2]] for t in fs_src.trees]
        so_dst = [list(fts_dst[t][1]) + [fts_dst[t][2]] for t in fs_dst.trees]
        w = U.size * int(np.prod(sz_dst))
        rows.append((w, U, sz_dst, so_dst, sz_src, so_src, fs_src, fs_dst))
    rows.sort(key=lambda r: r[0], reverse=True)
    return rows


def pack_abelian(tab):
    """The arrays `tkmi_abelian_plan_create` takes, from an abelian table."""
    return {"coeffs": np.ascontiguousarray(tab["coeffs"], dtype=np.float64),
            "dst": np.ascontiguousarray(np.asarray(tab["dst"]).reshape(-1), dtype=np.int64),
            "src": np.ascontiguousarray(np.asarray(tab["src"]).reshape(-1), dtype=np.int64)}


def pack_generic(rows):
    """The arrays `tkmi_generic_plan_create` takes, from generic-transformer rows; the
    per-tree weights dim(c) enter only the adjoint transform."""
    return {
        "rows": np.array([r[1].shape[0] for r in rows], dtype=np.int64),
        "cols": np.array([r[1].shape[1] for r in rows], dtype=np.int64),
        "U": np.concatenate([np.asarray(r[1], dtype=np.float64).reshape(-1, order="F") for r in rows]
                            + [np.zeros(0)]),
        "sz_dst": np.array([r[2] for r in rows], dtype=np.int64).reshape(-1),
        "sz_src": np.array([r[4] for r in rows], dtype=np.int64).reshape(-1),
        "so_dst": np.array([x for r in rows for x in r[3]], dtype=np.int64).reshape(-1),
        "so_src": np.array([x for r in rows for x in r[5]], dtype=np.int64).reshape(-1),
        "w_dst": np.array([t[0].coupled.dim() for r in rows for t in r[7].trees], dtype=np.float64),
        "w_src": np.array([t[0].coupled.dim() for r in rows for t in r[6].trees], dtype=np.float64),
    }


def plan_from_packed(kind, perm, packed, dst_len, src_len, table=None):
    """Create the device plan of a packed table (the same call whether the table was
    just built or read back from a table file)."""
    lib = _lib.ensure_device()
    h = _lib.C.c_void_p()
    pa, pp = _lib.i32(perm)
    n = len(perm)
    if kind == "abelian":
        nent = len(packed["coeffs"])
        ca, cp = _lib.f64(packed["coeffs"])
        da, dp = _lib.i64(packed["dst"])
        sa, sp = _lib.i64(packed["src"])
        _lib.check(lib.tkmi_abelian_plan_create(_lib.C.byref(h), n, pp, nent, cp, dp, sp, int(dst_len),
                                                int(src_len)), "tkmi_abelian_plan_create")
    else:
        nent = len(packed["rows"])
        keep = [_lib.i64(packed["rows"]), _lib.i64(packed["cols"]), _lib.f64(packed["U"]),
                _lib.i64(packed["sz_dst"]), _lib.i64(packed["sz_src"]), _lib.i64(packed["so_dst"]),
                _lib.i64(packed["so_src"])]
        wd = _lib.f64(packed["w_dst"]) if "w_dst" in packed else (None, None)
        ws = _lib.f64(packed["w_src"]) if "w_src" in packed else (None, None)
        _lib.check(lib.tkmi_generic_plan_create(_lib.C.byref(h), n, pp, nent, keep[0][1], keep[1][1], keep[2][1],
                                                keep[3][1], keep[4][1], keep[5][1], keep[6][1], wd[1], ws[1],
                                                int(dst_len), int(src_len)), "tkmi_generic_plan_create")
    plan = TransformPlan(h, kind, nent, None, table if table is not None else packed)
    plan.packed = packed
    plan.perm = tuple(int(x) for x in perm)
    plan.lens = (int(dst_len), int(src_len))
    return plan


def _make_abelian_plan(tab, perm, dst_len, src_len):
    return plan_from_packed("abelian", perm, pack_abelian(tab), dst_len, src_len, table=tab)


def _make_generic_plan(rows, perm, dst_len, src_len):
    return plan_from_packed("generic", perm, pack_generic(rows), dst_len, src_len, table=rows)


def treebraider(Vdst, Vsrc, p1, p2, levels=None, *, adjoint_src=False, dst_twist=()):
    """Cached device plan for `braid!/permute!` from `Vsrc` to `Vdst`
    (`treebraider`, `treetransformers.jl:159-167`).  With `adjoint_src`, the
    source is the lazy adjoint of a tensor with space `Vsrc.adjoint()` (its data
    is the parent's, sub-blocks permuted as in `src/tensors/adjoint.jl:44-48`)."""
    key = ("braid", Vdst, Vsrc, tuple(p1), tuple(p2), levels, adjoint_src, tuple(dst_twist))
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is not None:
        return plan
    if Vsrc.permute(p1, p2) != Vdst:
        from .spaces import SpaceMismatch
        raise SpaceMismatch("Incompatible spaces for permuting.")
    src_structure = adjoint_structure(Vsrc) if adjoint_src else None
    perm = _perm0(p1, p2)
    if Vsrc.unit.fusion_style == UNIQUE:
        tab = abelian_table(Vdst, Vsrc, p1, p2, src_structure, dst_twist)
        plan = _make_abelian_plan(tab, perm, Vdst.dim(), Vsrc.dim())
    else:
        lv1 = levels[0] if levels else None
        lv2 = levels[1] if levels else None
        rows = generic_table(Vdst, Vsrc, p1, p2, lv1, lv2, src_structure, dst_twist)
        plan = _make_generic_plan(rows, perm, Vdst.dim(), Vsrc.dim())
    _TRANSFORMER_CACHE.put(key, plan)
    return plan


def treetransposer(Vdst, Vsrc, p1, p2, *, adjoint_src=False):
    """Cached device plan for the planar `transpose!` from `Vsrc` to `Vdst` (`treetransposer`,
    `treetransformers.jl:169-180`): the sub-blocks move as in a permute, each scaled by the
    coefficient of `transpose((f₁, f₂), (p₁, p₂))`.  Multiplicity-free abelian sectors
    (UniqueFusion) of any braiding style."""
    key = ("transpose", Vdst, Vsrc, tuple(p1), tuple(p2), adjoint_src)
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is not None:
        return plan
    if Vsrc.permute(p1, p2) != Vdst:
        from .spaces import SpaceMismatch
        raise SpaceMismatch("Incompatible spaces for transposing.")
    if Vsrc.unit.fusion_style != UNIQUE:
        raise NotImplementedError("planar transpose of non-abelian sectors without a symmetric braiding")
    src_structure = adjoint_structure(Vsrc) if adjoint_src else None
    tab = abelian_table(Vdst, Vsrc, tuple(p1), tuple(p2), src_structure, move=transpose_pair)
    plan = _make_abelian_plan(tab, _perm0(p1, p2), Vdst.dim(), Vsrc.dim())
    _TRANSFORMER_CACHE.put(key, plan)
    return plan


def adjoint_structure(Vadj):
    """Sub-block structure of the lazy adjoint t′ of a tensor t with space
    `Vadj.adjoint()`: (f₁, f₂) of t′ is t's (f₂, f₁) sub-block viewed with the
    domain legs first (`subblock(::AdjointTensorMap)`, `src/tensors/adjoint.jl:44-48`)."""
    parent = Vadj.adjoint()
    ps = parent.subblockstructure()
    n1p = parent.numout
    out = {}
    for (f1, f2), (sz, st, off) in ps.items():
        out[(f2, f1)] = (tuple(sz[n1p:]) + tuple(sz[:n1p]), tuple(st[n1p:]) + tuple(st[:n1p]), off)
    return out


def flip_pair(f1, f2, inds, inv=False):
    """`flip((f₁, f₂), ind; inv)` (`braiding_manipulations.jl:388-436`): toggle the duality
    flag of every leg in `inds` (0-based, in order), accumulating per leg
    codomain leg a:  isdual → χₐ·θₐ, else 1        (inv: isdual → 1, else conj(χₐ·θₐ))
    domain leg a:    isdual → conj(χₐ), else θₐ    (inv: isdual → conj(θₐ), else χₐ)
    with χ the Frobenius–Schur phase and θ the twist (real for the sectors here)."""
    n1 = len(f1.uncoupled)
    d1, d2 = list(f1.isdual), list(f2.isdual)
    factor = 1.0
    for i in inds:
        if i < n1:
            a = f1.uncoupled[i]
            chi, th = a.frobenius_schur_phase(), a.twist()
            if not inv:
                factor *= chi * th if d1[i] else 1
            else:
                factor *= 1 if d1[i] else chi * th
            d1[i] = not d1[i]
        else:
            j = i - n1
            a = f2.uncoupled[j]
            chi, th = a.frobenius_schur_phase(), a.twist()
            if not inv:
                factor *= chi if d2[j] else th
            else:
                factor *= th if d2[j] else chi
            d2[j] = not d2[j]
    g1 = FusionTree(f1.uncoupled, f1.coupled, d1, f1.innerlines, f1.vertices)
    g2 = FusionTree(f2.uncoupled, f2.coupled, d2, f2.innerlines, f2.vertices)
    return (g1, g2), float(factor)


def flipper(Vdst, Vsrc, inds, inv=False):
    """Cached device plan of `flip(t, I; inv)` (`indexmanipulations.jl:21-29`): every
    tree pair keeps its sub-block (same sizes, identity leg order) at the flipped space's
    offset, scaled by its `flip_pair` coefficient — one entry per tree pair on the
    abelian transform kernel, for any fusion style (flip never mixes trees)."""
    key = ("flip", Vdst, Vsrc, tuple(inds), bool(inv))
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is not None:
        return plan
    fs, fd = Vsrc.subblockstructure(), Vdst.subblockstructure()
    coeffs, dst, src = [], [], []
    for (f1, f2), (sz, st, off) in fs.items():
        pair, c = flip_pair(f1, f2, inds, inv)
        szd, std, offd = fd[pair]
        if tuple(szd) != tuple(sz):
            raise AssertionError("flip: sub-block sizes differ")
        coeffs.append(c)
        dst.append(list(szd) + list(std) + [offd])
        src.append(list(sz) + list(st) + [off])
    n = Vsrc.numind
    tab = {"coeffs": np.array(coeffs, dtype=np.float64),
           "dst": np.array(dst, dtype=np.int64).reshape(-1, 2 * n + 1),
           "src": np.array(src, dtype=np.int64).reshape(-1, 2 * n + 1)}
    plan = _make_abelian_plan(tab, tuple(range(n)), Vdst.dim(), Vsrc.dim())
    _TRANSFORMER_CACHE.put(key, plan)
    return plan


def flat_plan(length):
    """Plan for an elementwise y = β·y + α·x over a flat vector (the identity
    transform used by add!/scale!)."""
    key = ("flat", int(length))
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is not None:
        return plan
    tab = {"coeffs": np.ones(1 if length else 0), "dst": np.array([[length, 1, 0]] if length else np.zeros((0, 3))),
           "src": np.array([[length, 1, 0]] if length else np.zeros((0, 3)))}
    plan = _make_abelian_plan(tab, (0,), length, length)
    _TRANSFORMER_CACHE.put(key, plan)
    return plan


# ---------------------------------------------------------------------------
# partial trace (trace_permute!) as a transform: every destination sub-block sums
# the traced diagonal slices of the source sub-blocks that reach it
# ---------------------------------------------------------------------------
def trace_table(Vdst, Vsrc, p1, p2, q1, q2, planar=False):
    """Rows of a trace transform, one per destination sub-block (in `fusiontrees(Vdst)`
    order): (coeffs, dst sizes, [dst (strides..., offset)], [source slice (strides...,
    offset)]).  Restates `_trace_permute!` (`src/tensors/tensoroperations.jl:245-300`):
    each source tree pair (or block column of the recoupling matrix) is permuted to
    ((p1, q1), (p2, q2)), split after the kept legs, and contributes when both traced
    remainders are the same tree, with the factor dim(coupled)/dim(first uncoupled) and
    the twists of the non-dual traced legs.  A contribution is the diagonal sum over the
    traced degeneracy indices, listed as one strided slice per index tuple, already in
    destination order."""
    fts_src = Vsrc.subblockstructure()
    fts_dst = Vdst.subblockstructure()
    contrib = _trace_contrib(Vsrc, p1, p2, q1, q2, planar)
    keep = tuple(p1) + tuple(p2)
    rows = []
    for dpair in Vdst.fusiontrees():
        szd, std, offd = fts_dst[dpair]
        coeffs, slices = [], []
        for c, spair in contrib.get(dpair, []):
            szs, sts, offs = fts_src[spair]
            if tuple(szs[x] for x in keep) != tuple(szd) or any(szs[a] != szs[b] for a, b in zip(q1, q2)):
                raise AssertionError("trace: sub-block sizes do not match")
            for sl in _diag_slices(szs, sts, offs, keep, q1, q2):
                slices.append(sl)
                coeffs.append(c)
        rows.append((np.array(coeffs, dtype=np.float64).reshape(1, len(coeffs)), tuple(szd),
                     [list(std) + [offd]], slices))
    return rows


def _diag_slices(szs, sts, offs, keep, q1, q2):
    """The traced-diagonal slices of one source sub-block: one strided view of the kept
    legs (in destination order) per tuple of traced degeneracy indices."""
    base = [sts[x] for x in keep]
    diag = [sts[a] + sts[b] for a, b in zip(q1, q2)]
    return [base + [offs + sum(tk * dk for tk, dk in zip(t, diag))]
            for t in itertools.product(*[range(szs[a]) for a in q1])]


def _trace_contrib(Vsrc, p1, p2, q1, q2, planar=False):
    """Destination tree pair → [(coefficient, source tree pair)] of a trace; with `planar`,
    the coefficients of the planar trace (`planar_trace`, `duality_manipulations.jl:575-629`,
    as `planartrace!` uses them for sectors without a symmetric braiding)."""
    n1, n2 = len(p1), len(p2)
    if planar:
        if Vsrc.unit.fusion_style != UNIQUE:
            raise NotImplementedError("planar trace of non-abelian sectors without a symmetric braiding")
        contrib = {}
        for spair in Vsrc.fusiontrees():
            for dpair, c in planar_trace_pair(spair[0], spair[1], p1, p2, q1, q2).items():
                contrib.setdefault(dpair, []).append((float(c), spair))
        return contrib
    r1, r2 = tuple(p1) + tuple(q1), tuple(p2) + tuple(q2)
    contrib = {}

    def add(coeff, g1full, g2full, spair):
        f1pp, g1 = split_tree(g1full, n1)
        f2pp, g2 = split_tree(g2full, n2)
        if g1 != g2:
            return
        c = coeff * g1.coupled.dim() / g1.uncoupled[0].dim()
        for i in range(1, len(g1.uncoupled)):
            if not g1.isdual[i]:
                c *= g1.uncoupled[i].twist()
        contrib.setdefault((f1pp, f2pp), []).append((float(c), spair))

    unit = Vsrc.unit
    if unit.fusion_style == UNIQUE:
        for spair in Vsrc.fusiontrees():
            (g1, g2), coeff = braid_pair(spair[0], spair[1], r1, r2)
            add(coeff, g1, g2, spair)
    else:
        for blk in Vsrc.fusionblocks():
            dst, U = permute_block(blk, r1, r2, unit)
            for i, spair in enumerate(blk.trees):
                for j, (g1, g2) in enumerate(dst.trees):
                    if U[j, i] != 0:
                        add(U[j, i], g1, g2, spair)
    return contrib


def trace_adjoint_table(Vdst, Vsrc, p1, p2, q1, q2):
    """Rows of the pullback of a trace (ΔA += α·T‡ΔC), one per source tree pair of `Vsrc`
    that feeds the trace: destinations are its traced-diagonal slices, sources are the
    `Vdst` sub-blocks it contributes to, and U[r, k] = coeff_k·dim(c_dst)/dim(c_src) for
    every r — the adjoint of the trace for the inner product Σ_c dim(c)·⟨x_c, y_c⟩ that
    TensorKit's AD rules use (`trace_permute_pullback_ΔA!`,
    src/pullbacks/tensoroperations.jl:50-64, computes the same map as a product with
    an identity).  Each destination slice lies in exactly one row, so a β = 1 launch
    accumulates without write conflicts."""
    fts_src = Vsrc.subblockstructure()
    fts_dst = Vdst.subblockstructure()
    keep = tuple(p1) + tuple(p2)
    by_src = {}
    for dpair, lst in _trace_contrib(Vsrc, p1, p2, q1, q2).items():
        for c, spair in lst:
            by_src.setdefault(spair, []).append((c, dpair))
    rows = []
    for spair in Vsrc.fusiontrees():
        lst = by_src.get(spair)
        if not lst:
            continue
        szs, sts, offs = fts_src[spair]
        szk = tuple(szs[x] for x in keep)
        dsts = _diag_slices(szs, sts, offs, keep, q1, q2)
        w_src = spair[0].coupled.dim()
        srcs, ucol = [], []
        for c, dpair in lst:
            szd, std, offd = fts_dst[dpair]
            if tuple(szd) != szk:
                raise AssertionError("trace: sub-block sizes do not match")
            srcs.append(list(std) + [offd])
            ucol.append(c * dpair[0].coupled.dim() / w_src)
        U = np.tile(np.array(ucol, dtype=np.float64), (len(dsts), 1))
        rows.append((U, szk, dsts, srcs))
    return rows


def pack_rows(rows):
    """Generic-plan arrays of rows (U as an R×C matrix, slice sizes, R destination slices,
    C source slices), identity permutation."""
    return {
        "rows": np.array([r[0].shape[0] for r in rows], dtype=np.int64),
        "cols": np.array([r[0].shape[1] for r in rows], dtype=np.int64),
        "U": np.concatenate([np.asarray(r[0], dtype=np.float64).reshape(-1, order="F") for r in rows]
                            + [np.zeros(0)]),
        "sz_dst": np.array([r[1] for r in rows], dtype=np.int64).reshape(-1),
        "sz_src": np.array([r[1] for r in rows], dtype=np.int64).reshape(-1),
        "so_dst": np.array([x for r in rows for x in r[2]], dtype=np.int64).reshape(-1),
        "so_src": np.array([x for r in rows for x in r[3]], dtype=np.int64).reshape(-1),
    }


def trace_adjointer(Vdst, Vsrc, p1, p2, q1, q2):
    """Cached device plan of `trace_adjoint_table` (source: a tensor in `Vdst`,
    destination: a tensor in `Vsrc`)."""
    key = ("trace_adj", Vdst, Vsrc, tuple(p1), tuple(p2), tuple(q1), tuple(q2))
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is None:
        rows = trace_adjoint_table(Vdst, Vsrc, p1, p2, q1, q2)
        plan = plan_from_packed("trace_adj", tuple(range(len(p1) + len(p2))), pack_rows(rows), Vsrc.dim(),
                                Vdst.dim(), table=rows)
        _TRANSFORMER_CACHE.put(key, plan)
    return plan


def pack_trace(rows):
    """Generic-plan arrays of trace rows (identity permutation, sources already in
    destination order, no adjoint weights)."""
    return {
        "rows": np.ones(len(rows), dtype=np.int64),
        "cols": np.array([r[0].shape[1] for r in rows], dtype=np.int64),
        "U": np.concatenate([r[0].reshape(-1) for r in rows] + [np.zeros(0)]),
        "sz_dst": np.array([r[1] for r in rows], dtype=np.int64).reshape(-1),
        "sz_src": np.array([r[1] for r in rows], dtype=np.int64).reshape(-1),
        "so_dst": np.array([x for r in rows for x in r[2]], dtype=np.int64).reshape(-1),
        "so_src": np.array([x for r in rows for x in r[3]], dtype=np.int64).reshape(-1),
    }


def _make_trace_plan(rows, n, dst_len, src_len):
    return plan_from_packed("trace", tuple(range(n)), pack_trace(rows), dst_len, src_len, table=rows)


def tracer(Vdst, Vsrc, p1, p2, q1, q2, planar=False):
    """Cached device plan for `trace_permute!` (with `planar`: `planartrace!`) from `Vsrc`
    to `Vdst`."""
    key = ("trace", Vdst, Vsrc, tuple(p1), tuple(p2), tuple(q1), tuple(q2), planar)
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is None:
        rows = trace_table(Vdst, Vsrc, p1, p2, q1, q2, planar)
        plan = _make_trace_plan(rows, len(p1) + len(p2), Vdst.dim(), Vsrc.dim())
        _TRANSFORMER_CACHE.put(key, plan)
    return plan
