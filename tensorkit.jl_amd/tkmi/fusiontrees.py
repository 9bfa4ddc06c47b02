"""Fusion trees and the coefficient-generating tree manipulations.

Host-side restatement of TensorKit's fusion-tree layer (L2), which produces
the coefficients that the device permute kernels apply:

- `FusionTree` data model             `src/fusiontrees/fusiontrees.jl:22-51`
- tree iterator order                 `src/fusiontrees/iterator.jl:98-228`
- `FusionTreeBlock`                   `src/fusiontrees/fusiontrees.jl:121-157`
- bendright / bendleft (B-moves)      `src/fusiontrees/duality_manipulations.jl:33-191`
- repartition                         `src/fusiontrees/duality_manipulations.jl:460-505`
- artin_braid (tree and block)        `src/fusiontrees/braiding_manipulations.jl:18-198`
- braid of a tree (symmetric)         `src/fusiontrees/braiding_manipulations.jl:217-251`
- braid of a pair / block (fsbraid)   `src/fusiontrees/braiding_manipulations.jl:283-340`
- linearizepermutation / permutation2swaps `src/fusiontrees/duality_manipulations.jl:571-582`,
  `src/auxiliary/auxiliary.jl:13-28`
- multi_Fmove / multi_Fmove_inv / multi_associator (UniqueFusion)
                                      `src/fusiontrees/basic_manipulations.jl:99-470`
- foldright / foldleft (A-moves), cycleclockwise / cycleanticlockwise, transpose (fstranspose),
  planar_trace / elementary_trace (tree pairs, UniqueFusion)
                                      `src/fusiontrees/duality_manipulations.jl:193-744`

Only multiplicity-free fusion (UniqueFusion, SimpleFusion) is supported:
all vertex labels are 1.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from .sectors import UNIQUE


class FusionTree:
    """Splitting tree: `uncoupled` sectors (before the Z isomorphisms marked
    by `isdual`) fused through `innerlines` to `coupled`."""

    __slots__ = ("uncoupled", "coupled", "isdual", "innerlines", "vertices", "_h")

    def __init__(self, uncoupled, coupled, isdual, innerlines=None, vertices=None):
        uncoupled = tuple(uncoupled)
        n = len(uncoupled)
        if innerlines is None:
            innerlines = _abelianinner(uncoupled + (coupled.dual(),))
        self.uncoupled = uncoupled
        self.coupled = coupled
        self.isdual = tuple(bool(x) for x in isdual)
        self.innerlines = tuple(innerlines)
        self.vertices = tuple(vertices) if vertices is not None else (1,) * max(0, n - 1)
        self._h = hash((self.uncoupled, self.coupled, self.isdual, self.innerlines))

    def __len__(self):
        return len(self.uncoupled)

    def __eq__(self, other):
        return (self.coupled == other.coupled and self.uncoupled == other.uncoupled
                and self.isdual == other.isdual and self.innerlines == other.innerlines)

    def __hash__(self):
        return self._h

    def __repr__(self):
        return (f"FusionTree({self.uncoupled}, {self.coupled}, {self.isdual}, "
                f"{self.innerlines})")


def _abelianinner(outer):
    """`_abelianinner` (`fusiontrees.jl:368-382`): inner lines of the unique
    abelian tree with outer legs `(uncoupled..., dual(coupled))`."""
    n = len(outer)
    if n <= 3:
        return ()
    c = outer[0].otimes(outer[1])[0]
    return (c,) + _abelianinner((c,) + tuple(outer[2:]))


# --------------------------------------------------------------------------
# iterator
# --------------------------------------------------------------------------
def _iter_raw(uncoupled_lists, coupled):
    """Yield (out, lines, vertices) in the reference's iteration order: the last
    uncoupled leg is the outermost loop, then the inner line into it, then
    (recursively) the front legs (`iterator.jl:98-228`)."""
    n = len(uncoupled_lists)
    if n == 0:
        if coupled.isunit():
            yield (), (), ()
        return
    if n == 1:
        if coupled in uncoupled_lists[0]:
            yield (coupled,), (), ()
        return
    if n == 2:
        for b in uncoupled_lists[1]:
            for a in uncoupled_lists[0]:
                nab = a.Nsymbol(b, coupled)
                for v in range(1, nab + 1):
                    yield (a, b), (), (v,)
        return
    front = uncoupled_lists[:-1]
    for b in uncoupled_lists[-1]:
        for a in coupled.otimes(b.dual()):
            nabc = a.Nsymbol(b, coupled)
            for v in range(1, nabc + 1):
                for rout, rlines, rverts in _iter_raw(front, a):
                    yield rout + (b,), rlines + (a,), rverts + (v,)


def fusiontrees(uncoupled_lists, coupled, isdual):
    """All fusion trees with uncoupled sectors drawn from `uncoupled_lists[i]`
    (ordered iterables), coupled sector `coupled` and duality flags `isdual`."""
    lists = tuple(tuple(l) for l in uncoupled_lists)
    return _fusiontrees_cached(lists, coupled, tuple(bool(x) for x in isdual))


@lru_cache(maxsize=1 << 16)
def _fusiontrees_cached(lists, coupled, isdual):
    out = []
    for unc, lines, verts in _iter_raw(lists, coupled):
        out.append(FusionTree(unc, coupled, isdual, lines, verts))
    return tuple(out)


def fuse_set(sectors, unit):
    """Set of all outcomes of ⊗(sectors...)."""
    if len(sectors) == 0:
        return {unit}
    cur = {sectors[0]}
    for s in sectors[1:]:
        nxt = set()
        for c in cur:
            nxt.update(c.otimes(s))
        cur = nxt
    return cur


class FusionTreeBlock:
    """All tree pairs with fixed uncoupled sectors (`fusiontrees.jl:121-157`).
    Tree order: coupled sector (sorted) outermost, then f₂, then f₁."""

    __slots__ = ("uncoupled", "isdual", "trees", "_index")

    def __init__(self, uncoupled, isdual, unit):
        u1, u2 = tuple(uncoupled[0]), tuple(uncoupled[1])
        d1, d2 = tuple(isdual[0]), tuple(isdual[1])
        self.uncoupled = (u1, u2)
        self.isdual = (d1, d2)
        n1, n2 = len(u1), len(u2)
        trees = []
        if n1 == 0 and n2 == 0:
            f = FusionTree((), unit, (), (), ())
            trees.append((f, f))
        else:
            if n1 == 0:
                cs = sorted(c for c in fuse_set(u2, unit) if c.isunit())
            elif n2 == 0:
                cs = sorted(c for c in fuse_set(u1, unit) if c.isunit())
            else:
                cs = sorted(fuse_set(u1, unit) & fuse_set(u2, unit))
            for c in cs:
                t1s = fusiontrees([(s,) for s in u1], c, d1)
                for f2 in fusiontrees([(s,) for s in u2], c, d2):
                    for f1 in t1s:
                        trees.append((f1, f2))
        self.trees = trees
        self._index = {t: i for i, t in enumerate(trees)}

    def __len__(self):
        return len(self.trees)

    def index(self, pair):
        return self._index[pair]


# --------------------------------------------------------------------------
# elementary moves on tree pairs (UniqueFusion) and blocks (SimpleFusion)
# --------------------------------------------------------------------------
def _bendright_treepair(f1, f2):
    """`_bendright_treepair` (`duality_manipulations.jl:33-54`)."""
    n1, n2 = len(f1), len(f2)
    c = f1.coupled
    if n1 == 1:
        a = f1.uncoupled[0].unit()
    elif n1 == 2:
        a = f1.uncoupled[0]
    else:
        a = f1.innerlines[-1]
    b = f1.uncoupled[n1 - 1]
    f1n = FusionTree(f1.uncoupled[:-1], a, f1.isdual[:-1],
                     f1.innerlines[:-1] if n1 > 2 else (),
                     f1.vertices[:-1] if n1 > 1 else ())
    f2n = FusionTree(f2.uncoupled + (b.dual(),), a, f2.isdual + (not f1.isdual[n1 - 1],),
                     f2.innerlines + (c,) if n2 > 1 else (),
                     f2.vertices + (1,) if n2 > 0 else ())
    return (a, b, c), (f1n, f2n)


def _bendright_coeff(f1, a, b, c):
    coeff = c.sqrtdim() / a.sqrtdim()
    if f1.isdual[len(f1) - 1]:
        coeff *= b.dual().frobenius_schur_phase()
    return coeff * a.Bsymbol(b, c)


def bendright_pair(f1, f2):
    """`bendright((f₁, f₂))` (`duality_manipulations.jl:56-68`)."""
    (a, b, c), (g1, g2) = _bendright_treepair(f1, f2)
    return (g1, g2), _bendright_coeff(f1, a, b, c)


def bendleft_pair(f1, f2):
    """`bendleft((f₁, f₂))` (`duality_manipulations.jl:140-145`): conj of the
    bendright of the swapped pair (all coefficients here are real)."""
    (g2, g1), coeff = bendright_pair(f2, f1)
    return (g1, g2), coeff


def repartition_pair(f1, f2, n):
    """`repartition` (`duality_manipulations.jl:460-505`) for a tree pair."""
    coeff = 1
    k = len(f1) - n
    while k > 0:
        (f1, f2), c = bendright_pair(f1, f2)
        coeff *= c
        k -= 1
    while k < 0:
        (f1, f2), c = bendleft_pair(f1, f2)
        coeff *= c
        k += 1
    return (f1, f2), coeff


def braid_tree_symmetric(f, p):
    """`braid(f::FusionTree, p, levels)` for UniqueFusion with symmetric
    braiding (`braiding_manipulations.jl:217-235`): product of R-symbols over
    the inversions of `p` (0-based here)."""
    n = len(f)
    coeff = 1
    u = f.uncoupled
    for i in range(n):
        for j in range(i):
            if p[j] > p[i]:
                a, b = u[p[j]], u[p[i]]
                coeff *= a.Rsymbol(b, a.otimes(b)[0])
    unc = tuple(u[k] for k in p)
    isd = tuple(f.isdual[k] for k in p)
    return FusionTree(unc, f.coupled, isd), coeff


def linearizepermutation(p1, p2, n1, n2):
    """`linearizepermutation` (`duality_manipulations.jl:571-582`), 0-based:
    domain legs are numbered in reverse after the codomain legs."""
    def lin(x):
        return n2 + 2 * n1 - 1 - x if x >= n1 else x
    return tuple(lin(x) for x in p1) + tuple(lin(x) for x in reversed(p2))


def permutation2swaps(perm):
    """`permutation2swaps` (`auxiliary.jl:13-28`), 0-based swap positions
    (swap s exchanges legs s and s+1)."""
    p = list(perm)
    n = len(p)
    swaps = []
    for k in range(n - 1):
        swaps.extend(range(p[k] - 1, k - 1, -1))
        for l in range(k + 1, n):
            if p[l] < p[k]:
                p[l] += 1
        p[k] = k
    return swaps


def braid_pair(f1, f2, p1, p2, levels1=None, levels2=None):
    """`fsbraid` pair branch (`braiding_manipulations.jl:304-312`): returns
    ((f1′, f2′), coeff) for UniqueFusion sectors.  0-based permutation."""
    n1, n2 = len(f1), len(f2)
    p = linearizepermutation(p1, p2, n1, n2)
    (f, f0), c1 = repartition_pair(f1, f2, n1 + n2)
    fp, c2 = braid_tree_symmetric(f, p)
    (g1, g2), c3 = repartition_pair(fp, f0, len(p1))
    return (g1, g2), c1 * c2 * c3


# ---------------- blocks ----------------
def _block_unit(block_unc):
    for side in block_unc:
        for s in side:
            return s.unit()
    raise ValueError("empty block has no sector type")


def bendright_block(src: FusionTreeBlock, unit):
    """`bendright(::FusionTreeBlock)` (`duality_manipulations.jl:69-115`)."""
    n1 = len(src.uncoupled[0])
    unc = (src.uncoupled[0][:-1], src.uncoupled[1] + (src.uncoupled[0][n1 - 1].dual(),))
    isd = (src.isdual[0][:-1], src.isdual[1] + (not src.isdual[0][n1 - 1],))
    dst = FusionTreeBlock(unc, isd, unit)
    U = np.zeros((len(dst), len(src)))
    for col, (f1, f2) in enumerate(src.trees):
        (a, b, c), (g1, g2) = _bendright_treepair(f1, f2)
        U[dst.index((g1, g2)), col] = _bendright_coeff(f1, a, b, c)
    return dst, U


def bendleft_block(src: FusionTreeBlock, unit):
    """`bendleft(::FusionTreeBlock)` (`duality_manipulations.jl:148-191`)."""
    n2 = len(src.uncoupled[1])
    unc = (src.uncoupled[0] + (src.uncoupled[1][n2 - 1].dual(),), src.uncoupled[1][:-1])
    isd = (src.isdual[0] + (not src.isdual[1][n2 - 1],), src.isdual[1][:-1])
    dst = FusionTreeBlock(unc, isd, unit)
    U = np.zeros((len(dst), len(src)))
    for col, (f1, f2) in enumerate(src.trees):
        (a, b, c), (g2, g1) = _bendright_treepair(f2, f1)
        U[dst.index((g1, g2)), col] = _bendright_coeff(f2, a, b, c)
    return dst, U


def repartition_block(src: FusionTreeBlock, n, unit):
    k = len(src.uncoupled[0]) - n
    U = np.eye(len(src))
    dst = src
    while k > 0:
        dst, Ut = bendright_block(dst, unit)
        U = Ut @ U
        k -= 1
    while k < 0:
        dst, Ut = bendleft_block(dst, unit)
        U = Ut @ U
        k += 1
    return dst, U


def artin_braid_block(src: FusionTreeBlock, i, unit, inv=False):
    """`artin_braid(::FusionTreeBlock{I,N,0}, i)` (`braiding_manipulations.jl:81-198`),
    multiplicity-free branch; `i` is 0-based (swaps legs i and i+1)."""
    unc = src.uncoupled[0]
    n = len(unc)
    a0, b0 = unc[i], unc[i + 1]
    unc2 = list(unc)
    unc2[i], unc2[i + 1] = b0, a0
    unc2 = tuple(unc2)
    coupled2 = unit
    isd = list(src.isdual[0])
    isd[i], isd[i + 1] = isd[i + 1], isd[i]
    isd = tuple(isd)
    dst = FusionTreeBlock((unc2, ()), (isd, ()), unit)
    U = np.zeros((len(dst), len(src)))
    if a0.isunit() or b0.isunit():
        for col, (f, f2) in enumerate(src.trees):
            inner = f.innerlines
            ext = (unc[0],) + inner + (coupled2,)
            inner2 = inner
            if i > 0:
                inner2 = list(inner)
                inner2[i - 1] = ext[i + 1] if a0.isunit() else ext[i - 1]
                inner2 = tuple(inner2)
            g = FusionTree(unc2, coupled2, isd, inner2, f.vertices)
            U[dst.index((g, f2)), col] = 1.0
        return dst, U
    for col, (f, f2) in enumerate(src.trees):
        inner = f.innerlines
        ext = (unc[0],) + inner + (coupled2,)
        if i == 0:
            c = inner[0] if n > 2 else coupled2
            R = b0.Rsymbol(a0, c) if inv else a0.Rsymbol(b0, c)
            g = FusionTree(unc2, coupled2, isd, inner, f.vertices)
            U[dst.index((g, f2)), col] = R
            continue
        b = unc[i]
        d = unc[i + 1]
        a = ext[i - 1]
        c = ext[i]
        e = ext[i + 1]
        cands = set(a.otimes(d)) & set(e.otimes(b.dual()))
        for cp in sorted(cands):
            if inv:
                coeff = d.Rsymbol(c, e) * d.Fsymbol(a, b, e, cp, c) * d.Rsymbol(a, cp)
            else:
                coeff = c.Rsymbol(d, e) * d.Fsymbol(a, b, e, cp, c) * a.Rsymbol(d, cp)
            if coeff == 0:
                continue
            inner2 = list(inner)
            inner2[i - 1] = cp
            g = FusionTree(unc2, coupled2, isd, tuple(inner2))
            U[dst.index((g, f2)), col] = coeff
    return dst, U


def braid_block(src: FusionTreeBlock, p1, p2, levels1, levels2, unit):
    """`fsbraid` block branch (`braiding_manipulations.jl:314-340`): returns
    (dst block, U) with U of shape (len(dst), len(src)).  0-based p."""
    n1, n2 = len(src.uncoupled[0]), len(src.uncoupled[1])
    p = linearizepermutation(p1, p2, n1, n2)
    levels = list(levels1) + list(reversed(levels2))
    dst, U = repartition_block(src, n1 + n2, unit)
    for s in permutation2swaps(p):
        inv = levels[s] > levels[s + 1]
        dst, Ut = artin_braid_block(dst, s, unit, inv=inv)
        U = Ut @ U
        levels[s], levels[s + 1] = levels[s + 1], levels[s]
    if len(p2) == 0:
        return dst, U
    dst, Ut = repartition_block(dst, len(p1), unit)
    return dst, Ut @ U


def permute_block(src, p1, p2, unit):
    n1, n2 = len(src.uncoupled[0]), len(src.uncoupled[1])
    return braid_block(src, p1, p2, tuple(range(n1)), tuple(range(n1, n1 + n2)), unit)


def permute_pair(f1, f2, p1, p2):
    return braid_pair(f1, f2, p1, p2)


def split_tree(f, M):
    """`split(f, M)` (`src/fusiontrees/basic_manipulations.jl:30-55`): the first M
    uncoupled sectors of f fused to the inner sector between legs M and M+1, and a
    second tree whose first uncoupled sector is that inner sector followed by the
    remaining legs, coupling to f.coupled.  0-based M, 0 ≤ M ≤ N."""
    N = len(f.uncoupled)
    if not 0 <= M <= N:
        raise ValueError(f"M should be between 0 and N = {N}")
    if N == 0:
        return f, FusionTree((f.coupled,), f.coupled, (False,), (), ())
    inner_ext = (f.uncoupled[0],) + tuple(f.innerlines) + (f.coupled,)
    vert_ext = (1,) + tuple(f.vertices)
    coupled1 = f.coupled.unit() if M == 0 else inner_ext[M - 1]
    f1 = FusionTree(f.uncoupled[:M], coupled1, f.isdual[:M], f.innerlines[:max(0, M - 2)],
                    f.vertices[:max(0, M - 1)])
    f2 = FusionTree((coupled1,) + f.uncoupled[M:], f.coupled, (False,) + f.isdual[M:],
                    inner_ext[M:M + max(0, N - M - 1)], vert_ext[M:N])
    return f1, f2


# --------------------------------------------------------------------------
# planar moves on tree pairs (UniqueFusion): A-moves, cyclic rotations, transpose and
# planar traces.  No braiding enters: the coefficients are products of F-, A- and
# B-symbols, quantum dimensions and Frobenius–Schur phases only.
# --------------------------------------------------------------------------
def _conj(x):
    return x.conjugate() if isinstance(x, complex) else x


def _vertex_info(f, k):
    """`vertex_info(f, k)` (`basic_manipulations.jl:103-111`), 1-based k ≥ 2: the sectors
    left and right of uncoupled leg k (the incoming line and the outgoing line of its
    vertex)."""
    N = len(f.uncoupled)
    left = f.uncoupled[0] if k == 2 else f.innerlines[k - 3]
    right = f.coupled if k == N else f.innerlines[k - 2]
    return left, right


def _multi_associator(long, short):
    """`multi_associator(long, short)` (`basic_manipulations.jl:137-167`), multiplicity-free:
    the product of F-symbols that recouples `long` into its first leg times `short`."""
    if long.uncoupled[1:] != short.uncoupled or long.isdual[1:] != short.isdual:
        return 0
    a = long.uncoupled[0]
    coeff = 1
    for k in range(2, len(long.uncoupled)):
        c = long.uncoupled[k]
        e, d = _vertex_info(long, k + 1)
        b, e2 = _vertex_info(short, k)
        coeff *= a.Fsymbol(b, c, d, e, e2)
    return coeff


def multi_Fmove_pair(f):
    """`multi_Fmove(f)` for UniqueFusion (`basic_manipulations.jl:191-211`): the one tree
    of the legs after the first, coupling to dual(a) ⊗ c, and its coefficient."""
    if not f.uncoupled:
        raise ValueError("multi_Fmove requires at least one uncoupled sector")
    coupled = f.uncoupled[0].dual().otimes(f.coupled)[0]
    g = FusionTree(f.uncoupled[1:], coupled, f.isdual[1:])
    return g, _multi_associator(f, g)


def multi_Fmove_inv_pair(a, c, f, isduala=False):
    """`multi_Fmove_inv(a, c, f, isduala)` for UniqueFusion (`basic_manipulations.jl:
    353-470`): a fused onto the left of f to coupled sector c, as the one standard tree,
    with the product of conjugated F-symbols."""
    b = f.coupled
    if not a.Nsymbol(b, c):
        raise ValueError(f"cannot fuse sectors {a} and {b} to {c}")
    N = len(f.uncoupled)
    g = FusionTree((a,) + f.uncoupled, c, (bool(isduala),) + f.isdual)
    if N <= 1:
        return g, 1
    coeff = 1
    for k in range(2, N + 1):
        cc = f.uncoupled[k - 1]
        bb, e2 = (f.uncoupled[0], _vertex_info(f, 2)[1]) if k == 2 else _vertex_info(f, k)
        e, d = _vertex_info(g, k + 1)
        coeff *= _conj(a.Fsymbol(bb, cc, d, e, e2))
    return g, coeff


def foldright_pair(f1, f2):
    """`foldright((f₁, f₂))` (`duality_manipulations.jl:217-232`): the first splitting
    vertex a ⊗ b ← c becomes the fusion vertex b ← dual(a) ⊗ c."""
    if not f1.uncoupled:
        raise ValueError("foldright needs a splitting leg")
    a = f1.uncoupled[0]
    isduala = f1.isdual[0]
    g1, c1 = multi_Fmove_pair(f1)
    b, c = g1.coupled, f1.coupled
    g2, c2 = multi_Fmove_inv_pair(a.dual(), b, f2, not isduala)
    coeff = c.sqrtdim() / b.sqrtdim() * c1 * a.Asymbol(b, c) * _conj(c2)
    if isduala:
        coeff *= a.frobenius_schur_phase()
    return (g1, g2), coeff


def foldleft_pair(f1, f2):
    """`foldleft((f₁, f₂))` (`duality_manipulations.jl:313-317`): conj of the foldright of
    the swapped pair."""
    (g2, g1), coeff = foldright_pair(f2, f1)
    return (g1, g2), _conj(coeff)


def cycleclockwise_pair(f1, f2):
    """`cycleclockwise` (`duality_manipulations.jl:401-410`): the last fusion leg bent to
    the splitting side, the first splitting leg folded to the fusion side."""
    if f1.uncoupled:
        (t1, t2), u1 = foldright_pair(f1, f2)
        (g1, g2), u2 = bendleft_pair(t1, t2)
    else:
        (t1, t2), u1 = bendleft_pair(f1, f2)
        (g1, g2), u2 = foldright_pair(t1, t2)
    return (g1, g2), u2 * u1


def cycleanticlockwise_pair(f1, f2):
    """`cycleanticlockwise` (`duality_manipulations.jl:430-440`)."""
    if f2.uncoupled:
        (t1, t2), u1 = foldleft_pair(f1, f2)
        (g1, g2), u2 = bendright_pair(t1, t2)
    else:
        (t1, t2), u1 = bendright_pair(f1, f2)
        (g1, g2), u2 = foldleft_pair(t1, t2)
    return (g1, g2), u2 * u1


def transpose_pair(f1, f2, p1, p2):
    """`transpose((f₁, f₂), (p₁, p₂))` = `fstranspose` (`duality_manipulations.jl:518-569`):
    repartition to len(p₁) splitting legs, then rotate the linearised legs one step at a
    time (anticlockwise while leg 1 sits in the first half, clockwise otherwise) until leg
    1 is where the cyclic permutation puts it.  0-based (p₁, p₂)."""
    N = len(p1) + len(p2)
    p = linearizepermutation(p1, p2, len(f1.uncoupled), len(f2.uncoupled))
    (g1, g2), U = repartition_pair(f1, f2, len(p1))
    if N == 0:
        return (g1, g2), U
    i1 = p.index(0) + 1
    if i1 == 1:
        return (g1, g2), U
    nhalf = N >> 1
    while 1 < i1 <= nhalf:
        (g1, g2), u = cycleanticlockwise_pair(g1, g2)
        U = u * U
        i1 -= 1
    while nhalf < i1:
        (g1, g2), u = cycleclockwise_pair(g1, g2)
        U = u * U
        i1 = i1 % N + 1
    return (g1, g2), U


def join_trees(f1, f2):
    """`join(f₁, f₂)` (`basic_manipulations.jl:57-97`) for UniqueFusion: f₂'s first leg is
    f₁'s coupled sector; the joined tree has f₁'s legs followed by f₂'s others."""
    if f2.uncoupled[0] != f1.coupled:
        raise ValueError("cannot connect the two fusion trees")
    return FusionTree(f1.uncoupled + f2.uncoupled[1:], f2.coupled, f1.isdual + f2.isdual[1:])


def elementary_trace_tree(f, i):
    """`elementary_trace(f, i)` (`duality_manipulations.jl:697-744`), 0-based i: the trace
    of legs i and (i+1) mod N of one splitting tree → {tree: coefficient}."""
    N = len(f.uncoupled)
    if not (N > 1 and 0 <= i < N):
        raise ValueError(f"cannot trace outputs {i} and {i + 1} out of only {N} outputs")
    if not (i < N - 1 or f.coupled.isunit()):
        raise ValueError("cannot trace the last and first output of a tree with a non-trivial coupled sector")
    j = (i + 1) % N
    b, b2 = f.uncoupled[i], f.uncoupled[j]
    if not (b == b2.dual() and f.isdual[i] != f.isdual[j]):
        return {}
    if i < N - 1:
        fleft, frem = split_tree(f, i)
        ftrace, fright = split_tree(frem, 3)
        a, d = ftrace.uncoupled[0], ftrace.coupled
        if a != d:
            return {}
        coeff = b.sqrtdim()
        if i > 0:
            coeff *= a.Fsymbol(b, b.dual(), a, ftrace.innerlines[0], a.unit())
        if ftrace.isdual[1]:
            coeff *= b.frobenius_schur_phase()
        return {join_trees(fleft, fright): coeff}
    fleft, _ = split_tree(f, N - 1)
    g, c0 = multi_Fmove_pair(fleft)
    if not g.coupled.isunit():
        return {}
    coeff = c0 * b.sqrtdim()
    if not f.isdual[N - 1]:
        coeff *= _conj(b.frobenius_schur_phase())
    return {g: coeff}


def planar_trace_tree(f, q1, q2):
    """`planar_trace(f, (q₁, q₂))` (`duality_manipulations.jl:631-688`), 0-based: trace the
    pairs (q₁[k], q₂[k]), innermost neighbouring pair first."""
    if not q1:
        return {f: 1}
    N = len(f.uncoupled)
    for i, j in zip(q1, q2):
        if not (f.uncoupled[i] == f.uncoupled[j].dual() and f.isdual[i] != f.isdual[j]):
            return {}
    k = None
    for kk in range(len(q1)):
        if (q1[kk] + 1) % N == q2[kk]:
            i, j, k = q1[kk], q2[kk], kk
            break
        if (q2[kk] + 1) % N == q1[kk]:
            i, j, k = q2[kk], q1[kk], kk
            break
    if k is None:
        raise ValueError(f"indices {q1} and {q2} do not form a valid planar trace on a fusion tree with {N} legs")
    q1r = tuple(l - (l > i) - (l > j) for n, l in enumerate(q1) if n != k)
    q2r = tuple(l - (l > i) - (l > j) for n, l in enumerate(q2) if n != k)
    out = {}
    for g, c1 in elementary_trace_tree(f, i).items():
        for h, c2 in planar_trace_tree(g, q1r, q2r).items():
            c = c1 * c2
            if c != 0:
                out[h] = out.get(h, 0) + c
    return out


def planar_trace_pair(f1, f2, p1, p2, q1, q2):
    """`planar_trace((f₁, f₂), (p₁, p₂), (q₁, q₂))` (`duality_manipulations.jl:575-629`),
    UniqueFusion branch, 0-based: repartition to one splitting tree, trace it planarly,
    transpose the rest to (p₁, p₂) → {(f₁′, f₂′): coefficient}."""
    n1, n2 = len(f1.uncoupled), len(f2.uncoupled)
    N = len(p1) + len(p2) + 2 * len(q1)
    if n1 + n2 != N:
        raise ValueError(f"fusion tree pair has {n1 + n2} indices, but the permutation expects {N}")
    if not q1:
        g, c = transpose_pair(f1, f2, p1, p2)
        return {g: c}

    def lin(x):
        return x if x < n1 else 2 * n1 + n2 - 1 - x

    q1l = tuple(lin(x) for x in q1)
    q2l = tuple(lin(x) for x in q2)
    qall = q1l + q2l
    p1l = tuple(lin(x) - sum(lin(x) > y for y in qall) for x in p1)
    p2l = tuple(lin(x) - sum(lin(x) > y for y in qall) for x in p2)
    (g1, g2), c1 = repartition_pair(f1, f2, N)
    out = {}
    for h, c2 in planar_trace_tree(g1, q1l, q2l).items():
        t, c3 = transpose_pair(h, g2, p1l, p2l)
        c = c1 * c2 * c3
        if c != 0:
            out[t] = out.get(t, 0) + c
    return out
