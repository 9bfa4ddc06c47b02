"""Oracle (test infrastructure): SU(2) fusion tensors and dense embedding.

Independent restatement used to pin the recoupling matrices of the generic
(non-abelian) permute:
  clebschgordan  textbook closed form (Condon-Shortley phase), exact rationals
  fusiontensor   `fusiontensor(f::FusionTree)` / pair form, including the Z
                 isomorphism of dual legs (src/fusiontrees/fusiontrees.jl:278-341)
  layout         block / sub-block layout of an SU(2) HomSpace: coupled spins
                 ascending, trees ordered with the last uncoupled leg outermost,
                 then the inner line, then the front legs (src/fusiontrees/iterator.jl:98-228),
                 f₂ outer / f₁ inner within a block (src/spaces/structure.jl:41-188)
  to_dense       `convert(Array, t)` = Σ_trees kron(sub-block, fusion tensor)
                 placed at the sectors' axis ranges (src/tensors/abstracttensor.jl:733-756,
                 `_kron` src/auxiliary/auxiliary.jl:45-57: degeneracy index slow,
                 magnetic index fast)
The reference's own check is `convert(Array, permute(t, p)) ≈ permutedims(convert(Array, t), p)`
(test/tensors/indexmanipulations.jl:76-98).  Spins are stored as 2j.
"""
from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

import numpy as np


def _f(n):
    return math.factorial(n)


def _tri(a, b, c):
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


@lru_cache(maxsize=None)
def clebschgordan(t1, tm1, t2, tm2, tJ, tM):
    """<j1 m1 j2 m2 | J M> with all arguments doubled (2j, 2m)."""
    if tm1 + tm2 != tM or not _tri(t1, t2, tJ):
        return 0.0
    if abs(tm1) > t1 or abs(tm2) > t2 or abs(tM) > tJ:
        return 0.0
    j1p, j1m = (t1 + tm1) // 2, (t1 - tm1) // 2
    j2p, j2m = (t2 + tm2) // 2, (t2 - tm2) // 2
    Jp, Jm = (tJ + tM) // 2, (tJ - tM) // 2
    a = (tJ + t1 - t2) // 2
    b = (tJ - t1 + t2) // 2
    c = (t1 + t2 - tJ) // 2
    s = (t1 + t2 + tJ) // 2 + 1
    pref = Fraction((tJ + 1) * _f(a) * _f(b) * _f(c), _f(s)) * (_f(Jp) * _f(Jm) * _f(j1m) * _f(j1p) * _f(j2m) * _f(j2p))
    tot = Fraction(0)
    for k in range(0, c + 1):
        d = [k, c - k, j1m - k, j2p - k, (tJ - t2 + tm1) // 2 + k, (tJ - t1 - tm2) // 2 + k]
        if min(d) < 0:
            continue
        den = 1
        for x in d:
            den *= _f(x)
        tot += Fraction((-1) ** k, den)
    val = tot * tot * pref
    sgn = 1 if tot > 0 else (-1 if tot < 0 else 0)
    return sgn * math.sqrt(float(val)) if sgn else 0.0


def mvals(tj):
    """Magnetic quantum numbers (doubled) in basis order m = j, j-1, ..., -j."""
    return [tj - 2 * k for k in range(tj + 1)]


@lru_cache(maxsize=None)
def cg_tensor(ta, tb, tc):
    X = np.zeros((ta + 1, tb + 1, tc + 1))
    for i, ma in enumerate(mvals(ta)):
        for j, mb in enumerate(mvals(tb)):
            for k, mc in enumerate(mvals(tc)):
                X[i, j, k] = clebschgordan(ta, ma, tb, mb, tc, mc)
    return X


def Z(tc):
    """Z isomorphism tensor of a dual leg: sqrt(d_c)·conj(CG(c̄, c → 0))."""
    return math.sqrt(tc + 1) * cg_tensor(tc, tc, 0)[:, :, 0]


def tree_tensor(unc, inner, coupled, isdual):
    """Splitting tensor of a fusion tree, shape (d(u1), ..., d(uN), d(c))."""
    n = len(unc)
    if n == 0:
        return np.ones((1,))
    if n == 1:
        return Z(coupled) if isdual[0] else np.eye(coupled + 1)
    if n == 2:
        X = cg_tensor(unc[0], unc[1], coupled)
        if isdual[0]:
            X = np.einsum("xa,abc->xbc", Z(unc[0]), X)
        if isdual[1]:
            X = np.einsum("yb,abc->ayc", Z(unc[1]), X)
        return X
    C1 = tree_tensor(unc[:2], (), inner[0], isdual[:2])
    Ct = tree_tensor((inner[0],) + tuple(unc[2:]), tuple(inner[1:]), coupled, (False,) + tuple(isdual[2:]))
    # contract C1's coupled index with Ct's first index
    return np.tensordot(C1, Ct, axes=([2], [0]))


def pair_tensor(t1, t2):
    F1 = tree_tensor(*t1)
    F2 = tree_tensor(*t2)
    d1, d2 = F1.shape[:-1], F2.shape[:-1]
    return (F1.reshape(-1, F1.shape[-1]) @ F2.reshape(-1, F2.shape[-1]).T).reshape(d1 + d2)


# ----------------------------------------------------------------------------
# layout
# ----------------------------------------------------------------------------
def leg_sectors(leg):
    degs, _ = leg
    return sorted(k for k, d in degs.items() if d > 0)


def trees(lists, c):
    """(uncoupled, inner) for all trees of the uncoupled lists fusing to c."""
    n = len(lists)
    out = []
    if n == 0:
        return [((), ())] if c == 0 else []
    if n == 1:
        return [((c,), ())] if c in lists[0] else []
    if n == 2:
        for b in lists[1]:
            for a in lists[0]:
                if _tri(a, b, c):
                    out.append(((a, b), ()))
        return out
    for b in lists[-1]:
        for a in range(abs(c - b), c + b + 1, 2):
            for u, inn in trees(lists[:-1], a):
                out.append((u + (b,), inn + (a,)))
    return out


def layout(cod, dom):
    cl = [leg_sectors(l) for l in cod]
    dl = [leg_sectors(l) for l in dom]
    # coupled spins reachable on both sides
    def reach(lists):
        cur = {0}
        for L in lists:
            nxt = set()
            for x in cur:
                for y in L:
                    nxt.update(range(abs(x - y), x + y + 1, 2))
            cur = nxt
        return cur
    cs = sorted(reach(cl) & reach(dl))
    blocks, subs = [], []
    off = 0
    for c in cs:
        t1s = trees(cl, c)
        t2s = trees(dl, c)
        if not t1s or not t2s:
            continue
        d1s = [[cod[i][0][u] for i, u in enumerate(t[0])] for t in t1s]
        d2s = [[dom[i][0][u] for i, u in enumerate(t[0])] for t in t2s]
        bd1 = sum(int(np.prod(d)) for d in d1s)
        o2 = 0
        for t2, e in zip(t2s, d2s):
            o1 = 0
            for t1, d in zip(t1s, d1s):
                st = list(np.cumprod([1] + d[:-1])) if d else []
                st += list(bd1 * np.cumprod([1] + e[:-1])) if e else []
                subs.append(((t1, t2, c), tuple(d) + tuple(e), tuple(int(x) for x in st), off + o2 * bd1 + o1))
                o1 += int(np.prod(d))
            o2 += int(np.prod(e))
        blocks.append((c, bd1, o2, off))
        off += bd1 * o2
    return off, blocks, subs


def leg_axes(leg):
    out, o = {}, 0
    for s in leg_sectors(leg):
        n = leg[0][s] * (s + 1)
        out[s] = (o, n)
        o += n
    return out, o


def to_dense(data, cod, dom):
    total, blocks, subs = layout(cod, dom)
    assert data.size == total
    legs = list(cod) + list(dom)
    axes = [leg_axes(l) for l in legs]
    A = np.zeros(tuple(a[1] for a in axes))
    for (t1, t2, c), sz, st, off in subs:
        blk = np.lib.stride_tricks.as_strided(data[off:], shape=sz, strides=[s * 8 for s in st])
        F = pair_tensor((t1[0], t1[1], c, tuple(l[1] for l in cod)), (t2[0], t2[1], c, tuple(l[1] for l in dom)))
        kr = np.multiply.outer(blk, F)  # (deg..., m...)
        n = len(sz)
        perm = [x for i in range(n) for x in (i, n + i)]
        kr = kr.transpose(perm).reshape(tuple(sz[i] * F.shape[i] for i in range(n)))
        secs = list(t1[0]) + list(t2[0])
        sl = tuple(slice(ax[0][s][0], ax[0][s][0] + ax[0][s][1]) for ax, s in zip(axes, secs))
        A[sl] += kr
    return A


def permuted_legs(cod, dom, p1, p2):
    n1 = len(cod)

    def W(n):
        return cod[n] if n < n1 else (dom[n - n1][0], not dom[n - n1][1])

    return [W(n) for n in p1], [(W(n)[0], not W(n)[1]) for n in p2]
