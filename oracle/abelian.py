"""Oracle (test infrastructure): abelian block-sparse layout and dense embedding.

Independent restatement, written from the reference sources directly:
  sector order          U1 `isless` (docs/src/man/sectors.md:262-264); product
                        sectors compared on the factor tuple from the last factor
  space sectors         sorted keys, mapped through `dual` for dual spaces
                        (src/spaces/gradedspace.jl:179-186)
  tree order            uncoupled tuples with the first leg fastest
                        (src/fusiontrees/iterator.jl:98-228, abelian case), trees of
                        one block ordered f₂ outer, f₁ inner (src/spaces/structure.jl:41-72)
  block layout          column-major (d₁ × d₂) blocks in sorted coupled-sector order,
                        sub-block offset = block offset + off₂·d₁ + off₁
                        (src/spaces/structure.jl:114-188)
  dense embedding       `convert(Array, t)` is a placement for abelian sectors
                        (src/tensors/abstracttensor.jl:733-756; fusion tensors are 1)
  fermionic coefficient Koszul sign of the linearised permutation
                        (src/fusiontrees/braiding_manipulations.jl:217-235,304-312 with
                        trivial bends for FermionParity)

Sectors are plain Python values: U1 → int (twice the charge); fZ2⊠U1 → tuple
(parity, twice the charge).  A leg is (dict sector→degeneracy, isdual).
"""
from __future__ import annotations

import itertools

import numpy as np


class U1Kind:
    unit = 0

    @staticmethod
    def key(s):
        return (abs(s), 1 if s < 0 else 0)

    @staticmethod
    def dual(s):
        return -s

    @staticmethod
    def fuse(a, b):
        return a + b

    @staticmethod
    def odd(s):
        return False


class FZ2U1Kind:
    unit = (0, 0)

    @staticmethod
    def key(s):
        p, tq = s
        return ((abs(tq), 1 if tq < 0 else 0), p)

    @staticmethod
    def dual(s):
        return (s[0], -s[1])

    @staticmethod
    def fuse(a, b):
        return (a[0] ^ b[0], a[1] + b[1])

    @staticmethod
    def odd(s):
        return s[0] == 1


def leg_sectors(leg, kind):
    degs, isdual = leg
    keys = sorted((k for k, d in degs.items() if d > 0), key=kind.key)
    return [kind.dual(k) if isdual else k for k in keys]


def leg_deg(leg, s, kind):
    degs, isdual = leg
    return degs.get(kind.dual(s) if isdual else s, 0)


def leg_dual(leg):
    return (leg[0], not leg[1])


def leg_axes(leg, kind):
    out, off = {}, 0
    for s in leg_sectors(leg, kind):
        d = leg_deg(leg, s, kind)
        out[s] = (off, d)
        off += d
    return out, off


def _fuse_all(secs, kind):
    c = kind.unit
    for s in secs:
        c = kind.fuse(c, s)
    return c


def sector_tuples(legs, kind):
    """All sector tuples, first leg fastest."""
    lists = [leg_sectors(l, kind) for l in legs]
    return [tuple(reversed(c)) for c in itertools.product(*reversed(lists))]


def layout(cod, dom, kind):
    """Returns (total, blocks, subs): blocks = [(c, d1, d2, offset)],
    subs = {(cod_secs, dom_secs): (sizes, strides, offset)} in layout order."""
    ct = sector_tuples(cod, kind)
    dt = sector_tuples(dom, kind)
    ccs = {_fuse_all(x, kind) for x in ct}
    dcs = {_fuse_all(x, kind) for x in dt}
    blocks_c = sorted(ccs & dcs, key=kind.key)
    blocks, subs = [], {}
    offset = 0
    for c in blocks_c:
        f1s = [x for x in ct if _fuse_all(x, kind) == c]
        f2s = [x for x in dt if _fuse_all(x, kind) == c]
        d1s = [[leg_deg(l, s, kind) for l, s in zip(cod, x)] for x in f1s]
        d2s = [[leg_deg(l, s, kind) for l, s in zip(dom, x)] for x in f2s]
        bd1 = sum(int(np.prod(d)) for d in d1s)
        bd2 = sum(int(np.prod(d)) for d in d2s)
        o2 = 0
        for x2, e in zip(f2s, d2s):
            o1 = 0
            for x1, d in zip(f1s, d1s):
                st = list(np.cumprod([1] + d[:-1])) if d else []
                st += list(bd1 * np.cumprod([1] + e[:-1])) if e else []
                subs[(x1, x2)] = (tuple(d) + tuple(e), tuple(int(v) for v in st), offset + o2 * bd1 + o1)
                o1 += int(np.prod(d))
            o2 += int(np.prod(e))
        blocks.append((c, bd1, bd2, offset))
        offset += bd1 * bd2
    return offset, blocks, subs


def _view(data, sz, st, off):
    return np.lib.stride_tricks.as_strided(data[off:], shape=sz, strides=[s * data.itemsize for s in st],
                                           writeable=False)


def to_dense(data, cod, dom, kind):
    total, blocks, subs = layout(cod, dom, kind)
    assert data.size == total
    legs = list(cod) + list(dom)
    axes = [leg_axes(l, kind) for l in legs]
    dense = np.zeros(tuple(a[1] for a in axes), dtype=data.dtype)
    for (x1, x2), (sz, st, off) in subs.items():
        secs = x1 + x2
        sl = tuple(slice(ax[0][s][0], ax[0][s][0] + ax[0][s][1]) for ax, s in zip(axes, secs))
        dense[sl] = _view(data, sz, st, off)
    return dense


def from_dense(dense, cod, dom, kind):
    total, blocks, subs = layout(cod, dom, kind)
    legs = list(cod) + list(dom)
    axes = [leg_axes(l, kind) for l in legs]
    out = np.zeros(total, dtype=dense.dtype)
    for (x1, x2), (sz, st, off) in subs.items():
        secs = x1 + x2
        sl = tuple(slice(ax[0][s][0], ax[0][s][0] + ax[0][s][1]) for ax, s in zip(axes, secs))
        v = np.lib.stride_tricks.as_strided(out[off:], shape=sz, strides=[s * 8 for s in st])
        v[...] = dense[sl]
    return out


def permuted_legs(cod, dom, p1, p2):
    """Legs of `permute(W, (p1, p2))` (src/spaces/homspace.jl:212-244)."""
    n1 = len(cod)

    def W(n):
        return cod[n] if n < n1 else leg_dual(dom[n - n1])

    return [W(n) for n in p1], [leg_dual(W(n)) for n in p2]


def linearize(p1, p2, n1, n2):
    f = (lambda x: n2 + 2 * n1 - 1 - x if x >= n1 else x)
    return [f(x) for x in p1] + [f(x) for x in reversed(p2)]


def koszul_sign(secs, p1, p2, n1, n2, kind):
    """(-1)^(number of crossings of two odd legs) in the linearised picture."""
    lin = linearize(p1, p2, n1, n2)
    n = n1 + n2

    def leg_of(k):  # old linear position -> leg index
        return k if k < n1 else n2 + 2 * n1 - 1 - k

    odd = [kind.odd(secs[leg_of(k)]) for k in range(n)]
    s = 1
    for i in range(n):
        for j in range(i + 1, n):
            if lin[i] > lin[j] and odd[lin[i]] and odd[lin[j]]:
                s = -s
    return s


def permute(data, cod, dom, p1, p2, kind, alpha=1.0):
    """Oracle `permute(t, (p1, p2))`: dense transpose (bosonic) with the Koszul
    sign per sub-block (fermionic).  Returns (data, cod', dom')."""
    ncod, ndom = permuted_legs(cod, dom, p1, p2)
    n1, n2 = len(cod), len(dom)
    p = list(p1) + list(p2)
    if not any(kind.odd(s) for l in list(cod) + list(dom) for s in leg_sectors(l, kind)):
        dense = to_dense(data, cod, dom, kind)
        out = from_dense(np.ascontiguousarray(dense.transpose(p)), ncod, ndom, kind)
        return alpha * out if alpha != 1.0 else out, ncod, ndom
    # fermionic: move each sub-block with its sign
    _, _, subs = layout(cod, dom, kind)
    total2, _, subs2 = layout(ncod, ndom, kind)
    out = np.zeros(total2, dtype=data.dtype)
    for (x1, x2), (sz, st, off) in subs.items():
        secs = x1 + x2
        # sector of each leg as seen from W[n]: domain legs appear dualised
        wsec = list(x1) + [kind.dual(s) for s in x2]
        y1 = tuple(wsec[n] for n in p1)
        y2 = tuple(kind.dual(wsec[n]) for n in p2)
        sgn = koszul_sign(secs, p1, p2, n1, n2, kind)
        sz2, st2, off2 = subs2[(y1, y2)]
        v = np.lib.stride_tricks.as_strided(out[off2:], shape=sz2, strides=[s * 8 for s in st2])
        v[...] = (alpha * sgn) * _view(data, sz, st, off).transpose(p)
    return out, ncod, ndom


def transpose(data, cod, dom, p1, p2, kind, alpha=1.0):
    """Oracle planar `transpose(t, (p1, p2))` for U1 and fZ2⊠U1.  A planar transpose is a
    product of bends and folds (src/fusiontrees/duality_manipulations.jl:193-569), whose
    coefficients are built from F-, A- and B-symbols, quantum dimensions and
    Frobenius–Schur phases only — no R-symbol, so no Koszul sign.  For these sectors every
    one of them is 1 (dimensions 1, F = A = B = 1, χ = 1), so each sub-block moves with
    coefficient 1 and the result is the dense permutation of axes, for either braiding.
    The permutation must be cyclic in the linearised leg order."""
    n1, n2 = len(cod), len(dom)
    lin = linearize(p1, p2, n1, n2)
    n = n1 + n2
    if n and any(lin[(i + 1) % n] != (lin[i] + 1) % n for i in range(n)):
        raise ValueError(f"{(p1, p2)} is not a cyclic permutation")
    ncod, ndom = permuted_legs(cod, dom, p1, p2)
    dense = to_dense(data, cod, dom, kind)
    out = from_dense(np.ascontiguousarray(dense.transpose(list(p1) + list(p2))), ncod, ndom, kind)
    return alpha * out if alpha != 1.0 else out, ncod, ndom


def blocks(data, cod, dom, kind):
    total, bl, _ = layout(cod, dom, kind)
    return {c: data[o:o + d1 * d2].reshape((d1, d2), order="F") for c, d1, d2, o in bl}


def mul(A, Acod, Adom, B, Bcod, Bdom, kind, transA=False, transB=False):
    """Oracle `mul!`: per common coupled sector C_c = A_c · B_c (numpy/BLAS)."""
    ba = blocks(A, Acod, Adom, kind)
    bb = blocks(B, Bcod, Bdom, kind)
    ccod = Adom if transA else Acod
    cdom = Bcod if transB else Bdom
    total, bl, _ = layout(ccod, cdom, kind)
    out = np.zeros(total)
    for c, d1, d2, o in bl:
        if c in ba and c in bb:
            a = ba[c].T if transA else ba[c]
            b = bb[c].T if transB else bb[c]
            out[o:o + d1 * d2] = (a @ b).reshape(-1, order="F")
    return out
