"""Shared test fixtures: the benchmark spaces and conversions between the
product's space objects and the oracle's raw leg descriptions."""
import math

import numpy as np

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, ProductSpace, HomSpace, U1Space

from oracle import abelian as O


def julia_round(x):
    """Julia's `round` (ties to even)."""
    return int(round(x))


def cfg1_spaces():
    """V = U1Space(q => round(16 exp(-q²/4)), q ∈ -4:4), P = U1Space(-1,0,1 => 1)."""
    V = U1Space(*[(q, julia_round(16 * math.exp(-q * q / 4))) for q in range(-4, 5)])
    P = U1Space((-1, 1), (0, 1), (1, 1))
    return V, P


def cfg2_degeneracies(total=2048, mode="round"):
    """Gaussian d_q ∝ exp(-q²/16), q ∈ -8:8, normalised to `total`.
    mode='round': Julia round of the normalised weights (D = 2052);
    mode='lr': largest-remainder apportionment (D = 2048 exactly)."""
    qs = list(range(-8, 9))
    w = np.array([math.exp(-q * q / 16) for q in qs])
    x = total * w / w.sum()
    if mode == "round":
        d = [julia_round(v) for v in x]
    else:
        fl = np.floor(x).astype(int)
        rem = total - fl.sum()
        order = np.argsort(-(x - fl), kind="stable")
        fl[order[:rem]] += 1
        d = [int(v) for v in fl]
    return dict(zip(qs, d))


def cfg2_spaces(mode="round"):
    deg = cfg2_degeneracies(mode=mode)
    V = U1Space(*deg.items())
    P = U1Space((-0.5, 1), (0.5, 1))
    W = U1Space((0, 3), (1, 1), (-1, 1))
    return V, P, W


def to_leg(V):
    """GradedSpace -> oracle leg (dict of raw labels, isdual)."""
    d = {}
    for k, n in V._dims.items():
        d[_raw(k)] = n
    return (d, V.isdual)


def _raw(s):
    if isinstance(s, U1Irrep):
        return s.two_q
    if isinstance(s, ProductSector):
        p, u = s.sectors
        return (int(p.isodd), u.two_q)
    raise TypeError(s)


def kind_of(W):
    u = W.unit
    return O.FZ2U1Kind if isinstance(u, ProductSector) else O.U1Kind


def oracle_legs(W: HomSpace):
    return [to_leg(V) for V in W.codomain], [to_leg(V) for V in W.domain]


def random_u1_space(rng, maxq=2, maxdeg=4, halfint=False):
    qs = range(-maxq, maxq + 1)
    pairs = []
    for q in qs:
        d = int(rng.integers(0, maxdeg + 1))
        if d:
            pairs.append((q + (0.5 if halfint else 0), d))
    if not pairs:
        pairs = [(0.5 if halfint else 0, 1)]
    return U1Space(*pairs, dual=bool(rng.integers(0, 2)))


def random_perm(rng, n1, n2):
    p = list(rng.permutation(n1 + n2))
    k = int(rng.integers(0, n1 + n2 + 1))
    return tuple(int(x) for x in p[:k]), tuple(int(x) for x in p[k:])


def dense_heff(L, x, M1, M2, R):
    """Dense two-site H_eff application y[a s t; b] = L[a w A]·x[A S T B]·M1[w s S v]
    ·M2[v t T u]·R[B u b] as pairwise BLAS contractions (the same sums as
    np.einsum("awA,ASTB,wsSv,vtTu,Bub->astb"), without its unblocked loops)."""
    t1 = np.tensordot(L, x, axes=([2], [0]))            # a w S T B
    t2 = np.tensordot(t1, M1, axes=([1, 2], [0, 2]))    # a T B s v
    t3 = np.tensordot(t2, M2, axes=([1, 4], [2, 0]))    # a B s t u
    t4 = np.tensordot(t3, R, axes=([1, 4], [0, 1]))     # a s t b
    return t4
