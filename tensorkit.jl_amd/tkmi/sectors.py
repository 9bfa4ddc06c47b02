"""Symmetry sectors: the label sets that grade TensorKit's vector spaces.

Mirrors the sector interface that TensorKit re-exports from TensorKitSectors
(`src/TensorKit.jl:43-47`; specification text in `docs/src/man/sectors.md:225-410`
for U1/SU2 and `:580-620` for fermion parity):
`unit, dual, otimes (⊗), Nsymbol, Fsymbol, Rsymbol, Bsymbol, twist, dim,
frobenius_schur_phase, isless`.

Sectors are immutable, hashable Python objects.  Ordering (`<`) implements
the reference's `isless`, which fixes every block and sub-block offset in a
tensor's flat data vector.

Half-integer labels (U1 charges, SU2 spins) are stored as twice their value
(`two_q`, `two_j`), like HalfIntegers.jl's `HalfInt`.
"""
from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

UNIQUE = "unique"        # FusionStyle UniqueFusion (abelian)
SIMPLE = "simple"        # FusionStyle SimpleFusion (multiplicity free, non-abelian)
BOSONIC = "bosonic"
FERMIONIC = "fermionic"
NOBRAIDING = "nobraiding"  # BraidingStyle NoBraiding: planar operations only


def _half(two):
    return Fraction(two, 2)


class Sector:
    """Base class.  Subclasses define `_key` (tuple used for ==/hash),
    `_order` (tuple used for isless) and the symbol methods."""

    __slots__ = ()
    fusion_style = UNIQUE
    braiding_style = BOSONIC
    scalartype = int

    # equality / ordering -------------------------------------------------
    def __eq__(self, other):
        return type(self) is type(other) and self._key == other._key

    def __ne__(self, other):
        return not self.__eq__(other)

    def __hash__(self):
        return hash((type(self).__name__, self._key))

    def __lt__(self, other):
        return self._order < other._order

    def __le__(self, other):
        return self == other or self < other

    # defaults --------------------------------------------------------------
    def isunit(self):
        return self == self.unit()

    def dim(self):
        return 1

    def sqrtdim(self):
        return math.sqrt(self.dim())

    def frobenius_schur_phase(self):
        return 1

    def twist(self):
        return 1

    # symbols: default definitions of TensorKitSectors for multiplicity-free fusion
    def Bsymbol(a, b, c):
        """B-move coefficient; TensorKitSectors' multiplicity-free default
        sqrt(da db / dc) * F(a, b, dual(b), a; c, unit)."""
        if not a.Nsymbol(b, c):
            return 0
        if a.fusion_style == UNIQUE and a.braiding_style == BOSONIC:
            return 1
        return (a.sqrtdim() * b.sqrtdim() / c.sqrtdim()) * a.Fsymbol(b, b.dual(), a, c, a.unit())

    def Asymbol(a, b, c):
        """A-move coefficient; multiplicity-free default
        sqrt(da db / dc) * conj(χ_a F(dual(a), a, b; b, unit, c))."""
        if not a.Nsymbol(b, c):
            return 0
        return (a.sqrtdim() * b.sqrtdim() / c.sqrtdim()) * (
            a.frobenius_schur_phase() * a.dual().Fsymbol(a, b, b, a.unit(), c))


class U1Irrep(Sector):
    """U(1) irrep with (half-)integer charge.  `docs/src/man/sectors.md:249-279`."""

    __slots__ = ("two_q",)

    def __init__(self, q=0, *, two_q=None):
        if two_q is None:
            tq = Fraction(q) * 2
            if tq.denominator != 1:
                raise ValueError(f"U1 charge {q} is not a half-integer")
            two_q = int(tq)
        object.__setattr__(self, "two_q", int(two_q))

    def __setattr__(self, k, v):
        raise AttributeError("sectors are immutable")

    @property
    def _key(self):
        return self.two_q

    @property
    def _order(self):
        # isless(c1, c2) = |q1| < |q2| || 0 < q1 == -q2  (sectors.md:262-264)
        return (abs(self.two_q), 1 if self.two_q < 0 else 0)

    @property
    def charge(self):
        return _half(self.two_q)

    @staticmethod
    def unit():
        return U1Irrep(two_q=0)

    def dual(self):
        return U1Irrep(two_q=-self.two_q)

    def otimes(self, b):
        return (U1Irrep(two_q=self.two_q + b.two_q),)

    def Nsymbol(a, b, c):
        return int(a.two_q + b.two_q == c.two_q)

    def Fsymbol(a, b, c, d, e, f):
        return int(a.Nsymbol(b, e) and e.Nsymbol(c, d) and b.Nsymbol(c, f) and a.Nsymbol(f, d))

    def Rsymbol(a, b, c):
        return int(a.Nsymbol(b, c))

    def Bsymbol(a, b, c):
        return int(a.Nsymbol(b, c))

    def Asymbol(a, b, c):
        return int(a.Nsymbol(b, c))

    def __repr__(self):
        q = self.charge
        return f"U1Irrep({q.numerator}/{q.denominator})" if q.denominator != 1 else f"U1Irrep({q.numerator})"


class PlanarTrivial(Sector):
    """The trivial sector without a braiding (TensorKit's `PlanarTrivial`, the sector of
    `ℙ^n` spaces, `src/spaces/planarspace.jl`): one sector, every symbol 1, and only
    planar operations (transpose, planar traces and contractions) are defined."""

    __slots__ = ()
    braiding_style = NOBRAIDING

    @property
    def _key(self):
        return ()

    @property
    def _order(self):
        return ()

    @staticmethod
    def unit():
        return PlanarTrivial()

    def dual(self):
        return self

    def otimes(self, b):
        return (PlanarTrivial(),)

    def Nsymbol(a, b, c):
        return 1

    def Fsymbol(a, b, c, d, e, f):
        return 1

    def Rsymbol(a, b, c):
        raise ValueError("PlanarTrivial sectors have no braiding")

    def Bsymbol(a, b, c):
        return 1

    def Asymbol(a, b, c):
        return 1

    def __repr__(self):
        return "PlanarTrivial()"


class FermionParity(Sector):
    """Fermion parity fZ2.  `docs/src/man/sectors.md:580-607`."""

    __slots__ = ("isodd",)
    braiding_style = FERMIONIC

    def __init__(self, isodd=False):
        object.__setattr__(self, "isodd", bool(isodd))

    def __setattr__(self, k, v):
        raise AttributeError("sectors are immutable")

    @property
    def _key(self):
        return self.isodd

    @property
    def _order(self):
        return (int(self.isodd),)

    @staticmethod
    def unit():
        return FermionParity(False)

    def dual(self):
        return self

    def otimes(self, b):
        return (FermionParity(self.isodd ^ b.isodd),)

    def Nsymbol(a, b, c):
        return int((a.isodd ^ b.isodd) == c.isodd)

    def Fsymbol(a, b, c, d, e, f):
        return int(a.Nsymbol(b, e) and e.Nsymbol(c, d) and b.Nsymbol(c, f) and a.Nsymbol(f, d))

    def Rsymbol(a, b, c):
        n = a.Nsymbol(b, c)
        return -n if (a.isodd and b.isodd) else n

    def twist(self):
        return -1 if self.isodd else 1

    def Bsymbol(a, b, c):
        return int(a.Nsymbol(b, c))

    def Asymbol(a, b, c):
        return int(a.Nsymbol(b, c))

    def __repr__(self):
        return f"FermionParity({int(self.isodd)})"


# --------------------------------------------------------------------------
# SU(2): Racah W / Wigner 6j and Clebsch-Gordan closed forms
# --------------------------------------------------------------------------
def _fact(n):
    return math.factorial(n)


def _triangle_ok(two_a, two_b, two_c):
    return (two_a + two_b + two_c) % 2 == 0 and abs(two_a - two_b) <= two_c <= two_a + two_b


@lru_cache(maxsize=None)
def _delta_sq(two_a, two_b, two_c):
    # Δ(abc)^2 = (a+b-c)!(a-b+c)!(-a+b+c)!/(a+b+c+1)! as an exact Fraction
    return Fraction(_fact((two_a + two_b - two_c) // 2) * _fact((two_a - two_b + two_c) // 2)
                    * _fact((-two_a + two_b + two_c) // 2), _fact((two_a + two_b + two_c) // 2 + 1))


@lru_cache(maxsize=None)
def racahW_sq_signed(two_a, two_b, two_c, two_d, two_e, two_f):
    """Racah's closed form for W(abcd;ef) (Racah 1942), returned as
    (sign, exact square) so callers can take one correctly rounded sqrt.
    W = Δ(abe)Δ(cde)Δ(acf)Δ(bdf) Σ_z (-1)^{z+a+b+c+d}(z+1)!/[(z-a-b-e)!(z-c-d-e)!
        (z-a-c-f)!(z-b-d-f)!(a+b+c+d-z)!(a+d+e+f-z)!(b+c+e+f-z)!]."""
    if not (_triangle_ok(two_a, two_b, two_e) and _triangle_ok(two_c, two_d, two_e)
            and _triangle_ok(two_a, two_c, two_f) and _triangle_ok(two_b, two_d, two_f)):
        return 0, Fraction(0)
    a1 = (two_a + two_b + two_e) // 2
    a2 = (two_c + two_d + two_e) // 2
    a3 = (two_a + two_c + two_f) // 2
    a4 = (two_b + two_d + two_f) // 2
    b1 = (two_a + two_b + two_c + two_d) // 2
    b2 = (two_a + two_d + two_e + two_f) // 2
    b3 = (two_b + two_c + two_e + two_f) // 2
    zmin = max(a1, a2, a3, a4)
    zmax = min(b1, b2, b3)
    s = 0
    for z in range(zmin, zmax + 1):
        term = _fact(z + 1) // (_fact(z - a1) * _fact(z - a2) * _fact(z - a3) * _fact(z - a4)
                                * _fact(b1 - z) * _fact(b2 - z) * _fact(b3 - z))
        # (-1)^(z + a + b + c + d) with a+b+c+d = b1 (integer since W nonzero)
        s += term if (z + b1) % 2 == 0 else -term
    if s == 0:
        return 0, Fraction(0)
    pref = (_delta_sq(two_a, two_b, two_e) * _delta_sq(two_c, two_d, two_e)
            * _delta_sq(two_a, two_c, two_f) * _delta_sq(two_b, two_d, two_f))
    return (1 if s > 0 else -1), pref * s * s


def _signed_sqrt(sign, sq: Fraction):
    if sign == 0:
        return 0.0
    # correctly rounded sqrt of a rational: scale to integers
    num, den = sq.numerator, sq.denominator
    # sqrt(num/den) = sqrt(num*den)/den, with 128 guard bits before one rounding
    v = math.isqrt((num * den) << 256)
    return sign * (v / (den << 128))


class SU2Irrep(Sector):
    """SU(2) irrep of spin j.  `docs/src/man/sectors.md:349-409`; F-symbol via
    Racah W exactly as TensorKitSectors (`Fsymbol = sqrtdim(e) sqrtdim(f)
    racahW(a, b, d, c, e, f)`), R = ±1 by parity of a+b-c."""

    __slots__ = ("two_j",)
    fusion_style = SIMPLE
    scalartype = float

    def __init__(self, j=0, *, two_j=None):
        if two_j is None:
            tj = Fraction(j) * 2
            if tj.denominator != 1 or tj < 0:
                raise ValueError(f"invalid SU2 spin {j}")
            two_j = int(tj)
        if two_j < 0:
            raise ValueError("Not a valid SU₂ irrep")
        object.__setattr__(self, "two_j", int(two_j))

    def __setattr__(self, k, v):
        raise AttributeError("sectors are immutable")

    @property
    def _key(self):
        return self.two_j

    @property
    def _order(self):
        return (self.two_j,)

    @property
    def j(self):
        return _half(self.two_j)

    @staticmethod
    def unit():
        return SU2Irrep(two_j=0)

    def dual(self):
        return self

    def dim(self):
        return self.two_j + 1

    def otimes(self, b):
        lo = abs(self.two_j - b.two_j)
        hi = self.two_j + b.two_j
        return tuple(SU2Irrep(two_j=t) for t in range(lo, hi + 1, 2))

    def Nsymbol(a, b, c):
        return int(_triangle_ok(a.two_j, b.two_j, c.two_j))

    def Fsymbol(s1, s2, s3, s4, s5, s6):
        return _su2_F(s1.two_j, s2.two_j, s3.two_j, s4.two_j, s5.two_j, s6.two_j)

    def Rsymbol(a, b, c):
        if not a.Nsymbol(b, c):
            return 0.0
        return 1.0 if ((a.two_j + b.two_j - c.two_j) // 2) % 2 == 0 else -1.0

    def frobenius_schur_phase(self):
        return 1 if self.two_j % 2 == 0 else -1

    def Bsymbol(a, b, c):
        return _su2_B(a.two_j, b.two_j, c.two_j)

    def Asymbol(a, b, c):
        return _su2_A(a.two_j, b.two_j, c.two_j)

    def __repr__(self):
        j = self.j
        return f"SU2Irrep({j.numerator}/{j.denominator})" if j.denominator != 1 else f"SU2Irrep({j.numerator})"


@lru_cache(maxsize=None)
def _su2_F_exact(t1, t2, t3, t4, t5, t6):
    """(sign, square) of F = sqrt(d5 d6) W(s1 s2 s4 s3; s5 s6)."""
    if all(t == 0 for t in (t1, t2, t3, t4, t5, t6)):
        return 1, Fraction(1)
    sg, sq = racahW_sq_signed(t1, t2, t4, t3, t5, t6)
    return sg, sq * (t5 + 1) * (t6 + 1)


@lru_cache(maxsize=None)
def _su2_F(t1, t2, t3, t4, t5, t6):
    sg, sq = _su2_F_exact(t1, t2, t3, t4, t5, t6)
    return _signed_sqrt(sg, sq)


@lru_cache(maxsize=None)
def _su2_B(ta, tb, tc):
    if not _triangle_ok(ta, tb, tc):
        return 0.0
    # sqrt(da db/dc) * F(a, b, b, a; c, 0), combined exactly before the sqrt
    sg, sq = _su2_F_exact(ta, tb, tb, ta, tc, 0)
    return _signed_sqrt(sg, sq * Fraction((ta + 1) * (tb + 1), tc + 1))


@lru_cache(maxsize=None)
def _su2_A(ta, tb, tc):
    if not _triangle_ok(ta, tb, tc):
        return 0.0
    sg, sq = _su2_F_exact(ta, ta, tb, tb, 0, tc)
    chi = 1 if ta % 2 == 0 else -1
    return _signed_sqrt(sg * chi, sq * Fraction((ta + 1) * (tb + 1), tc + 1))


class ProductSector(Sector):
    """Deligne product a ⊠ b of sectors (`docs/src/man/sectors.md:477-520`).
    Symbols are products of the factors' symbols.  Ordering compares the
    factor tuple lexicographically from the LAST factor (TensorKitSectors'
    `isless(reverse(p1.sectors), reverse(p2.sectors))`) -- this is not stated
    in the reference's own sources; see the product-order note in the tests."""

    __slots__ = ("sectors",)

    def __init__(self, *sectors):
        object.__setattr__(self, "sectors", tuple(sectors))

    def __setattr__(self, k, v):
        raise AttributeError("sectors are immutable")

    @property
    def fusion_style(self):
        styles = [s.fusion_style for s in self.sectors]
        return UNIQUE if all(s == UNIQUE for s in styles) else SIMPLE

    @property
    def braiding_style(self):
        return FERMIONIC if any(s.braiding_style == FERMIONIC for s in self.sectors) else BOSONIC

    @property
    def scalartype(self):
        return float if any(s.scalartype is float for s in self.sectors) else int

    @property
    def _key(self):
        return tuple((type(s).__name__, s._key) for s in self.sectors)

    @property
    def _order(self):
        return tuple(s._order for s in reversed(self.sectors))

    def unit(self):
        return ProductSector(*(s.unit() for s in self.sectors))

    def dual(self):
        return ProductSector(*(s.dual() for s in self.sectors))

    def dim(self):
        d = 1
        for s in self.sectors:
            d *= s.dim()
        return d

    def otimes(self, b):
        # iteration order: first factor fastest (Iterators.product)
        res = []
        for combo in _product_first_fastest([s.otimes(t) for s, t in zip(self.sectors, b.sectors)]):
            res.append(ProductSector(*combo))
        return tuple(res)

    def Nsymbol(a, b, c):
        n = 1
        for x, y, z in zip(a.sectors, b.sectors, c.sectors):
            n *= x.Nsymbol(y, z)
        return n

    def Fsymbol(a, b, c, d, e, f):
        v = 1
        for xs in zip(a.sectors, b.sectors, c.sectors, d.sectors, e.sectors, f.sectors):
            v *= xs[0].Fsymbol(*xs[1:])
        return v

    def Rsymbol(a, b, c):
        v = 1
        for x, y, z in zip(a.sectors, b.sectors, c.sectors):
            v *= x.Rsymbol(y, z)
        return v

    def Bsymbol(a, b, c):
        v = 1
        for x, y, z in zip(a.sectors, b.sectors, c.sectors):
            v *= x.Bsymbol(y, z)
        return v

    def Asymbol(a, b, c):
        v = 1
        for x, y, z in zip(a.sectors, b.sectors, c.sectors):
            v *= x.Asymbol(y, z)
        return v

    def twist(self):
        v = 1
        for s in self.sectors:
            v *= s.twist()
        return v

    def frobenius_schur_phase(self):
        v = 1
        for s in self.sectors:
            v *= s.frobenius_schur_phase()
        return v

    def __repr__(self):
        return " ⊠ ".join(repr(s) for s in self.sectors)


def _product_first_fastest(lists):
    """Iterators.product order: first factor varies fastest."""
    if not lists:
        yield ()
        return
    for rest in _product_first_fastest(lists[1:]):
        for x in lists[0]:
            yield (x,) + rest


def fZ2U1(p, q):
    """Convenience constructor for FermionParity ⊠ U1Irrep sectors (p, q)."""
    return ProductSector(FermionParity(p), U1Irrep(q))


def fuse_all(sectors, unit):
    """All coupled sectors of ⊗(s...) (sorted unique)."""
    cur = {unit} if not sectors else {sectors[0]}
    for s in sectors[1:]:
        nxt = set()
        for c in cur:
            nxt.update(c.otimes(s))
        cur = nxt
    return sorted(cur)
