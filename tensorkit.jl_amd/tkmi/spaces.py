"""Graded vector spaces, tensor-product spaces, HomSpaces and their block layout.

Host-side restatement of TensorKit's L1 layer:

- `GradedSpace` (sectors stored sorted; dual spaces map keys through `dual`)
  `src/spaces/gradedspace.jl:26-101,179-186`
- `ProductSpace` sectors / blocksectors / fusiontrees
  `src/spaces/productspace.jl:95-211`
- `HomSpace` indexing, `permute`/`select`, `compose`, `blocksectors`,
  `fusionblocks` `src/spaces/homspace.jl:94-271`
- `sectorstructure` / `degeneracystructure` / `_subblock_strides`
  `src/spaces/structure.jl:41-196`

The layout (block order, block ranges, sub-block sizes/strides/offsets) is
what every device kernel addresses; it must match the reference bit for bit.
"""
from __future__ import annotations

from collections import OrderedDict
from functools import lru_cache

from .fusiontrees import FusionTreeBlock, fusiontrees, fuse_set
from .sectors import U1Irrep, SU2Irrep, ProductSector


class SpaceMismatch(ValueError):
    """Mirror of TensorKit's `SpaceMismatch` (`src/TensorKit.jl:170-197`)."""


class SectorMismatch(ValueError):
    """Mirror of TensorKit's `SectorMismatch`."""


class IndexError_(IndexError):
    """Mirror of TensorKit's `IndexError`."""


class GradedSpace:
    """⊕_c ℂ^{n_c} ⊗ R_c.  `dims` maps (non-dual) sector keys to degeneracies;
    zero degeneracies are dropped, like the `SectorDict` constructor."""

    __slots__ = ("_dims", "isdual", "_h", "_keys")

    def __init__(self, dims, dual=False):
        items = dims.items() if isinstance(dims, dict) else dims
        d = {}
        for c, n in items:
            if c in d:
                raise ValueError(f"Sector {c} appears multiple times")
            if n < 0:
                raise ValueError(f"Sector {c} has negative dimension {n}")
            if n != 0:
                d[c] = int(n)
        self._keys = tuple(sorted(d))
        self._dims = {k: d[k] for k in self._keys}
        self.isdual = bool(dual)
        self._h = hash((tuple(self._dims.items()), self.isdual))

    @property
    def sectortype(self):
        return type(self._keys[0]) if self._keys else None

    def sectors(self):
        """Sectors in storage order (sorted keys), mapped through `dual` for a
        dual space (`gradedspace.jl:179-182`)."""
        if self.isdual:
            return tuple(k.dual() for k in self._keys)
        return self._keys

    def dim(self, c=None):
        if c is None:
            return sum(k.dim() * n for k, n in self._dims.items())
        key = c.dual() if self.isdual else c
        return self._dims.get(key, 0)

    def hassector(self, c):
        return self.dim(c) != 0

    def dual(self):
        return GradedSpace(self._dims, not self.isdual)

    conj = dual

    def flip(self):
        """`flip(V)` (`gradedspace.jl:95-101`)."""
        if self.isdual:
            return GradedSpace({c: self.dim(c) for c in self.sectors()})
        return GradedSpace({c.dual(): self.dim(c) for c in self.sectors()}).dual()

    def __eq__(self, other):
        return isinstance(other, GradedSpace) and self.isdual == other.isdual and self._dims == other._dims

    def __hash__(self):
        return self._h

    def __repr__(self):
        body = ", ".join(f"{c!r}=>{n}" for c, n in self._dims.items())
        return f"Vect[{body}]" + ("'" if self.isdual else "")

    # tensor product sugar
    def __mul__(self, other):
        return ProductSpace(self) * other

    def otimes(self, other):
        return ProductSpace(self) * other


def U1Space(*pairs, dual=False):
    if len(pairs) == 1 and isinstance(pairs[0], dict):
        pairs = tuple(pairs[0].items())
    return GradedSpace({U1Irrep(q) if not isinstance(q, U1Irrep) else q: n for q, n in pairs}, dual)


def SU2Space(*pairs, dual=False):
    if len(pairs) == 1 and isinstance(pairs[0], dict):
        pairs = tuple(pairs[0].items())
    return GradedSpace({SU2Irrep(j) if not isinstance(j, SU2Irrep) else j: n for j, n in pairs}, dual)


def PlanarSpace(n, dual=False):
    """`ℙ^n` (`src/spaces/planarspace.jl`): n dimensions of the braiding-free trivial
    sector; its tensors admit the planar operations only."""
    from .sectors import PlanarTrivial
    return GradedSpace({PlanarTrivial(): int(n)}, dual)


def Vect(pairs, dual=False):
    if isinstance(pairs, dict):
        pairs = tuple(pairs.items())
    return GradedSpace(dict(pairs), dual)


# Structure of spaces, shared by every equal space object: TensorKit keeps it in global
# caches keyed by the space (`src/spaces/structure.jl`, `src/auxiliary/caches.jl`), so the
# HomSpace objects a contraction builds afresh for its operands and result do not
# recompute their block and sub-block structure.  Bounded LRU tables.
_SPACE_CACHE_SIZE = 10_000
_HOM_STRUCTURE = OrderedDict()  # (codomain, domain, unit of a scalar space) -> structure dict
_PRODUCT_BLOCKSECTORS = OrderedDict()  # (ProductSpace, unit) -> sorted coupled sectors


def _shared(table, key, make):
    v = table.get(key)
    if v is None:
        v = make()
        table[key] = v
        if len(table) > _SPACE_CACHE_SIZE:
            table.popitem(last=False)
    else:
        table.move_to_end(key)
    return v


def clear_space_caches():
    """Drop the shared space structure (part of `empty_globalcaches`)."""
    _HOM_STRUCTURE.clear()
    _PRODUCT_BLOCKSECTORS.clear()


def _product_first_fastest(lists):
    if not lists:
        yield ()
        return
    for rest in _product_first_fastest(lists[1:]):
        for x in lists[0]:
            yield (x,) + rest


class ProductSpace:
    __slots__ = ("spaces", "_h")

    def __init__(self, *spaces):
        if len(spaces) == 1 and isinstance(spaces[0], (tuple, list)):
            spaces = tuple(spaces[0])
        self.spaces = tuple(spaces)
        self._h = hash(self.spaces)

    def __len__(self):
        return len(self.spaces)

    def __getitem__(self, i):
        return self.spaces[i]

    def __iter__(self):
        return iter(self.spaces)

    def __eq__(self, other):
        return isinstance(other, ProductSpace) and self.spaces == other.spaces

    def __hash__(self):
        return self._h

    def __mul__(self, other):
        if isinstance(other, GradedSpace):
            return ProductSpace(*(self.spaces + (other,)))
        if isinstance(other, ProductSpace):
            return ProductSpace(*(self.spaces + other.spaces))
        raise TypeError(other)

    def dims(self, s):
        return tuple(V.dim(c) for V, c in zip(self.spaces, s))

    def dim(self, s=None):
        if s is None:
            d = 1
            for V in self.spaces:
                d *= V.dim()
            return d
        d = 1
        for x in self.dims(s):
            d *= x
        return d

    def sectors(self):
        """`product(map(sectors, P)...)`: first leg fastest."""
        return _product_first_fastest([V.sectors() for V in self.spaces])

    def isdual(self):
        return tuple(V.isdual for V in self.spaces)

    def blocksectors(self, unit):
        n = len(self.spaces)
        if n == 0:
            return [unit]

        def make():
            bs = set()
            if n == 1:
                bs.update(self.spaces[0].sectors())
            else:
                for s in self.sectors():
                    bs.update(fuse_set(s, unit))
            return tuple(sorted(bs))

        return list(_shared(_PRODUCT_BLOCKSECTORS, (self, unit), make))

    def fusiontrees(self, c):
        return fusiontrees([V.sectors() for V in self.spaces], c, self.isdual())

    def hasblock(self, c):
        return len(self.fusiontrees(c)) > 0

    def blockdim(self, c):
        return sum(self.dim(f.uncoupled) for f in self.fusiontrees(c))

    def dual(self):
        return ProductSpace(*(V.dual() for V in reversed(self.spaces)))

    def __repr__(self):
        return "(" + " ⊗ ".join(repr(V) for V in self.spaces) + ")"


class HomSpace:
    """codomain ← domain."""

    __slots__ = ("codomain", "domain", "_h", "_cache", "_unit")

    def __init__(self, codomain, domain, unit=None):
        """`unit`: the unit sector, needed only when both sides are empty (a scalar:
        TensorKit's empty ProductSpace still carries its sector type)."""
        if isinstance(codomain, GradedSpace):
            codomain = ProductSpace(codomain)
        if isinstance(domain, GradedSpace):
            domain = ProductSpace(domain)
        self.codomain = codomain
        self.domain = domain
        self._h = hash((codomain, domain))
        scalar = len(codomain) == 0 and len(domain) == 0
        self._cache = _shared(_HOM_STRUCTURE, (codomain, domain, unit if scalar else None), dict)
        self._unit = unit

    numout = property(lambda self: len(self.codomain))
    numin = property(lambda self: len(self.domain))
    numind = property(lambda self: len(self.codomain) + len(self.domain))

    def __eq__(self, other):
        return isinstance(other, HomSpace) and self.codomain == other.codomain and self.domain == other.domain

    def __hash__(self):
        return self._h

    def __repr__(self):
        return f"{self.codomain!r} ← {self.domain!r}"

    def __getitem__(self, n):
        """`W[n]`: codomain space, or the dual of a domain space (0-based)."""
        if n < self.numout:
            return self.codomain[n]
        return self.domain[n - self.numout].dual()

    @property
    def unit(self):
        for V in list(self.codomain) + list(self.domain):
            for k in V._keys:
                return k.unit()
        if self._unit is not None:
            return self._unit
        raise ValueError("HomSpace without sectors")

    @property
    def sectortype(self):
        return type(self.unit)

    def adjoint(self):
        return HomSpace(self.domain, self.codomain)

    def select(self, p1, p2):
        """`select(W, (p₁, p₂))` (`homspace.jl:240-244`)."""
        cod = ProductSpace(*(self[n] for n in p1))
        dom = ProductSpace(*(self[n].dual() for n in p2))
        return HomSpace(cod, dom, unit=self.unit)

    def permute(self, p1, p2):
        """`permute(W, (p₁, p₂))` (`homspace.jl:212-217`); 0-based indices."""
        p = tuple(p1) + tuple(p2)
        if sorted(p) != list(range(self.numind)):
            raise ValueError(f"{(p1, p2)} is not a valid permutation for {self}")
        return self.select(p1, p2)

    def flip(self, inds):
        """`flip(W, I)` (`homspace.jl:252-261`): flip the codomain and domain spaces whose
        linear (0-based) index is in `inds`."""
        inds = set(inds)
        n1 = self.numout
        cod = ProductSpace(*[V.flip() if i in inds else V for i, V in enumerate(self.codomain)])
        dom = ProductSpace(*[V.flip() if i + n1 in inds else V for i, V in enumerate(self.domain)])
        return HomSpace(cod, dom, unit=self.unit)

    def compose(self, other):
        if self.domain != other.codomain:
            raise SpaceMismatch(f"{self.domain} ≠ {other.codomain}")
        # a full contraction composes to a scalar space, which keeps the sector type
        return HomSpace(self.codomain, other.domain, unit=self.unit)

    # ---------------- structure ----------------
    def blocksectors(self):
        """`_blocksectors` (`homspace.jl:96-117`)."""
        if "bs" in self._cache:
            return self._cache["bs"]
        cod, dom = self.codomain, self.domain
        n1, n2 = len(cod), len(dom)
        u = self.unit
        if n1 == 0 and n2 == 0:
            bs = [u]
        elif n1 == 0:
            bs = [c for c in dom.blocksectors(u) if c.isunit()]
        elif n2 == 0:
            bs = [c for c in cod.blocksectors(u) if c.isunit()]
        elif n2 <= n1:
            bs = [c for c in dom.blocksectors(u) if cod.hasblock(c)]
        else:
            bs = [c for c in cod.blocksectors(u) if dom.hasblock(c)]
        self._cache["bs"] = bs
        return bs

    def sectorstructure(self):
        """`sectorstructure` (`structure.jl:41-72`): list of tree pairs
        (f₁ fastest within each f₂, coupled sectors sorted)."""
        if "ss" in self._cache:
            return self._cache["ss"]
        trees = []
        for c in self.blocksectors():
            t1 = self.codomain.fusiontrees(c)
            for f2 in self.domain.fusiontrees(c):
                for f1 in t1:
                    trees.append((f1, f2))
        self._cache["ss"] = trees
        return trees

    def fusiontrees(self):
        return self.sectorstructure()

    def degeneracystructure(self):
        """`degeneracystructure` (`structure.jl:114-188`).

        Returns (totaldim, blocks, subblocks) with blocks a list of
        (c, (d₁, d₂), offset) and subblocks a dict tree-pair →
        (sizes, strides, offset); all offsets 0-based."""
        if "ds" in self._cache:
            return self._cache["ds"]
        cod, dom = self.codomain, self.domain
        blocks = []
        subs = {}
        offset = 0
        for c in self.blocksectors():
            t1 = cod.fusiontrees(c)
            t2 = dom.fusiontrees(c)
            d1s = [cod.dims(f1.uncoupled) for f1 in t1]
            d1 = [_prod(x) for x in d1s]
            bd1 = sum(d1)
            off2 = 0
            for f2 in t2:
                d2s = dom.dims(f2.uncoupled)
                d2 = _prod(d2s)
                off1 = 0
                for f1, ds1, dd1 in zip(t1, d1s, d1):
                    sz = tuple(ds1) + tuple(d2s)
                    strides = _subblock_strides(ds1, d2s, bd1)
                    subs[(f1, f2)] = (sz, strides, offset + off2 * bd1 + off1)
                    off1 += dd1
                off2 += d2
            bd2 = off2
            blocks.append((c, (bd1, bd2), offset))
            offset += bd1 * bd2
        res = (offset, blocks, subs)
        self._cache["ds"] = res
        return res

    def dim(self):
        return self.degeneracystructure()[0]

    def blockstructure(self):
        return {c: (sz, off) for c, sz, off in self.degeneracystructure()[1]}

    def subblockstructure(self):
        return self.degeneracystructure()[2]

    def fusionblocks(self):
        """`fusionblocks(W)` (`homspace.jl:178-188`): domain uncoupled outer,
        codomain uncoupled inner, empty blocks skipped."""
        if "fb" in self._cache:
            return self._cache["fb"]
        isd = (self.codomain.isdual(), self.domain.isdual())
        out = []
        u = self.unit
        for ds in self.domain.sectors():
            for cs in self.codomain.sectors():
                b = FusionTreeBlock((cs, ds), isd, u)
                if len(b):
                    out.append(b)
        self._cache["fb"] = out
        return out


def _prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


def _subblock_strides(d1s, d2s, blockdim1):
    """Column-major strides of a sub-block inside its (d₁ × d₂) block with
    block strides (1, blockdim₁) (`structure.jl:190-196`).  Strides of size-1
    dims carry no information; they are set to the running product."""
    strides = []
    s = 1
    for d in d1s:
        strides.append(s)
        s *= d
    s = blockdim1
    for d in d2s:
        strides.append(s)
        s *= d
    return tuple(strides)


def otimes(*spaces):
    out = []
    for V in spaces:
        if isinstance(V, ProductSpace):
            out.extend(V.spaces)
        else:
            out.append(V)
    return ProductSpace(*out)
