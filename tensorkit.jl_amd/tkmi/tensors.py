"""Device TensorMap and the index-manipulation / contraction entry points.

Host-side mirror of TensorKit's L4 layer for the hot path.  A `TensorMap`
owns one flat device vector (a torch tensor on the HIP device, used purely as
device memory) plus its HomSpace (`src/tensors/tensor.jl:10-20`); blocks and
sub-blocks are views defined by `degeneracystructure`.  Every arithmetic
operation goes through libtkmi.so.

Entry points and the reference functions they mirror:
  permute / permute_ / braid_ / add_transform_   src/tensors/indexmanipulations.jl:217-318,557-580
  tensoradd_                                     src/tensors/tensoroperations.jl:44-62
  mul_                                           src/tensors/linalg.jl:330-370
  tensorcontract_ / contract_ / blas_contract_   src/tensors/tensoroperations.jl:130-161,318-462
  twist_                                         src/tensors/indexmanipulations.jl:62-78
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .spaces import HomSpace, SpaceMismatch, IndexError_
from .sectors import BOSONIC, FERMIONIC, NOBRAIDING
from .spaces import ProductSpace, SectorMismatch
from . import treetransformers as TT
from .treetransformers import (treebraider, flat_plan, twist_factor, tracer, _make_trace_plan, _TRANSFORMER_CACHE,
                               _LRU, pinning, DEFAULT_GLOBALCACHE_SIZE)
from . import fusedcontract as _FC

# contractions with a tiny operand and skinny per-sector GEMMs run as one fused
# transform (see fusedcontract); False restores permute + mul! + permute
FUSE_SMALL_OPERAND = True

_DT = {torch.float64: _lib.TKMI_F64, torch.float32: _lib.TKMI_F32,
       torch.complex128: _lib.TKMI_C128, torch.complex64: _lib.TKMI_C64}
_COMPLEX = (torch.complex128, torch.complex64)
_NP = {torch.float64: np.float64, torch.float32: np.float32, torch.complex128: np.complex128,
       torch.complex64: np.complex64}

DEFAULT_PRECISION = _lib.PREC_F64


def _device():
    _lib.ensure_device(0)
    if not torch.cuda.is_available():
        raise RuntimeError("tkmi: no HIP device available")
    return torch.device("cuda", 0)


def _stream():
    return torch.cuda.current_stream().cuda_stream


# Optional operation recorder: when `_TRACE` is a list, every device launch
# appends (kind, plan, buffers, α, β) so a caller can replay or account for the
# exact sequence of kernels a high-level call issued.
_TRACE = None


def start_trace():
    global _TRACE
    _TRACE = []


def stop_trace():
    global _TRACE
    tr, _TRACE = _TRACE, None
    return tr


def _texec(plan, dtype, src_ptr, dst_ptr, alpha, beta, conj=False, what="transform"):
    """One transform launch, dst = β·dst + α·transform(src); complex element types take
    complex α and β and read the source conjugated when `conj` (the data of the adjoint of
    a complex tensor)."""
    lib = _lib.load()
    a, b = complex(alpha), complex(beta)
    if dtype in _COMPLEX:
        _lib.check(lib.tkmi_transform_exec_ex(plan.handle, _DT[dtype], int(bool(conj)), src_ptr, dst_ptr, a.real,
                                              a.imag, b.real, b.imag, _stream()), what)
        return
    if a.imag or b.imag:
        raise TypeError(f"{what}: complex scalars need complex tensors")
    _lib.check(lib.tkmi_transform_exec(plan.handle, _DT[dtype], src_ptr, dst_ptr, a.real, b.real, _stream()), what)


def _scalar(x):
    """A scalar as the trace records it (complex only when it has an imaginary part)."""
    z = complex(x)
    return z if z.imag else z.real


def _buf(t):
    # the recorded entry holds the device tensor itself, so a recorded buffer
    # stays allocated (and its address valid) for as long as the trace lives
    return (t.data, t.data.numel())


class TensorMap:
    """Block-sparse tensor: flat device data vector + HomSpace."""

    __slots__ = ("data", "space")

    def __init__(self, data, space: HomSpace):
        if data.dim() != 1 or data.numel() != space.dim():
            raise SpaceMismatch(f"data length {data.numel()} does not match dim {space.dim()} of {space}")
        if data.dtype not in _DT:
            raise TypeError("tkmi TensorMap data must be float64, float32, complex128 or complex64")
        self.data = data
        self.space = space

    # ---------------- construction ----------------
    @staticmethod
    def empty(space, dtype=torch.float64):
        return TensorMap(torch.empty(space.dim(), dtype=dtype, device=_device()), space)

    @staticmethod
    def zeros(space, dtype=torch.float64):
        t = TensorMap.empty(space, dtype)
        scale_(t, 0.0)
        return t

    @staticmethod
    def from_numpy(arr, space, dtype=torch.float64):
        arr = np.ascontiguousarray(arr, dtype=_NP[dtype])
        return TensorMap(torch.from_numpy(arr).to(_device()), space)

    @staticmethod
    def randn(space, rng: np.random.Generator, dtype=torch.float64):
        """Seeded standard-normal data drawn on the host (the seed stream is
        this package's, not Julia's Xoshiro); complex types draw the real parts, then
        the imaginary parts."""
        x = rng.standard_normal(space.dim())
        if dtype in _COMPLEX:
            x = x + 1j * rng.standard_normal(space.dim())
        return TensorMap.from_numpy(x, space, dtype)

    def numpy(self):
        return self.data.detach().cpu().numpy()

    # ---------------- properties ----------------
    codomain = property(lambda self: self.space.codomain)
    domain = property(lambda self: self.space.domain)
    numout = property(lambda self: self.space.numout)
    numin = property(lambda self: self.space.numin)
    numind = property(lambda self: self.space.numind)
    dtype = property(lambda self: self.data.dtype)

    def dim(self):
        return self.space.dim()

    def adjoint(self):
        return AdjointTensorMap(self)

    def codomainind(self):
        return tuple(range(self.numout))

    def domainind(self):
        return tuple(range(self.numout, self.numind))

    def similar(self, space=None, dtype=None):
        return TensorMap.empty(space or self.space, dtype or self.dtype)

    def copy(self):
        t = self.similar()
        add_(t, self, 1.0, 0.0)
        return t

    def blocks_host(self):
        """Host copies of all blocks, keyed by coupled sector (column-major)."""
        d = self.numpy()
        out = {}
        for c, (d1, d2), off in self.space.degeneracystructure()[1]:
            out[c] = d[off:off + d1 * d2].reshape((d1, d2), order="F")
        return out

    def __repr__(self):
        return f"TensorMap({self.space}, dtype={self.dtype})"


class AdjointTensorMap:
    """Lazy adjoint (`src/tensors/adjoint.jl:9-53`): shares the parent's data."""

    __slots__ = ("parent",)

    def __init__(self, parent: TensorMap):
        self.parent = parent

    space = property(lambda self: self.parent.space.adjoint())
    numout = property(lambda self: self.parent.numin)
    numin = property(lambda self: self.parent.numout)
    numind = property(lambda self: self.parent.numind)
    dtype = property(lambda self: self.parent.dtype)
    data = property(lambda self: self.parent.data)

    def dim(self):
        return self.parent.dim()

    def adjoint(self):
        return self.parent

    def codomainind(self):
        return tuple(range(self.numout))

    def domainind(self):
        return tuple(range(self.numout, self.numind))


def adjointtensorindices(t, p):
    """`adjointtensorindex`: index i of t ↦ index of t′ (0-based)."""
    n1, n2 = t.numout, t.numin
    return tuple(n2 + i if i < n1 else i - n1 for i in p)


def _is_identity(t, p1, p2):
    return tuple(p1) == t.codomainind() and tuple(p2) == t.domainind()


# ---------------------------------------------------------------------------
# elementwise
# ---------------------------------------------------------------------------
def _exec_flat(dst, src, alpha, beta):
    n = dst.data.numel()
    if n == 0:
        return
    plan = flat_plan(n)
    if _TRACE is not None:
        _TRACE.append(("transform", plan, (0,), _buf(src), _buf(dst), _scalar(alpha), _scalar(beta)))
    _texec(plan, dst.dtype, src.data.data_ptr(), dst.data.data_ptr(), alpha, beta, what="add!")


def add_(tdst, tsrc, alpha=1.0, beta=1.0):
    """`add!(tdst, tsrc, α, β)`: tdst = β·tdst + α·tsrc (same space)."""
    if tdst.space != tsrc.space:
        raise SpaceMismatch(f"{tdst.space} ≠ {tsrc.space}")
    if tdst.dtype != tsrc.dtype:
        raise TypeError("add!: dtype mismatch")
    _exec_flat(tdst, tsrc, alpha, beta)
    return tdst


def scale_(t, alpha):
    """`scale!(t, α)` in place (α = 0 writes zeros without reading t)."""
    _exec_flat(t, t, alpha, 0.0)
    return t


# ---------------------------------------------------------------------------
# permute / braid
# ---------------------------------------------------------------------------
def spacecheck_transform(tdst, tsrc, p1, p2):
    p = tuple(p1) + tuple(p2)
    if sorted(p) != list(range(tsrc.numind)):
        raise IndexError_(f"{(p1, p2)} is not a valid permutation for {tsrc.space}")
    if tdst.space != tsrc.space.permute(p1, p2):
        raise SpaceMismatch(f"incompatible spaces for permute: {tdst.space} ≠ permute({tsrc.space}, {(p1, p2)})")


def add_transform_(tdst, tsrc, p1, p2, alpha, beta, levels=None, dst_twist=()):
    """`add_transform!` (`indexmanipulations.jl:557-580`) with the device kernel."""
    spacecheck_transform(tdst, tsrc, p1, p2)
    if tsrc.space.unit.braiding_style == NOBRAIDING and not _is_identity(tsrc, p1, p2):
        raise SectorMismatch("permute and braid need a braiding; use the planar operations (transpose)")
    if tdst.dtype != tsrc.dtype:
        raise TypeError("permute!: dtype mismatch")
    is_adj = isinstance(tsrc, AdjointTensorMap)
    if not is_adj and not dst_twist and _is_identity(tsrc, p1, p2):
        return add_(tdst, tsrc, alpha, beta)
    plan = treebraider(tdst.space, tsrc.space, tuple(p1), tuple(p2), levels, adjoint_src=is_adj,
                       dst_twist=tuple(dst_twist))
    if _TRACE is not None:
        _TRACE.append(("transform", plan, tuple(p1) + tuple(p2), _buf(tsrc), _buf(tdst), _scalar(alpha),
                       _scalar(beta)))
    _texec(plan, tdst.dtype, tsrc.data.data_ptr(), tdst.data.data_ptr(), alpha, beta, conj=is_adj,
           what="add_transform!")
    return tdst


def braid_(tdst, tsrc, p1, p2, levels, alpha=1.0, beta=0.0):
    """`braid!(tdst, tsrc, (p₁, p₂), levels, α, β)` (`indexmanipulations.jl:304-318`)."""
    n1 = tsrc.numout
    lv = (tuple(levels[:n1]), tuple(levels[n1:]))
    if len(set(levels)) != len(levels) and tsrc.space.unit.braiding_style not in ("bosonic", "fermionic"):
        raise ValueError("ambiguous braid levels")
    return add_transform_(tdst, tsrc, p1, p2, alpha, beta, levels=lv)


def permute_(tdst, tsrc, p1, p2, alpha=1.0, beta=0.0):
    """`permute!(tdst, tsrc, p, α, β)` (`indexmanipulations.jl:217-225`)."""
    return braid_(tdst, tsrc, p1, p2, tuple(range(tsrc.numind)), alpha, beta)


def permute(t, p1, p2=(), copy=False):
    """`permute(t, (p₁, p₂); copy)` (`indexmanipulations.jl:242-258`): shares
    data for the identity permutation unless `copy`."""
    p1, p2 = tuple(p1), tuple(p2)
    if isinstance(t, AdjointTensorMap):
        q1 = adjointtensorindices(t, p2)
        q2 = adjointtensorindices(t, p1)
        return permute(t.parent, q1, q2, copy=copy).adjoint()
    if not copy and _is_identity(t, p1, p2):
        return t
    tdst = TensorMap.empty(t.space.permute(p1, p2), t.dtype)
    return permute_(tdst, t, p1, p2, 1.0, 0.0)


def braid(t, p1, p2, levels, copy=False):
    p1, p2 = tuple(p1), tuple(p2)
    if not copy and _is_identity(t, p1, p2):
        return t
    tdst = TensorMap.empty(t.space.permute(p1, p2), t.dtype)
    return braid_(tdst, t, p1, p2, levels)


def _iscyclic(p):
    n = len(p)
    return n == 0 or all(p[(i + 1) % n] == (p[i] + 1) % n for i in range(n))


def transpose(t, p1=None, p2=None):
    """`transpose(t, (p₁, p₂))` (`indexmanipulations.jl:380-420`): a planar
    (cyclic, in the linearised leg order) repartition, into a new tensor.  Default:
    (reverse(domain), reverse(codomain))."""
    if p1 is None:
        p1 = tuple(reversed(t.domainind()))
        p2 = tuple(reversed(t.codomainind()))
    p1, p2 = _check_transpose(t, p1, p2)
    tdst = TensorMap.empty(t.space.permute(p1, p2), t.dtype)
    return transpose_(tdst, t, p1, p2, 1.0, 0.0)


def _check_transpose(t, p1, p2):
    p1, p2 = tuple(p1), tuple(p2)
    n1, n2 = t.numout, t.numin
    if sorted(p1 + p2) != list(range(t.numind)):
        raise IndexError_(f"{(p1, p2)} is not a valid permutation for {t.space}")
    lin = [x if x < n1 else n2 + 2 * n1 - 1 - x for x in p1] + \
          [x if x < n1 else n2 + 2 * n1 - 1 - x for x in reversed(p2)]
    if not _iscyclic(lin):
        raise ValueError(f"{(p1, p2)} is not a cyclic permutation for {t.space}")
    return p1, p2


def _repartition_perm(tsrc, n1):
    """(p₁, p₂) of a repartition (`indexmanipulations.jl:452-454`): the legs in the
    cyclic order codomain…, reverse(domain)…, the first n1 into the codomain."""
    inds = tsrc.codomainind() + tuple(reversed(tsrc.domainind()))
    return inds[:n1], tuple(reversed(inds[n1:]))


def repartition_(tdst, tsrc, alpha=1.0, beta=0.0):
    """`repartition!(tdst, tsrc, α, β)` (`indexmanipulations.jl:445-457`): move legs
    between codomain and domain without changing their cyclic order — a `transpose!`
    (bends only: the linearised leg order is unchanged)."""
    if tsrc.numind != tdst.numind:
        raise ValueError("tsrc and tdst should have an equal amount of indices")
    p1, p2 = _repartition_perm(tsrc, tdst.numout)
    return transpose_(tdst, tsrc, p1, p2, alpha, beta)


def repartition(tsrc, n1, n2=None, copy=False):
    """`repartition(tsrc, N₁, N₂; copy)` (`indexmanipulations.jl:474-484`)."""
    n2 = tsrc.numind - n1 if n2 is None else n2
    if n1 + n2 != tsrc.numind:
        raise ValueError(f"({n1}, {n2}) is not a partition of the {tsrc.numind} indices of {tsrc.space}")
    p1, p2 = _repartition_perm(tsrc, n1)
    if not copy and _is_identity(tsrc, p1, p2):
        return tsrc
    tdst = TensorMap.empty(tsrc.space.permute(p1, p2), tsrc.dtype)
    return transpose_(tdst, tsrc, p1, p2, 1.0, 0.0)


def _deprecated(old, new):
    import warnings
    warnings.warn(f"`{old}` is deprecated, use `{new}` instead", DeprecationWarning, stacklevel=3)


def add_permute_(tdst, tsrc, p, alpha, beta):
    """Deprecated `add_permute!` (`indexmanipulations.jl:542-545`) → `permute!`."""
    _deprecated("add_permute!", "permute!")
    return permute_(tdst, tsrc, p[0], p[1], alpha, beta)


def add_braid_(tdst, tsrc, p, levels, alpha, beta):
    """Deprecated `add_braid!` (`indexmanipulations.jl:546-549`) → `braid!`."""
    _deprecated("add_braid!", "braid!")
    return braid_(tdst, tsrc, p[0], p[1], levels, alpha, beta)


def add_transpose_(tdst, tsrc, p, alpha, beta):
    """Deprecated `add_transpose!` (`indexmanipulations.jl:550-553`) → `transpose!`."""
    _deprecated("add_transpose!", "transpose!")
    return transpose_(tdst, tsrc, p[0], p[1], alpha, beta)


def transpose_(tdst, tsrc, p1, p2, alpha=1.0, beta=0.0):
    """`transpose!(tdst, tsrc, (p₁, p₂), α, β)` (`indexmanipulations.jl:380-392`): the
    planar transpose.  Bosonic sectors: the planar coefficients are the braided ones, a
    cyclic `permute!`.  Otherwise (fermions, sectors without a braiding) the transform
    applies the bend/fold coefficients of `transpose((f₁, f₂), p)`; the data moves as in
    a permute, with the same rounding (tensoradd! of α·coeff·src onto β·dst)."""
    p1, p2 = _check_transpose(tsrc, p1, p2)
    if tsrc.space.unit.braiding_style == BOSONIC:
        return permute_(tdst, tsrc, p1, p2, alpha, beta)
    spacecheck_transform(tdst, tsrc, p1, p2)
    if tdst.dtype != tsrc.dtype:
        raise TypeError("transpose!: dtype mismatch")
    is_adj = isinstance(tsrc, AdjointTensorMap)
    if not is_adj and _is_identity(tsrc, p1, p2):
        return add_(tdst, tsrc, alpha, beta)
    plan = TT.treetransposer(tdst.space, tsrc.space, p1, p2, adjoint_src=is_adj)
    if _TRACE is not None:
        _TRACE.append(("transform", plan, p1 + p2, _buf(tsrc), _buf(tdst), _scalar(alpha), _scalar(beta)))
    _texec(plan, tdst.dtype, tsrc.data.data_ptr(), tdst.data.data_ptr(), alpha, beta, conj=is_adj,
           what="transpose!")
    return tdst


def planaradd_(C, A, p, alpha=1.0, beta=0.0):
    """`planaradd!(C, A, (p₁, p₂), α, β)` (`src/planar/planaroperations.jl:29-36`): the
    planar transpose `transpose!(C, A, p, α, β)`."""
    return transpose_(C, A, p[0], p[1], alpha, beta)


def _cyclicpermute(t):
    return tuple(t[1:]) + tuple(t[:1])


def reorder_indices(codA, domA, codB, domB, oindA, cindA, oindB, cindB, p1, p2):
    """`reorder_indices` (`src/planar/planaroperations.jl:196-251`), 0-based: the open and
    contracted legs of A and B in the cyclic orders a planar contraction needs —
    A as (oindA2 ← cindA2), B as (cindB2 ← oindB2), the contracted legs paired as before
    and C's legs (oindA2..., oindB2...) in the order (p₁, p₂) asks for."""
    n1, n2 = len(oindA), len(oindB)
    if len(p1) != n1 or sorted(p1) != list(range(n1)) or len(p2) != n2 or \
            sorted(p2) != list(range(n1, n1 + n2)):
        raise IndexError_(f"{(p1, p2)} does not place A's open legs in the codomain and B's in the domain")
    oindA2 = tuple(oindA[i] for i in p1)
    oindB2 = tuple(oindB[i - n1] for i in p2)
    indA = tuple(codA) + tuple(reversed(domA))
    indB = tuple(codB) + tuple(reversed(domB))
    # cycle indA to (oindA2..., reverse(cindA2)...) and indB to (cindB2..., reverse(oindB2)...)
    for _ in range(len(indA)):
        if not oindA2 or indA[0] == oindA2[0]:
            break
        indA = _cyclicpermute(indA)
    for _ in range(len(indB)):
        if not oindB2 or indB[-1] == oindB2[0]:
            break
        indB = _cyclicpermute(indB)
    if tuple(indA[:n1]) != oindA2 or any(indB[len(indB) - 1 - j] != oindB2[j] for j in range(n2)):
        raise ValueError("the contraction is not planar: open legs are not in cyclic order")
    nc = len(indA) - n1
    if nc != len(indB) - n2:
        raise IndexError_("numbers of contracted legs of A and B differ")
    cindA2 = tuple(reversed(indA[n1:n1 + nc]))
    cindB2 = tuple(indB[:nc])
    # with no open legs on one side the cyclic order fixes the contracted legs only up to
    # a rotation: align them with the pairing given by (cindA, cindB)
    if not oindA2 and cindA:
        hit = cindA[list(cindB).index(cindB2[0])]
        for _ in range(len(cindA2)):
            if hit == cindA2[0]:
                break
            cindA2 = _cyclicpermute(cindA2)
    if not oindB2 and cindB:
        hit = cindB[list(cindA).index(cindA2[0])]
        for _ in range(len(cindB2)):
            if hit == cindB2[0]:
                break
            cindB2 = _cyclicpermute(cindB2)
    if sorted(cindA2) != sorted(cindA) or sorted(zip(cindA2, cindB2)) != sorted(zip(cindA, cindB)):
        raise ValueError("the contraction is not planar: contracted legs cannot be paired in cyclic order")
    return oindA2, cindA2, oindB2, cindB2


def planarcontract_(C, A, pA, B, pB, pAB, alpha=1.0, beta=0.0, precision=None):
    """`planarcontract!` (`src/planar/planaroperations.jl:150-190`).  Bosonic sectors
    forward to `contract!`.  Otherwise: `reorder_indices`, the planar transposes of A into
    (oindA ← cindA) and of B into (cindB ← oindB) where they are not already in that
    form, and one grouped `mul!` into C."""
    if C.space.unit.braiding_style == BOSONIC:
        return contract_(C, A, pA, B, pB, pAB, alpha, beta, precision)
    codA, domA = A.codomainind(), A.domainind()
    codB, domB = B.codomainind(), B.domainind()
    (oindA, cindA), (cindB, oindB) = (tuple(pA[0]), tuple(pA[1])), (tuple(pB[0]), tuple(pB[1]))
    oindA, cindA, oindB, cindB = reorder_indices(codA, domA, codB, domB, oindA, cindA, oindB, cindB,
                                                 tuple(pAB[0]), tuple(pAB[1]))
    if oindA == codA and cindA == domA and not isinstance(A, AdjointTensorMap):
        A2 = A
    else:
        A2 = transpose(A, oindA, cindA)
    if cindB == codB and oindB == domB and not isinstance(B, AdjointTensorMap):
        B2 = B
    else:
        B2 = transpose(B, cindB, oindB)
    return mul_(C, A2, B2, alpha, beta, precision)


def _legs(t, inds):
    inds = (inds,) if isinstance(inds, int) else tuple(inds)
    if not all(0 <= i < t.numind for i in inds):
        raise ValueError(f"Can't twist indices {inds} of a tensor with only {t.numind} indices.")
    return inds


def has_shared_twist(t, inds):
    """`has_shared_twist(t, inds)` (`indexmanipulations.jl:34-51`): the twist of those legs
    is the identity (bosonic braiding, or every sector on them has θ = 1)."""
    if t.space.unit.braiding_style == BOSONIC:
        return True
    W = t.space
    return all(c.twist() == 1 for i in inds for c in W[i].sectors())


def twist_(t, inds, inv=False):
    """`twist!(t, inds; inv)` (`indexmanipulations.jl:62-78`): scale each sub-block by the
    product of the twists θ of the legs in `inds` (θ⁻¹ with `inv`; the twists here are
    ±1, their own inverses).  Nothing is launched when the twist is the identity."""
    inds = _legs(t, inds)
    if not inds or has_shared_twist(t, inds):
        return t
    tmp = t.copy()
    add_transform_(t, tmp, t.codomainind(), t.domainind(), 1.0, 0.0, dst_twist=inds)
    return t


def twist(t, inds, inv=False, copy=False):
    """`twist(t, inds; inv, copy)` (`indexmanipulations.jl:90-97`): a twisted copy, or `t`
    itself when the twist is the identity and `copy` is false."""
    inds = _legs(t, inds)
    if not copy and has_shared_twist(t, inds):
        return t
    return twist_(t.copy(), inds, inv)


def flip(t, inds, inv=False):
    """`flip(t, I; inv)` (`indexmanipulations.jl:21-29`): a tensor isomorphic to `t` whose
    legs in I have their arrows flipped (`space(t′, i) = flip(space(t, i))`); each tree
    pair's sub-block is scaled by the coefficient of `flip((f₁, f₂), I; inv)`.  Not an
    involution: flip(flip(t, I), I; inv=True) == t."""
    inds = (inds,) if isinstance(inds, int) else tuple(inds)
    if not all(0 <= i < t.numind for i in inds):
        raise ValueError(f"Can't flip indices {inds} of a tensor with only {t.numind} indices.")
    if isinstance(t, AdjointTensorMap):
        t = permute(t, t.codomainind(), t.domainind(), copy=True)
    W2 = t.space.flip(inds)
    plan = TT.flipper(W2, t.space, inds, inv)
    out = TensorMap.empty(W2, t.dtype)
    _texec(plan, t.dtype, t.data.data_ptr(), out.data.data_ptr(), 1.0, 0.0, what="flip")
    return out


# ---------------------------------------------------------------------------
# partial trace: trace_permute! / TO.tensortrace! / planartrace! / tr
# ---------------------------------------------------------------------------
def trace_permute_(tdst, tsrc, p, q, alpha=1.0, beta=0.0):
    """`trace_permute!(tdst, tsrc, (p₁, p₂), (q₁, q₂), α, β)`
    (`src/tensors/tensoroperations.jl:203-300`): tdst = β·tdst + α·(tsrc with the legs
    q₁[i], q₂[i] traced and the rest permuted to (p₁, p₂)).  One device transform:
    each destination sub-block sums the traced diagonal slices of the source sub-blocks
    whose traced remainders match, with the reference's coefficients."""
    (p1, p2), (q1, q2) = (tuple(p[0]), tuple(p[1])), (tuple(q[0]), tuple(q[1]))
    if tsrc.space.unit.braiding_style not in (BOSONIC, FERMIONIC):
        raise SectorMismatch("only tensors with symmetric braiding rules can be contracted; try `@planar` instead")
    if len(q1) != len(q2):
        raise IndexError_("number of trace indices does not match")
    if sorted(p1 + p2 + q1 + q2) != list(range(tsrc.numind)):
        raise IndexError_(f"{(p1, p2)}, {(q1, q2)} do not cover the indices of {tsrc.space}")
    W = tsrc.space
    if tdst.space != W.select(p1, p2):
        raise SpaceMismatch(f"trace: tsrc = {W}, tdst = {tdst.space}, p₁ = {p1}, p₂ = {p2}")
    if any(W[a] != W[b].dual() for a, b in zip(q1, q2)):
        raise SpaceMismatch(f"trace: tsrc = {W}, q₁ = {q1}, q₂ = {q2}")
    if tdst.dtype != tsrc.dtype:
        raise TypeError("trace_permute!: dtype mismatch")
    if isinstance(tsrc, AdjointTensorMap):
        tsrc = permute(tsrc, tsrc.codomainind(), tsrc.domainind(), copy=True)
    plan = tracer(tdst.space, W, p1, p2, q1, q2)
    _texec(plan, tdst.dtype, tsrc.data.data_ptr(), tdst.data.data_ptr(), alpha, beta, what="trace_permute!")
    return tdst


def tensortrace_(C, A, p, q, conjA=False, alpha=1.0, beta=0.0):
    """`TO.tensortrace!(C, A, p, q, conjA, α, β)` (`tensoroperations.jl:82-98`)."""
    if conjA:
        Ap = A.adjoint()
        pc = _canonicalize(p, C)
        p2 = (adjointtensorindices(A, pc[0]), adjointtensorindices(A, pc[1]))
        q2 = (adjointtensorindices(A, q[0]), adjointtensorindices(A, q[1]))
        return trace_permute_(C, Ap, p2, q2, alpha, beta)
    return trace_permute_(C, A, _canonicalize(p, C), q, alpha, beta)


def tensortrace(A, p, q, conjA=False, alpha=1.0):
    """Allocating form: the result's space is `select(space(A), p)`."""
    Asp = A.space if not conjA else A.space.adjoint()
    C = TensorMap.empty(Asp.select(tuple(p[0]), tuple(p[1])), A.dtype)
    return tensortrace_(C, A, p, q, conjA, alpha, 0.0)


def planartrace_(C, A, p, q, alpha=1.0, beta=0.0):
    """`planartrace!` (`src/planar/planaroperations.jl:66-111`).  Bosonic sectors forward
    to `trace_permute!`.  Otherwise every source tree pair contributes through the
    coefficients of `planar_trace((f₁, f₂), p, q)` — one device transform summing the
    traced diagonal slices, C = β·C + α·Σ coeff·(trace of A[f₁, f₂])."""
    if A.space.unit.braiding_style == BOSONIC:
        return trace_permute_(C, A, p, q, alpha, beta)
    (p1, p2), (q1, q2) = (tuple(p[0]), tuple(p[1])), (tuple(q[0]), tuple(q[1]))
    if len(q1) != len(q2):
        raise IndexError_("number of trace indices does not match")
    if sorted(p1 + p2 + q1 + q2) != list(range(A.numind)):
        raise IndexError_(f"{(p1, p2)}, {(q1, q2)} do not cover the indices of {A.space}")
    W = A.space
    if C.space != W.select(p1, p2):
        raise SpaceMismatch(f"trace: A = {W}, C = {C.space}, p₁ = {p1}, p₂ = {p2}")
    if any(W[a] != W[b].dual() for a, b in zip(q1, q2)):
        raise SpaceMismatch(f"trace: A = {W}, q₁ = {q1}, q₂ = {q2}")
    if C.dtype != A.dtype:
        raise TypeError("planartrace!: dtype mismatch")
    if isinstance(A, AdjointTensorMap):
        A = transpose(A, A.codomainind(), A.domainind())
    plan = tracer(C.space, W, p1, p2, q1, q2, planar=True)
    _texec(plan, C.dtype, A.data.data_ptr(), C.data.data_ptr(), alpha, beta, what="planartrace!")
    return C


def tr(t):
    """`tr(t)` = Σ_c dim(c)·tr(block_c) (`src/tensors/linalg.jl:319-327`), for
    domain(t) == codomain(t): one device transform summing every block diagonal (with
    weight dim(c)) into a scalar."""
    W = t.space
    if W.domain != W.codomain:
        raise SpaceMismatch("Trace of a tensor only exist when domain == codomain")
    if isinstance(t, AdjointTensorMap):
        t = permute(t, t.codomainind(), t.domainind(), copy=True)
    key = ("tr", W)
    plan = _TRANSFORMER_CACHE.get(key)
    if plan is None:
        coeffs, slices = [], []
        for c, (d1, d2), off in W.degeneracystructure()[1]:
            for i in range(d1):
                slices.append([off + i * (d1 + 1)])
                coeffs.append(float(c.dim()))
        rows = [(np.array(coeffs, dtype=np.float64).reshape(1, -1), (), [[0]], slices)]
        plan = _make_trace_plan(rows, 0, 1, W.dim())
        _TRANSFORMER_CACHE.put(key, plan)
    out = torch.empty(1, dtype=t.dtype, device=_device())
    _texec(plan, t.dtype, t.data.data_ptr(), out.data_ptr(), 1.0, 0.0, what="tr")
    return complex(out.item()) if t.dtype in _COMPLEX else float(out.item())


# ---------------------------------------------------------------------------
# tensoradd!
# ---------------------------------------------------------------------------
def tensoradd_(C, A, pA, conjA, alpha=1.0, beta=0.0):
    """`TO.tensoradd!(C, A, pA, conjA, α, β)` (`tensoroperations.jl:44-62`)."""
    p1, p2 = _canonicalize(pA, C)
    if conjA:
        Ap = A.adjoint()
        q1 = adjointtensorindices(A, p1)
        q2 = adjointtensorindices(A, p2)
        return permute_(C, Ap, q1, q2, alpha, beta)
    return permute_(C, A, p1, p2, alpha, beta)


def _canonicalize(p, C):
    p1, p2 = p
    if len(p1) == C.numout and len(p2) == C.numin:
        return tuple(p1), tuple(p2)
    lin = tuple(p1) + tuple(p2)
    return lin[:C.numout], lin[C.numout:]


# ---------------------------------------------------------------------------
# mul! (grouped GEMM per coupled sector)
# ---------------------------------------------------------------------------
class GemmPlan:
    def __init__(self, handle, flops, ngroups):
        self.handle = handle
        self._destroy = _lib.load().tkmi_gemm_plan_destroy
        self.flops = flops
        self.ngroups = ngroups

    def __del__(self):
        # the destroy function is bound at creation: module globals may already be
        # torn down when the interpreter exits
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self.handle = None


_GEMM_CACHE = _LRU(DEFAULT_GLOBALCACHE_SIZE)


def _block_layout(t):
    """(c -> (rows, cols, offset, ld, trans)) of op(block) as seen by mul!."""
    if isinstance(t, AdjointTensorMap):
        out = {}
        for c, (d1, d2), off in t.parent.space.degeneracystructure()[1]:
            out[c] = (d2, d1, off, max(d1, 1), 1)
        return out
    return {c: (d1, d2, off, max(d1, 1), 0) for c, (d1, d2), off in t.space.degeneracystructure()[1]}


def gemm_plan(tC, tA, tB, with_beta_groups=True):
    key = (tC.space, tA.space, tB.space, isinstance(tA, AdjointTensorMap), isinstance(tB, AdjointTensorMap))
    plan = _GEMM_CACHE.get(key)
    if plan is not None:
        return plan
    LA, LB = _block_layout(tA), _block_layout(tB)
    rows = []
    for c, (d1, d2), off in tC.space.degeneracystructure()[1]:
        if c in LA and c in LB:
            ma, ka, oa, lda, ta = LA[c]
            kb, nb, ob, ldb, tb = LB[c]
            if ma != d1 or nb != d2 or ka != kb:
                raise SpaceMismatch("mul!: block sizes disagree")
            rows.append((d1, d2, ka, oa, lda, ta, ob, ldb, tb, off, max(d1, 1)))
        else:
            # no matching block: C_c = β·C_c (rmul!)
            rows.append((d1, d2, 0, 0, 1, 0, 0, 1, 0, off, max(d1, 1)))
    arr = np.array(rows, dtype=np.int64).reshape(-1, 11)
    lib = _lib.ensure_device()
    h = _lib.C.c_void_p()
    cols = [_lib.i64(arr[:, i]) for i in range(11)]
    ta = _lib.i32(arr[:, 5])
    tb = _lib.i32(arr[:, 8])
    _lib.check(lib.tkmi_gemm_plan_create(_lib.C.byref(h), len(rows), cols[0][1], cols[1][1], cols[2][1],
                                         cols[3][1], cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1],
                                         cols[9][1], cols[10][1]), "tkmi_gemm_plan_create")
    flops = float(np.sum(2.0 * arr[:, 0] * arr[:, 1] * arr[:, 2])) if len(rows) else 0.0
    plan = GemmPlan(h, flops, len(rows))
    plan.table = arr
    _GEMM_CACHE.put(key, plan)
    return plan


def gemm_precision(dtype, precision=None):
    """The arithmetic of a grouped GEMM on tensors of element type `dtype`: float64 runs
    in DEFAULT_PRECISION unless told otherwise; float32 runs in FP32 (the sgemm of a
    Float32 mul!) or, when asked, in 3×BF16 — FP64 arithmetic on float32 data is refused."""
    if dtype == torch.float32:
        prec = _lib.PREC_F32 if precision is None else precision
        if prec == _lib.PREC_F64:
            raise TypeError("mul!: float32 tensors multiply in FP32 or 3×BF16, not in FP64")
        return prec
    if dtype in _COMPLEX:
        # zgemm / cgemm: ComplexF64 in FP64, ComplexF32 in FP32 unless told otherwise
        prec = (_lib.PREC_F64 if dtype == torch.complex128 else _lib.PREC_F32) if precision is None else precision
        if prec == _lib.PREC_BF16X3:
            raise TypeError("mul!: complex tensors multiply in FP64 or FP32")
        return prec
    return DEFAULT_PRECISION if precision is None else precision


def mul_(tC, tA, tB, alpha=1.0, beta=0.0, precision=None):
    """`mul!(tC, tA, tB, α, β)` (`linalg.jl:330-370`): one grouped launch."""
    if tA.space.compose(tB.space) != tC.space:
        raise SpaceMismatch(f"{tC.space} ≠ {tA.space} * {tB.space}")
    if not (tA.dtype == tB.dtype == tC.dtype):
        raise TypeError(f"mul!: element types differ (A {tA.dtype}, B {tB.dtype}, C {tC.dtype})")
    plan = gemm_plan(tC, tA, tB)
    prec = gemm_precision(tC.dtype, precision)
    lib = _lib.load()
    if tC.dtype in _COMPLEX:
        # an adjoint operand's blocks are conjugate-transposed
        a, b = complex(alpha), complex(beta)
        if _TRACE is not None:
            _TRACE.append(("gemm", plan, int(prec), _buf(tA), _buf(tB), _buf(tC), _scalar(alpha), _scalar(beta)))
        _lib.check(lib.tkmi_gemm_exec_z(plan.handle, int(prec), _DT[tC.dtype], int(isinstance(tA, AdjointTensorMap)),
                                        int(isinstance(tB, AdjointTensorMap)), a.real, a.imag, b.real, b.imag,
                                        tA.data.data_ptr(), tB.data.data_ptr(), tC.data.data_ptr(), _stream()),
                   "mul!")
        return tC
    if _TRACE is not None:
        _TRACE.append(("gemm", plan, int(prec), _buf(tA), _buf(tB), _buf(tC), float(alpha), float(beta)))
    _lib.check(lib.tkmi_gemm_exec_dtype(plan.handle, int(prec), _DT[tC.dtype], 3, float(alpha), float(beta),
                                        tA.data.data_ptr(), tB.data.data_ptr(), tC.data.data_ptr(), _stream()),
               "mul!")
    return tC


# ---------------------------------------------------------------------------
# contraction
# ---------------------------------------------------------------------------
def _sortperm(t):
    return tuple(sorted(range(len(t)), key=lambda i: t[i]))


def _contract_candidates(pA, pB, pAB):
    """`_contract_candidates` (`tensoroperations.jl:360-376`), 0-based."""
    n1, n2 = len(pA[0]), len(pB[1])
    qA = _sortperm(pA[1])
    pA1 = (pA[0], tuple(pA[1][i] for i in qA))
    pB1 = (tuple(pB[0][i] for i in qA), pB[1])
    qB = _sortperm(pB[0])
    pA2 = (pA[0], tuple(pA[1][i] for i in qB))
    pB2 = (tuple(pB[0][i] for i in qB), pB[1])
    f = (lambda x: x - n1 if x >= n1 else x + n2)
    pABp = (tuple(f(x) for x in pAB[0]), tuple(f(x) for x in pAB[1]))
    return pA1, pB1, pA2, pB2, pABp


def _invperm(p):
    inv = [0] * len(p)
    for i, x in enumerate(p):
        inv[x] = i
    return tuple(inv)


def oindABinC(pAB, pA, pB):
    """TensorOperations `oindABinC`: positions in C of A's and B's open indices."""
    ip = _invperm(tuple(pAB[0]) + tuple(pAB[1]))
    na, nb = len(pA[0]), len(pB[1])
    return tuple(ip[i] for i in range(na)), tuple(ip[na + i] for i in range(nb))


def _has_shared_permute(t, p):
    if isinstance(t, AdjointTensorMap):
        q1 = adjointtensorindices(t, p[1])
        q2 = adjointtensorindices(t, p[0])
        return _has_shared_permute(t.parent, (q1, q2))
    return _is_identity(t, p[0], p[1])


def _contract_memcost(dA, dB, dC, C, A, pA, B, pB, pAB):
    ipAB = oindABinC(pAB, pA, pB)
    return (dA * (not _has_shared_permute(A, pA) or A.dtype != C.dtype)
            + dB * (not _has_shared_permute(B, pB) or B.dtype != C.dtype)
            + dC * (not _has_shared_permute(C, ipAB)))


def _rev(p):
    return (p[1], p[0])


def contract_(C, A, pA, B, pB, pAB, alpha=1.0, beta=0.0, precision=None):
    """`contract!` (`tensoroperations.jl:318-357`): pick operand order and index
    sorting by memory cost, then `blas_contract!`."""
    if len(pA[1]) != len(pB[0]):
        raise IndexError_("number of contracted indices does not match")
    pA1, pB1, pA2, pB2, pABp = _contract_candidates(pA, pB, pAB)
    dA, dB, dC = A.dim(), B.dim(), C.dim()
    m1 = _contract_memcost(dA, dB, dC, C, A, pA1, B, pB1, pAB)
    m2 = _contract_memcost(dA, dB, dC, C, A, pA2, B, pB2, pAB)
    m3 = _contract_memcost(dB, dA, dC, C, B, _rev(pB1), A, _rev(pA1), pABp)
    m4 = _contract_memcost(dB, dA, dC, C, B, _rev(pB2), A, _rev(pA2), pABp)
    if min(m1, m2) <= min(m3, m4):
        if m1 <= m2:
            return blas_contract_(C, A, pA1, B, pB1, pAB, alpha, beta, precision)
        return blas_contract_(C, A, pA2, B, pB2, pAB, alpha, beta, precision)
    if m3 <= m4:
        return blas_contract_(C, B, _rev(pB1), A, _rev(pA1), pABp, alpha, beta, precision)
    return blas_contract_(C, B, _rev(pB2), A, _rev(pA2), pABp, alpha, beta, precision)


def _space_isdual(t, i):
    return t.space[i].isdual


def blas_contract_(C, A, pA, B, pB, pAB, alpha, beta, precision=None):
    """`blas_contract!` (`tensoroperations.jl:392-462`): permute operands into
    GEMM form (fermionic twists folded into the permute), grouped GEMM, and a
    final permute of the result when C is not in GEMM layout."""
    TC = C.dtype
    if A.dtype != TC or B.dtype != TC:
        # the kernels read every operand with C's element type: no promotion on the device
        raise TypeError(f"contract: element types differ (A {A.dtype}, B {B.dtype}, C {TC})")
    copyA = not (_has_shared_permute(A, pA) and A.dtype == TC)
    copyB = not (_has_shared_permute(B, pB) and B.dtype == TC)
    fermionic = C.space.unit.braiding_style == FERMIONIC
    if fermionic and any(_space_isdual(B, i) for i in pB[0]):
        if not (copyA ^ copyB):
            twistA = A.dim() < B.dim()
        else:
            twistA = copyA
        twistB = not twistA
        copyA |= twistA
        copyB |= twistB
    else:
        twistA = twistB = False

    if FUSE_SMALL_OPERAND and _FC.eligible(C, A, pA, B, pB):
        # the permutes of both operands, the skinny per-sector GEMMs and the result
        # permute as one pass
        n1A = len(pA[0])
        nA2 = n1A + len(pA[1])
        WA2, WB2 = A.space.permute(*pA), B.space.permute(*pB)
        twA = tuple(i for i in range(n1A, nA2) if not WA2[i].isdual) if twistA else ()
        twB = tuple(i for i in range(len(pB[0])) if WB2[i].isdual) if twistB else ()
        pC = _canonicalize(pAB, C)
        plan = _FC.fused_plan(C, A, pA, B, pB, pC, twA, twB)
        if _TRACE is not None:
            _TRACE.append(("fused", plan, (tuple(pA), tuple(pB), tuple(pC), twA, twB, A.space, B.space, C.space),
                           _buf(A), _buf(B), _buf(C), float(alpha), float(beta)))
        _lib.check(_lib.load().tkmi_transform_exec_gather(plan.handle, _DT[TC], A.data.data_ptr(), C.data.data_ptr(),
                                                          B.data.data_ptr(), float(alpha), float(beta), _stream()),
                   "contract (small operand)")
        return C

    if copyA:
        Anew = TensorMap.empty(A.space.permute(*pA), TC)
        tw = ()
        if twistA:
            tw = tuple(i for i in Anew.domainind() if not Anew.space[i].isdual)
        _transform_into(Anew, A, pA, tw)
    else:
        Anew = permute(A, *pA)
    if copyB:
        Bnew = TensorMap.empty(B.space.permute(*pB), TC)
        tw = ()
        if twistB:
            tw = tuple(i for i in Bnew.codomainind() if Bnew.space[i].isdual)
        _transform_into(Bnew, B, pB, tw)
    else:
        Bnew = permute(B, *pB)

    ipAB = oindABinC(pAB, (Anew.codomainind(), Anew.domainind()), (Bnew.codomainind(), Bnew.domainind()))
    copyC = not _has_shared_permute(C, ipAB)
    if copyC:
        Cnew = TensorMap.empty(C.space.permute(*ipAB), TC)
        mul_(Cnew, Anew, Bnew, 1.0, 0.0, precision)
        tensoradd_(C, Cnew, pAB, False, alpha, beta)
    else:
        Cnew = permute(C, *ipAB)
        mul_(Cnew, Anew, Bnew, alpha, beta, precision)
    return C


def _transform_into(dst, src, p, twist_legs):
    p1, p2 = tuple(p[0]), tuple(p[1])
    if isinstance(src, AdjointTensorMap) or twist_legs or not _is_identity(src, p1, p2):
        add_transform_(dst, src, p1, p2, 1.0, 0.0, dst_twist=twist_legs)
    else:
        add_(dst, src, 1.0, 0.0)


def spacecheck_contract(C, A, pA, conjA, B, pB, conjB, pAB):
    sA = A.space.adjoint() if conjA else A.space
    sB = B.space.adjoint() if conjB else B.space
    qA = adjointtensorindices(A, pA[0]), adjointtensorindices(A, pA[1])
    qB = adjointtensorindices(B, pB[0]), adjointtensorindices(B, pB[1])
    VA = sA.permute(*(qA if conjA else pA))
    VB = sB.permute(*(qB if conjB else pB))
    VC = VA.compose(VB).permute(*pAB)
    if VC != C.space:
        raise SpaceMismatch(f"incompatible spaces for tensorcontract: {VC} ≠ {C.space}")


def tensorcontract_(C, A, pA, conjA, B, pB, conjB, pAB, alpha=1.0, beta=0.0, precision=None):
    """`TO.tensorcontract!` (`tensoroperations.jl:130-161`).  Index tuples are
    0-based; conj operands become lazy adjoints."""
    pAB = _canonicalize(pAB, C)
    spacecheck_contract(C, A, pA, conjA, B, pB, conjB, pAB)
    if conjA:
        A2, pA2 = A.adjoint(), (adjointtensorindices(A, pA[0]), adjointtensorindices(A, pA[1]))
    else:
        A2, pA2 = A, (tuple(pA[0]), tuple(pA[1]))
    if conjB:
        B2, pB2 = B.adjoint(), (adjointtensorindices(B, pB[0]), adjointtensorindices(B, pB[1]))
    else:
        B2, pB2 = B, (tuple(pB[0]), tuple(pB[1]))
    return contract_(C, A2, pA2, B2, pB2, pAB, alpha, beta, precision)


def tensorcontract_space(A, pA, conjA, B, pB, conjB, pAB):
    """`TO.tensorcontract(HomSpace...)` (`homspace.jl:273-295`)."""
    sA = A.space.adjoint() if conjA else A.space
    sB = B.space.adjoint() if conjB else B.space
    qA = (adjointtensorindices(A, pA[0]), adjointtensorindices(A, pA[1])) if conjA else pA
    qB = (adjointtensorindices(B, pB[0]), adjointtensorindices(B, pB[1])) if conjB else pB
    return sA.permute(*qA).compose(sB.permute(*qB)).permute(*pAB)


def tensorcontract(A, pA, conjA, B, pB, conjB, pAB, alpha=1.0, precision=None):
    """Allocating contraction: returns a new TensorMap."""
    VC = tensorcontract_space(A, pA, conjA, B, pB, conjB, pAB)
    C = TensorMap.empty(VC, A.dtype)
    return tensorcontract_(C, A, pA, conjA, B, pB, conjB, pAB, alpha, 0.0, precision)


# ---------------------------------------------------------------------------
# label-based pairwise contraction (the index bookkeeping @tensor lowers to)
# ---------------------------------------------------------------------------
def contraction_indices(la, lb, lc, nout):
    """TensorOperations-style (pA, pB, pAB) for C[lc] = A[la] * B[lb] with the
    first `nout` labels of `lc` in C's codomain.  Open indices keep their order
    of appearance in A and B; contracted ones follow A's order."""
    la, lb, lc = list(la), list(lb), list(lc)
    shared = [x for x in la if x in lb]
    oA = [i for i, x in enumerate(la) if x not in lb]
    cA = [la.index(x) for x in shared]
    cB = [lb.index(x) for x in shared]
    oB = [i for i, x in enumerate(lb) if x not in la]
    ab = [la[i] for i in oA] + [lb[i] for i in oB]
    if sorted(ab) != sorted(lc):
        raise IndexError_(f"open labels {ab} do not match output labels {lc}")
    pos = [ab.index(x) for x in lc]
    return (tuple(oA), tuple(cA)), (tuple(cB), tuple(oB)), (tuple(pos[:nout]), tuple(pos[nout:]))


def contract_labels(A, la, B, lb, lc, nout, conjA=False, conjB=False, out=None, alpha=1.0, beta=0.0,
                    precision=None):
    """C[lc] = α·A[la]·B[lb] (+ β·C): one `tensorcontract!` call."""
    pA, pB, pAB = contraction_indices(la, lb, lc, nout)
    if out is None:
        VC = tensorcontract_space(A, pA, conjA, B, pB, conjB, pAB)
        out = TensorMap.empty(VC, A.dtype)
        beta = 0.0
    return tensorcontract_(out, A, pA, conjA, B, pB, conjB, pAB, alpha, beta, precision)


def autotune_trace(trace, precisions=(DEFAULT_PRECISION,), reps=20):
    """Run the GEMM autotuner on every grouped GEMM of a recorded trace, with the
    recorded operands (outputs are overwritten; re-run the computation after)."""
    lib = _lib.load()
    done = set()
    for op in trace:
        if op[0] != "gemm":
            continue
        _, plan, prec, (tA, _), (tB, _), (tC, _), a, b = op
        for p in precisions:
            if (id(plan), p) in done:
                continue
            done.add((id(plan), p))
            if tC.dtype == torch.float32 and p == _lib.PREC_F64:
                continue
            _lib.check(lib.tkmi_gemm_plan_autotune_dtype(plan.handle, int(p), _DT[tC.dtype], tA.data_ptr(),
                                                         tB.data_ptr(), tC.data_ptr(), int(reps), _stream()),
                       "autotune")
