"""Mutating-AD rules for the hot path: the Mooncake front-end (`rrule`, `frule` — the
reference's `Mooncake.rrule!!` / `Mooncake.frule!!`) and the Enzyme front-end
(`augmented_primal`, `reverse`, `forward` — `EnzymeRules.*`), over the shared in-place
pullbacks of `pullbacks.py` and the device kernels they launch.

Rules exist for the in-place primals of the path:

  tensors.permute_    ext/TensorKitMooncakeExt/indexmanipulations.jl:1-60,
  tensors.transpose_  ext/TensorKitEnzymeExt/indexmanipulations.jl:1-59
  tensors.braid_      ext/TensorKitMooncakeExt/indexmanipulations.jl:62-116,
                      ext/TensorKitEnzymeExt/indexmanipulations.jl:61-117
  tensors.twist_      ext/TensorKitMooncakeExt/indexmanipulations.jl:118-161
  tensors.blas_contract_   ext/TensorKitMooncakeExt/tensoroperations.jl:9-100,
                           ext/TensorKitEnzymeExt/tensoroperations.jl:10-118
  tensors.trace_permute_   ext/TensorKitMooncakeExt/tensoroperations.jl:104-187
  tensors.mul_        ext/TensorKitMooncakeExt/linalg.jl:5-62, ext/TensorKitEnzymeExt/linalg.jl:4-91
  tensors.tr          ext/TensorKitMooncakeExt/linalg.jl:85-105, ext/TensorKitEnzymeExt/linalg.jl:93-153
  tensors.flip        ext/TensorKitEnzymeExt/indexmanipulations.jl:119-170 (out of place)

Argument annotations (the two packages' vocabulary):
  Duplicated(x, dx)  a tensor with its tangent storage (Mooncake CoDual / Dual, Enzyme
                     Duplicated); in forward mode also a scalar with its tangent dx
  Active(x)          a scalar whose derivative the reverse pass returns
  Const(x)           no derivative (Mooncake NoTangent / NoFData, Enzyme Const)
A bare tensor counts as Const; a bare float counts as Active in reverse mode (Mooncake
gives every Float64 a tangent) and as Const in forward mode.

Reverse mode follows the reference step for step: the primal overwrites its output after
caching it; when α carries a derivative the product is formed separately (C = β·C +
α·AB) and kept for Δα = ⟨AB, ΔC⟩; the reverse pass restores the output (Mooncake only),
accumulates ΔA/ΔB (+=), returns Δα and Δβ = ⟨C_old, ΔC⟩, and finally scales ΔC by β.
Forward mode propagates ΔC′ = β·ΔC + Δβ·C + Δα·AB + α·(ΔA·B + A·ΔB) before the primal
call overwrites C.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from . import tensors as T
from . import pullbacks as PB


@dataclass
class Duplicated:
    val: Any
    dval: Any


@dataclass
class Active:
    val: float


@dataclass
class Const:
    val: Any


def _v(x):
    return x.val if isinstance(x, (Duplicated, Active, Const)) else x


def _tangent(x):
    """Tangent storage of a tensor argument (None when it has none)."""
    return x.dval if isinstance(x, Duplicated) else None


def _rev_active(x):
    """A scalar argument whose derivative the reverse pass returns."""
    return isinstance(x, Active) or (not isinstance(x, (Const, Duplicated)) and isinstance(x, (int, float)))


def _fwd_tangent(x):
    """The forward-mode tangent of a scalar argument (None: no tangent)."""
    return float(x.dval) if isinstance(x, Duplicated) else None


# --------------------------------------------------------------------------- rules
class _TransformRule:
    """permute! / transpose! / braid!: args (C, A, p1, p2, [levels,] α, β)."""

    def __init__(self, kind):
        self.kind = kind
        self.nidx = 3 if kind == "braid" else 2  # index-tuple arguments after A

    def _split(self, args):
        C, A = args[0], args[1]
        idx = [_v(a) for a in args[2:2 + self.nidx]]
        return C, A, idx, args[2 + self.nidx], args[3 + self.nidx]

    def _new(self, A, idx):
        a = _v(A)
        if self.kind == "braid":
            return T.braid(a, idx[0], idx[1], idx[2], copy=True)
        if self.kind == "transpose":
            return T.transpose(a, idx[0], idx[1])
        return T.permute(a, idx[0], idx[1], copy=True)

    def _primal(self, C, A, idx, a, b):
        fn = {"permute": T.permute_, "transpose": T.transpose_, "braid": T.braid_}[self.kind]
        return fn(C, A, *idx, a, b)

    def augmented(self, args):
        C, A, idx, al, be = self._split(args)
        c, a, b = _v(C), _v(al), _v(be)
        cache_C = c.copy()
        if _rev_active(al):
            Ap = self._new(A, idx)
            T.add_(c, Ap, a, b)
        else:
            Ap = None
            self._primal(c, _v(A), idx, a, b)
        return c, (cache_C, Ap)

    def reverse(self, args, cache, dret, restore):
        C, A, idx, al, be = self._split(args)
        cache_C, Ap = cache
        c, dC, dA = _v(C), _tangent(C), _tangent(A)
        if restore:
            T.add_(c, cache_C, 1.0, 0.0)
        out = [None] * len(args)
        if dC is None:
            return tuple(out)
        a = _v(al)
        p = (idx[0], idx[1])
        if dA is not None:
            if self.kind == "braid":
                PB.braid_pb(dA, _v(A), dC, c, p, idx[2], a)
            elif self.kind == "transpose":
                PB.transpose_pullback_dA(dA, _v(A), dC, c, p, a)
            else:
                PB.permute_pullback_dA(dA, _v(A), dC, c, p, a)
        if _rev_active(al):
            out[len(args) - 2] = PB.pullback_dalpha(dC, Ap)
        if _rev_active(be):
            out[len(args) - 1] = PB.pullback_dbeta(dC, cache_C)
        PB.pullback_dC(dC, _v(be))
        return tuple(out)

    def forward(self, args):
        C, A, idx, al, be = self._split(args)
        c, dC, dA = _v(C), _tangent(C), _tangent(A)
        a, b = _v(al), _v(be)
        da, db = _fwd_tangent(al), _fwd_tangent(be)
        if dC is not None:
            if db is None:
                T.scale_(dC, b)
            else:
                T.add_(dC, c, db, b)
            if da is not None:
                self._primal(dC, _v(A), idx, da, 1.0)
            if dA is not None:
                self._primal(dC, dA, idx, a, 1.0)
        self._primal(c, _v(A), idx, a, b)
        return c, dC


class _TwistRule:
    """twist!(t, inds): args (t, inds).  The twists of the path's sectors are ±1, so the
    inverse twist is the twist itself."""

    def augmented(self, args):
        t = _v(args[0])
        cache = t.copy()
        T.twist_(t, tuple(_v(args[1])))
        return t, cache

    def reverse(self, args, cache, dret, restore):
        t, dt = _v(args[0]), _tangent(args[0])
        if restore:
            T.add_(t, cache, 1.0, 0.0)
        if dt is not None:
            T.twist_(dt, tuple(_v(args[1])))
        return (None, None)

    def forward(self, args):
        t, dt = _v(args[0]), _tangent(args[0])
        T.twist_(t, tuple(_v(args[1])))
        if dt is not None:
            T.twist_(dt, tuple(_v(args[1])))
        return t, dt


class _ContractRule:
    """blas_contract!: args (C, A, pA, B, pB, pAB, α, β[, precision])."""

    def _unpack(self, args):
        C, A, pA, B, pB, pAB, al, be = args[:8]
        prec = _v(args[8]) if len(args) > 8 else None
        return C, A, _v(pA), B, _v(pB), _v(pAB), al, be, prec

    def augmented(self, args):
        C, A, pA, B, pB, pAB, al, be, prec = self._unpack(args)
        c, a, b = _v(C), _v(al), _v(be)
        cache_C = c.copy()
        if _rev_active(al):
            AB = T.tensorcontract(_v(A), pA, False, _v(B), pB, False, pAB, 1.0, prec)
            T.add_(c, AB, a, b)
        else:
            AB = None
            T.blas_contract_(c, _v(A), pA, _v(B), pB, pAB, a, b, prec)
        return c, (cache_C, AB)

    def reverse(self, args, cache, dret, restore):
        C, A, pA, B, pB, pAB, al, be, prec = self._unpack(args)
        cache_C, AB = cache
        c, dC = _v(C), _tangent(C)
        if restore:
            T.add_(c, cache_C, 1.0, 0.0)
        out = [None] * len(args)
        if dC is None:
            return tuple(out)
        a = _v(al)
        if _tangent(A) is not None:
            PB.blas_contract_pullback_dA(_tangent(A), dC, _v(A), pA, _v(B), pB, pAB, a, prec)
        if _tangent(B) is not None:
            PB.blas_contract_pullback_dB(_tangent(B), dC, _v(A), pA, _v(B), pB, pAB, a, prec)
        if _rev_active(al):
            out[6] = PB.pullback_dalpha(dC, AB)
        if _rev_active(be):
            out[7] = PB.pullback_dbeta(dC, cache_C)
        PB.pullback_dC(dC, _v(be))
        return tuple(out)

    def forward(self, args):
        C, A, pA, B, pB, pAB, al, be, prec = self._unpack(args)
        c, dC = _v(C), _tangent(C)
        a, b = _v(al), _v(be)
        da, db = _fwd_tangent(al), _fwd_tangent(be)
        A_, B_ = _v(A), _v(B)
        if dC is not None:
            if db is None:
                T.scale_(dC, b)
            else:
                T.add_(dC, c, db, b)
            if da is not None:
                T.blas_contract_(dC, A_, pA, B_, pB, pAB, da, 1.0, prec)
            if _tangent(A) is not None:
                T.blas_contract_(dC, _tangent(A), pA, B_, pB, pAB, a, 1.0, prec)
            if _tangent(B) is not None:
                T.blas_contract_(dC, A_, pA, _tangent(B), pB, pAB, a, 1.0, prec)
        T.blas_contract_(c, A_, pA, B_, pB, pAB, a, b, prec)
        return c, dC


class _TraceRule:
    """trace_permute!: args (C, A, p, q, α, β)."""

    def augmented(self, args):
        C, A, p, q, al, be = args
        c, p, q, a, b = _v(C), _v(p), _v(q), _v(al), _v(be)
        cache_C = c.copy()
        if _rev_active(al):
            At = T.TensorMap.empty(c.space, c.dtype)
            T.trace_permute_(At, _v(A), p, q, 1.0, 0.0)
            T.add_(c, At, a, b)
        else:
            At = None
            T.trace_permute_(c, _v(A), p, q, a, b)
        return c, (cache_C, At)

    def reverse(self, args, cache, dret, restore):
        C, A, p, q, al, be = args
        cache_C, At = cache
        c, dC = _v(C), _tangent(C)
        if restore:
            T.add_(c, cache_C, 1.0, 0.0)
        out = [None] * len(args)
        if dC is None:
            return tuple(out)
        if _tangent(A) is not None:
            PB.trace_permute_pullback_dA(_tangent(A), dC, _v(A), _v(p), _v(q), _v(al))
        if _rev_active(al):
            out[4] = PB.pullback_dalpha(dC, At)
        if _rev_active(be):
            out[5] = PB.pullback_dbeta(dC, cache_C)
        PB.pullback_dC(dC, _v(be))
        return tuple(out)

    def forward(self, args):
        C, A, p, q, al, be = args
        c, dC, p, q = _v(C), _tangent(C), _v(p), _v(q)
        a, b = _v(al), _v(be)
        da, db = _fwd_tangent(al), _fwd_tangent(be)
        if dC is not None:
            if db is None:
                T.scale_(dC, b)
            else:
                T.add_(dC, c, db, b)
            if da is not None:
                T.trace_permute_(dC, _v(A), p, q, da, 1.0)
            if _tangent(A) is not None:
                T.trace_permute_(dC, _tangent(A), p, q, a, 1.0)
        T.trace_permute_(c, _v(A), p, q, a, b)
        return c, dC


class _MulRule:
    """mul!: args (C, A, B, α, β[, precision])."""

    def augmented(self, args):
        C, A, B, al, be = args[:5]
        prec = _v(args[5]) if len(args) > 5 else None
        c, a, b = _v(C), _v(al), _v(be)
        cache_C = c.copy()
        if _rev_active(al):
            AB = T.TensorMap.empty(c.space, c.dtype)
            T.mul_(AB, _v(A), _v(B), 1.0, 0.0, prec)
            T.add_(c, AB, a, b)
        else:
            AB = None
            T.mul_(c, _v(A), _v(B), a, b, prec)
        return c, (cache_C, AB)

    def reverse(self, args, cache, dret, restore):
        C, A, B, al, be = args[:5]
        prec = _v(args[5]) if len(args) > 5 else None
        cache_C, AB = cache
        c, dC = _v(C), _tangent(C)
        if restore:
            T.add_(c, cache_C, 1.0, 0.0)
        out = [None] * len(args)
        if dC is None:
            return tuple(out)
        a = _v(al)
        if _tangent(A) is not None:
            PB.project_mul(_tangent(A), dC, _v(B).adjoint(), a, 1.0, prec)
        if _tangent(B) is not None:
            PB.project_mul(_tangent(B), _v(A).adjoint(), dC, a, 1.0, prec)
        if _rev_active(al):
            out[3] = PB.pullback_dalpha(dC, AB)
        if _rev_active(be):
            out[4] = PB.pullback_dbeta(dC, cache_C)
        PB.pullback_dC(dC, _v(be))
        return tuple(out)

    def forward(self, args):
        C, A, B, al, be = args[:5]
        prec = _v(args[5]) if len(args) > 5 else None
        c, dC = _v(C), _tangent(C)
        a, b = _v(al), _v(be)
        da, db = _fwd_tangent(al), _fwd_tangent(be)
        if dC is not None:
            T.scale_(dC, b)
            if db is not None:
                T.add_(dC, c, db, 1.0)
            if da is not None:
                PB.project_mul(dC, _v(A), _v(B), da, 1.0, prec)
            if _tangent(A) is not None:
                PB.project_mul(dC, _tangent(A), _v(B), a, 1.0, prec)
            if _tangent(B) is not None:
                PB.project_mul(dC, _v(A), _tangent(B), a, 1.0, prec)
        T.mul_(c, _v(A), _v(B), a, b, prec)
        return c, dC


class _TrRule:
    """tr(A): args (A,); the output is a scalar and the reverse pass takes its cotangent."""

    def augmented(self, args):
        return T.tr(_v(args[0])), None

    def reverse(self, args, cache, dret, restore):
        dA = _tangent(args[0])
        if dA is not None and dret is not None:
            PB.tr_pullback_dA(dA, dret)
        return (None,)

    def forward(self, args):
        dA = _tangent(args[0])
        return T.tr(_v(args[0])), (T.tr(dA) if dA is not None else None)


class _FlipRule:
    """flip(t, inds[, inv]): out of place.  The tape is the output's shadow, created zero,
    into which the caller accumulates the output's cotangent (Enzyme); a cotangent given
    to the reverse pass directly (Mooncake's pullback argument) takes its place.  The
    reverse pass adds flip(Δt′, inds; inv = !inv) to Δt: flip is unitary for the inner
    product Σ_c dim(c)·⟨b₁, b₂⟩, so its adjoint is its inverse, the opposite flip."""

    @staticmethod
    def _split(args):
        inds = _v(args[1])
        inds = (inds,) if isinstance(inds, int) else tuple(inds)
        return _v(args[0]), inds, bool(_v(args[2])) if len(args) > 2 else False

    def augmented(self, args):
        t, inds, inv = self._split(args)
        out = T.flip(t, inds, inv)
        return out, T.TensorMap.zeros(out.space, out.dtype)

    def reverse(self, args, cache, dret, restore):
        t, inds, inv = self._split(args)
        dt = _tangent(args[0])
        cot = dret if dret is not None else cache
        if dt is not None and cot is not None:
            T.add_(dt, T.flip(cot, inds, not inv), 1.0, 1.0)
        return (None,) * len(args)

    def forward(self, args):
        t, inds, inv = self._split(args)
        out = T.flip(t, inds, inv)
        dt = _tangent(args[0])
        return out, (T.flip(dt, inds, inv) if dt is not None else T.TensorMap.zeros(out.space, out.dtype))


_RULES = {
    T.permute_: _TransformRule("permute"),
    T.transpose_: _TransformRule("transpose"),
    T.braid_: _TransformRule("braid"),
    T.twist_: _TwistRule(),
    T.blas_contract_: _ContractRule(),
    T.trace_permute_: _TraceRule(),
    T.mul_: _MulRule(),
    T.tr: _TrRule(),
    T.flip: _FlipRule(),
}


def _rule(f):
    try:
        return _RULES[f]
    except KeyError:
        raise NotImplementedError(f"no in-place AD rule for {getattr(f, '__name__', f)}") from None


# --------------------------------------------------------------------------- Mooncake
def rrule(f, *args):
    """`Mooncake.rrule!!(f, args...)`: runs the primal (overwriting the output argument)
    and returns (output, pullback).  `pullback(Δout=None)` restores the overwritten
    output, accumulates into the tangents of the Duplicated arguments, scales the
    output's tangent by β and returns one entry per argument: the derivative of each
    Active scalar, None elsewhere."""
    rule = _rule(f)
    out, cache = rule.augmented(args)

    def pullback(dout=None):
        return rule.reverse(args, cache, dout, True)

    return out, pullback


def frule(f, *args):
    """`Mooncake.frule!!(f, args...)`: (output, output tangent), propagating the
    tangents of the Duplicated arguments."""
    return _rule(f).forward(args)


# --------------------------------------------------------------------------- Enzyme
def augmented_primal(f, *args):
    """`EnzymeRules.augmented_primal`: runs the primal, returns (output, tape)."""
    return _rule(f).augmented(args)


def reverse(f, tape, *args, dret=None):
    """`EnzymeRules.reverse`: accumulates into the Duplicated tangents and returns the
    derivatives of the Active scalars (None elsewhere).  The primal output is left as
    the primal call wrote it."""
    return _rule(f).reverse(args, tape, dret, False)


def forward(f, *args):
    """`EnzymeRules.forward`: (output, output tangent)."""
    return _rule(f).forward(args)
