"""Reverse-mode rules for the hot path (ChainRulesCore-style `rrule`s).

Each `rrule_*` returns (primal result, pullback); the pullback maps the
cotangent of the result to cotangents of the inputs, using the device kernels:

- permute:        adjoint transform (inverse permutation, transposed recoupling
                  matrices) of the same cached plan — `tkmi_transform_adjoint_exec`
                  (`rrule(permute)`, ext/TensorKitChainRulesCoreExt/linalg.jl:62-70;
                  `permute_pullback_dA`, src/pullbacks/indexmanipulations.jl:1-17)
- flip:           the opposite flip, flip(Δt′, I; inv = !inv)
                  (ext/TensorKitChainRulesCoreExt/linalg.jl:99-103)
- tensoradd!:     ext/TensorKitChainRulesCoreExt/tensoroperations.jl:8-58
- tensorcontract!: ext/TensorKitChainRulesCoreExt/tensoroperations.jl:60-150
                  (dA = ΔC ⋆ twist(B), dB = twist(A) ⋆ ΔC as two more
                  tensorcontract! calls: permutes + grouped GEMMs)

Inner products use TensorKit's `inner` (Σ_c dim(c)·⟨b₁,b₂⟩), evaluated on the
device with the deterministic two-pass `tkmi_dot` reduction.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import tensors as T


def _invperm(p):
    inv = [0] * len(p)
    for i, x in enumerate(p):
        inv[x] = i
    return tuple(inv)


def _repartition(p, n):
    return tuple(p[:n]), tuple(p[n:])


def _twist_legs(t, legs):
    """twist(t, legs; copy=false): returns t itself when nothing is twisted."""
    if t.space.unit.braiding_style != "fermionic":
        return t
    legs = tuple(legs)
    if not legs:
        return t
    out = t.copy() if isinstance(t, T.TensorMap) else t
    T.twist_(out, legs)
    return out


def rrule_permute(t, p1, p2):
    p1, p2 = tuple(p1), tuple(p2)
    y = T.permute(t, p1, p2, copy=True)
    plan = T.treebraider(y.space, t.space, p1, p2)

    def pullback(dy):
        dt = T.TensorMap.empty(t.space, dy.dtype)
        lib = _lib.load()
        _lib.check(lib.tkmi_transform_adjoint_exec(plan.handle, T._DT[dy.dtype], dy.data.data_ptr(),
                                                   dt.data.data_ptr(), 1.0, 0.0, T._stream()), "permute pullback")
        return dt

    return y, pullback


def rrule_flip(t, inds, inv=False):
    y = T.flip(t, inds, inv)

    def pullback(dy):
        return T.flip(dy, inds, not inv)

    return y, pullback


def rrule_tensoradd(C, A, pA, conjA, alpha, beta):
    Cp = C.copy()
    T.tensoradd_(Cp, A, pA, conjA, alpha, beta)

    def pullback(dC):
        dCc = None
        if beta != 0:
            dCc = dC.copy()
            T.scale_(dCc, beta)
        ipA = _invperm(tuple(pA[0]) + tuple(pA[1]))
        pdA = _repartition(ipA, A.numout)
        VdA = dC.space.adjoint().permute(*(_adj_idx(dC, pdA))) if conjA else dC.space.permute(*pdA)
        dA = T.TensorMap.empty(VdA, dC.dtype)
        T.tensoradd_(dA, dC, pdA, conjA, alpha, 0.0)
        return dCc, dA

    return Cp, pullback


def _adj_idx(t, p):
    return T.adjointtensorindices(t, p[0]), T.adjointtensorindices(t, p[1])


def rrule_tensorcontract(C, A, pA, conjA, B, pB, conjB, pAB, alpha=1.0, beta=0.0, precision=None):
    pA = (tuple(pA[0]), tuple(pA[1]))
    pB = (tuple(pB[0]), tuple(pB[1]))
    pAB = (tuple(pAB[0]), tuple(pAB[1]))
    Cp = C.copy()
    T.tensorcontract_(Cp, A, pA, conjA, B, pB, conjB, pAB, alpha, beta, precision)

    def pullback(dC):
        ipAB = _invperm(pAB[0] + pAB[1])
        pdC = _repartition(ipAB, len(pA[0]))
        dCc = None
        if beta != 0:
            dCc = dC.copy()
            T.scale_(dCc, beta)
        # dA
        ipA = _repartition(_invperm(pA[0] + pA[1]), A.numout)
        conjB2 = conjB if conjA else (not conjB)
        tB = _twist_legs(B, tuple(x for x in pB[0] if not B.space[x].isdual)
                         + tuple(x for x in pB[1] if B.space[x].isdual))
        revB = (pB[1], pB[0])
        VA = T.tensorcontract_space(dC, pdC, conjA, tB, revB, conjB2, ipA)
        dA = T.TensorMap.empty(VA, dC.dtype)
        T.tensorcontract_(dA, dC, pdC, conjA, tB, revB, conjB2, ipA, alpha, 0.0, precision)
        # dB
        ipB = _repartition(_invperm(pB[0] + pB[1]), B.numout)
        pdC2 = _repartition(ipAB, len(pA[0]))
        conjA2 = conjA if conjB else (not conjA)
        tA = _twist_legs(A, tuple(x for x in pA[0] if A.space[x].isdual)
                         + tuple(x for x in pA[1] if not A.space[x].isdual))
        revA = (pA[1], pA[0])
        VB = T.tensorcontract_space(tA, revA, conjA2, dC, pdC2, conjB, ipB)
        dB = T.TensorMap.empty(VB, dC.dtype)
        T.tensorcontract_(dB, tA, revA, conjA2, dC, pdC2, conjB, ipB, alpha, 0.0, precision)
        return dCc, dA, dB

    return Cp, pullback


_WORK = {}


def inner(a, b):
    """TensorKit `inner(a, b)` = Σ_c dim(c)·⟨block_a(c), block_b(c)⟩ (real data),
    one device dot product per group of blocks sharing dim(c)."""
    if a.space != b.space:
        raise T.SpaceMismatch("inner: spaces differ")
    lib = _lib.load()
    work = _WORK.get("w")
    if work is None:
        work = torch.empty(1025, dtype=torch.float64, device=T._device())
        _WORK["w"] = work
    blocks = a.space.degeneracystructure()[1]
    total = 0.0
    esz = a.data.element_size()
    if all(c.dim() == 1 for c, _, _ in blocks):
        ranges = [(1, 0, a.data.numel())]
    else:
        ranges = [(c.dim(), off, d1 * d2) for c, (d1, d2), off in blocks]
    for w, off, n in ranges:
        if n == 0:
            continue
        _lib.check(lib.tkmi_dot(T._DT[a.dtype], a.data.data_ptr() + off * esz, b.data.data_ptr() + off * esz, n,
                                work.data_ptr(), T._stream()), "inner")
        total += w * float(work[0].item())
    return total


def rrule_contract_labels(A, la, B, lb, lc, nout, conjA=False, conjB=False, precision=None):
    """rrule of C[lc] = A[la]·B[lb] (one `tensorcontract!`), label form."""
    pA, pB, pAB = T.contraction_indices(la, lb, lc, nout)
    VC = T.tensorcontract_space(A, pA, conjA, B, pB, conjB, pAB)
    C0 = T.TensorMap.zeros(VC, A.dtype)
    C, pb = rrule_tensorcontract(C0, A, pA, conjA, B, pB, conjB, pAB, 1.0, 0.0, precision)

    def pullback(dC):
        _, dA, dB = pb(dC)
        return dA, dB

    return C, pullback


def add(a, b, alpha=1.0, beta=1.0):
    """New tensor α·b + β·a (same space)."""
    out = a.copy()
    T.add_(out, b, alpha, beta)
    return out
