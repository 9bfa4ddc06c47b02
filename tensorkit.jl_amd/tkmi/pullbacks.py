"""In-place pullbacks shared by the mutating AD front-ends.

Each function ACCUMULATES the cotangent contribution into existing tangent storage
(ΔA += …), the convention of the reference's core pullbacks that its Mooncake and
Enzyme extensions call:

- `permute_pullback_dA`, `transpose_pullback_dA`, `braid_pb`
      src/pullbacks/indexmanipulations.jl:1-31 — ΔA += α·permute(ΔC, p⁻¹).  Run as
      the adjoint transform of the forward plan (`tkmi_transform_adjoint_exec`, β = 1),
      which for TensorKit's inner product Σ_c dim(c)·⟨x_c, y_c⟩ is the inverse
      permutation.
- `blas_contract_pullback_dA`, `blas_contract_pullback_dB`
      src/pullbacks/tensoroperations.jl:1-48 with `project_contract!`
      (src/auxiliary/ad.jl:11-21): ΔA += α·ΔC ⋆ twist(B)ᴴ, ΔB += α·twist(A)ᴴ ⋆ ΔC as
      `tensorcontract!` calls with β = 1 (permutes + grouped GEMM).
- `trace_permute_pullback_dA`
      src/pullbacks/tensoroperations.jl:50-64: ΔA += α·(ΔC ⊗ twisted identity on the
      traced legs), here one generic-transform launch of the trace's adjoint table
      (`treetransformers.trace_adjoint_table`).
- `project_mul`  (src/auxiliary/ad.jl:1-9): C += α·A·B, one grouped GEMM.
- `pullback_dalpha`, `pullback_dbeta` (src/auxiliary/ad.jl:23-24): ⟨AB, ΔC⟩, ⟨C, ΔC⟩
  with `inner` (device dot products).
- `pullback_dC` (ext/TensorKitMooncakeExt/linalg.jl:3): ΔC ← β·ΔC.

Data are real (float64/float32), so conj(α) = α and the `project` steps never need to
take a real part.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import tensors as T
from . import treetransformers as TT


def _invperm(p):
    inv = [0] * len(p)
    for i, x in enumerate(p):
        inv[x] = i
    return tuple(inv)


def _repartition(p, n):
    return tuple(p[:n]), tuple(p[n:])


def _plain(t, what):
    if not isinstance(t, T.TensorMap):
        raise TypeError(f"{what}: in-place pullbacks take TensorMap tangents and primals")
    return t


def _transform_adjoint(dA, A, dC, C, p1, p2, alpha, levels=None):
    """ΔA += α·M‡·ΔC for the transform M: A → C of `permute!/braid!` (the cached
    forward plan run backwards with β = 1)."""
    _plain(dA, "pullback")
    _plain(dC, "pullback")
    if dA.space != A.space or dC.space != C.space:
        raise T.SpaceMismatch("pullback: tangent spaces differ from the primal spaces")
    if dA.dtype != dC.dtype:
        raise TypeError("pullback: dtype mismatch")
    if T._is_identity(A, tuple(p1), tuple(p2)):
        return T.add_(dA, dC, alpha, 1.0)
    if levels is None:  # the levels `permute!` passes, so the forward plan is shared
        n1 = A.numout
        levels = (tuple(range(n1)), tuple(range(n1, A.numind)))
    plan = TT.treebraider(C.space, A.space, tuple(p1), tuple(p2), levels)
    _lib.check(_lib.load().tkmi_transform_adjoint_exec(plan.handle, T._DT[dC.dtype], dC.data.data_ptr(),
                                                       dA.data.data_ptr(), float(alpha), 1.0, T._stream()),
               "permute pullback")
    return dA


def permute_pullback_dA(dA, A, dC, C, p, alpha):
    """`permute_pullback_dA(ΔA, A, ΔC, C, p, α)`: ΔA += α·permute(ΔC, p⁻¹)."""
    return _transform_adjoint(dA, A, dC, C, p[0], p[1], alpha)


def transpose_pullback_dA(dA, A, dC, C, p, alpha):
    """`transpose_pullback_dA`: the inverse of a planar transpose, ΔA += α·transpose(ΔC, p⁻¹)
    (bosonic sectors, where `transpose` is a cyclic `permute`)."""
    if A.space.unit.braiding_style == T.FERMIONIC:
        raise NotImplementedError("planar transpose of fermionic tensors")
    return _transform_adjoint(dA, A, dC, C, p[0], p[1], alpha)


def braid_pb(dA, A, dC, C, p, levels, alpha):
    """`braid_pb(ΔA, A, ΔC, C, p, levels, α)`: ΔA += α·braid(ΔC, p⁻¹, permuted levels);
    with symmetric braiding the levels do not change the coefficients."""
    n1 = A.numout
    lv = (tuple(levels[:n1]), tuple(levels[n1:]))
    return _transform_adjoint(dA, A, dC, C, p[0], p[1], alpha, levels=lv)


def _twist_copy(t, legs):
    """twist(t, legs; copy = false): t itself when nothing changes."""
    legs = tuple(legs)
    if t.space.unit.braiding_style != T.FERMIONIC or not legs:
        return t
    out = t.copy()
    T.twist_(out, legs)
    return out


def blas_contract_pullback_dA(dA, dC, A, pA, B, pB, pAB, alpha, precision=None):
    """`blas_contract_pullback_ΔA!`: ΔA += α·ΔC[pΔC] ⋆ conj(twist(B))[reverse(pB)] → ipA."""
    _plain(dA, "blas_contract pullback")
    ipAB = _invperm(tuple(pAB[0]) + tuple(pAB[1]))
    pdC = _repartition(ipAB, len(pA[0]))
    ipA = _repartition(_invperm(tuple(pA[0]) + tuple(pA[1])), A.numout)
    tB = _twist_copy(B, tuple(i for i in pB[0] if not B.space[i].isdual) +
                     tuple(i for i in pB[1] if B.space[i].isdual))
    T.tensorcontract_(dA, dC, pdC, False, tB, (tuple(pB[1]), tuple(pB[0])), True, ipA, alpha, 1.0, precision)
    return dA


def blas_contract_pullback_dB(dB, dC, A, pA, B, pB, pAB, alpha, precision=None):
    """`blas_contract_pullback_ΔB!`: ΔB += α·conj(twist(A))[reverse(pA)] ⋆ ΔC[pΔC] → ipB."""
    _plain(dB, "blas_contract pullback")
    ipAB = _invperm(tuple(pAB[0]) + tuple(pAB[1]))
    pdC = _repartition(ipAB, len(pA[0]))
    ipB = _repartition(_invperm(tuple(pB[0]) + tuple(pB[1])), B.numout)
    tA = _twist_copy(A, tuple(i for i in pA[0] if A.space[i].isdual) +
                     tuple(i for i in pA[1] if not A.space[i].isdual))
    T.tensorcontract_(dB, tA, (tuple(pA[1]), tuple(pA[0])), True, dC, pdC, False, ipB, alpha, 1.0, precision)
    return dB


def trace_permute_pullback_dA(dA, dC, A, p, q, alpha):
    """`trace_permute_pullback_ΔA!`: ΔA += α·T‡·ΔC for the trace T of
    `trace_permute!(C, A, p, q)` — every traced-diagonal slice of ΔA gets the ΔC
    sub-blocks it fed, weighted by coefficient·dim(c_C)/dim(c_A)."""
    _plain(dA, "trace pullback")
    _plain(dC, "trace pullback")
    (p1, p2), (q1, q2) = (tuple(p[0]), tuple(p[1])), (tuple(q[0]), tuple(q[1]))
    if dA.space != A.space or dC.space != A.space.select(p1, p2):
        raise T.SpaceMismatch("trace pullback: tangent spaces differ from the primal spaces")
    if dA.dtype != dC.dtype:
        raise TypeError("trace pullback: dtype mismatch")
    plan = TT.trace_adjointer(dC.space, A.space, p1, p2, q1, q2)
    _lib.check(_lib.load().tkmi_transform_exec(plan.handle, T._DT[dC.dtype], dC.data.data_ptr(), dA.data.data_ptr(),
                                               float(alpha), 1.0, T._stream()), "trace pullback")
    return dA


def tr_pullback_dA(dA, dtr):
    """`tr` pullback (ext/TensorKitMooncakeExt/linalg.jl:87-98): the diagonal of every
    block of ΔA gets += Δtr (one transform launch from a one-element source)."""
    _plain(dA, "tr pullback")
    W = dA.space
    if W.domain != W.codomain:
        raise T.SpaceMismatch("Trace of a tensor only exist when domain == codomain")
    key = ("tr_adj", W)
    plan = TT._TRANSFORMER_CACHE.get(key)
    if plan is None:
        rows = []
        for c, (d1, d2), off in W.degeneracystructure()[1]:
            if d1 == 0:
                continue
            rows.append((np.ones((d1, 1)), (), [[off + i * (d1 + 1)] for i in range(d1)], [[0]]))
        plan = TT.plan_from_packed("tr_adj", (), TT.pack_rows(rows), W.dim(), 1, table=rows)
        TT._TRANSFORMER_CACHE.put(key, plan)
    src = torch.full((1,), float(dtr), dtype=dA.dtype, device=T._device())
    _lib.check(_lib.load().tkmi_transform_exec(plan.handle, T._DT[dA.dtype], src.data_ptr(), dA.data.data_ptr(),
                                               1.0, 1.0, T._stream()), "tr pullback")
    return dA


def project_mul(C, A, B, alpha, beta=1.0, precision=None):
    """`project_mul!(C, A, B, α, β)`: C = β·C + α·A·B (real data: no projection)."""
    return T.mul_(C, A, B, alpha, beta, precision)


def pullback_dalpha(dC, AB):
    """∂/∂α of ⟨ΔC, α·AB + β·C⟩: inner(AB, ΔC)."""
    from .chainrules import inner
    return inner(AB, dC)


def pullback_dbeta(dC, C):
    """∂/∂β of ⟨ΔC, α·AB + β·C⟩: inner(C, ΔC) with the C the call overwrote."""
    from .chainrules import inner
    return inner(C, dC)


def pullback_dC(dC, beta):
    """ΔC ← β·ΔC: the part of ΔC that flows to the overwritten C."""
    T.scale_(dC, beta)
    return dC
