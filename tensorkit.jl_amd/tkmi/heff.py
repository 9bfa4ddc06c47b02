"""Two-site effective-Hamiltonian application: the chain of contractions of one
DMRG sweep step, built from the TensorKit-style API in `tensors`.

    y[a s1 s2; b] = L[a w1; a'] · x[a' s1' s2'; b'] · M1[w1 s1; s1' w2]
                    · M2[w2 s2; s2' w3] · R[b' w3; b]

spaces: x ∈ V⊗P⊗P ← V, L ∈ V⊗W' ← V, M1, M2 ∈ W⊗P ← P⊗W, R ∈ V⊗W ← V.
Pairwise order ((((L·x)·M1)·M2)·R); each step is one `tensorcontract!`
(permutes of the operands + one grouped GEMM launch + permute of the result
when needed).  Intermediates are allocated on first use and the whole chain
can be captured in a HIP graph (torch.cuda.CUDAGraph on ROCm).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import tensors as T
from .spaces import HomSpace, ProductSpace, U1Space

# (A-labels, B-labels, C-labels, C codomain size) of each pairwise step
STEPS = [
    (("a", "w1", "ap"), ("ap", "s1p", "s2p", "bp"), ("a", "w1", "s1p", "s2p", "bp"), 2),
    (("a", "w1", "s1p", "s2p", "bp"), ("w1", "s1", "s1p", "w2"), ("a", "s1", "s2p", "bp", "w2"), 2),
    (("a", "s1", "s2p", "bp", "w2"), ("w2", "s2", "s2p", "w3"), ("a", "s1", "s2", "bp", "w3"), 3),
    (("a", "s1", "s2", "bp", "w3"), ("bp", "w3", "b"), ("a", "s1", "s2", "b"), 3),
]


def gaussian_degeneracies(qs, total, width2, mode="round"):
    """d_q ∝ exp(-q²/width2) normalised to `total`; mode 'round' (Julia round,
    ties to even) or 'lr' (largest remainder, exact total)."""
    w = np.array([math.exp(-q * q / width2) for q in qs])
    x = total * w / w.sum()
    if mode == "round":
        d = [int(round(v)) for v in x]
    else:
        fl = np.floor(x).astype(int)
        rem = total - int(fl.sum())
        order = np.argsort(-(x - fl), kind="stable")
        fl[order[:rem]] += 1
        d = [int(v) for v in fl]
    return dict(zip(qs, d))


def two_site_spaces(D=2048, qmax=8, mode="round"):
    deg = gaussian_degeneracies(list(range(-qmax, qmax + 1)), D, 16.0, mode)
    V = U1Space(*[(q, d) for q, d in deg.items()])
    P = U1Space((-0.5, 1), (0.5, 1))
    W = U1Space((0, 3), (1, 1), (-1, 1))
    return V, P, W


class TwoSiteHeff:
    """Operands and the contraction chain of one H_eff application."""

    def __init__(self, V, P, W, rng, dtype=torch.float64, precision=None):
        self.V, self.P, self.W = V, P, W
        self.spaces = {
            "x": HomSpace(ProductSpace(V, P, P), ProductSpace(V)),
            "L": HomSpace(ProductSpace(V, W.dual()), ProductSpace(V)),
            "M": HomSpace(ProductSpace(W, P), ProductSpace(P, W)),
            "R": HomSpace(ProductSpace(V, W), ProductSpace(V)),
        }
        self.L = T.TensorMap.randn(self.spaces["L"], rng, dtype)
        self.M1 = T.TensorMap.randn(self.spaces["M"], rng, dtype)
        self.M2 = T.TensorMap.randn(self.spaces["M"], rng, dtype)
        self.R = T.TensorMap.randn(self.spaces["R"], rng, dtype)
        self.x = T.TensorMap.randn(self.spaces["x"], rng, dtype)
        self.precision = precision
        self.outs = [None] * len(STEPS)
        self.graph = None

    def operands(self):
        return [self.L, self.M1, self.M2, self.R]

    def astype(self, dtype):
        """The same operator with its operands (and x) converted to `dtype`, each element
        rounded once.  A float32 operator keeps every intermediate of the chain in FP32 (the
        permutes move half the bytes); its grouped GEMMs run in FP32 or 3×BF16."""
        H = TwoSiteHeff.__new__(TwoSiteHeff)
        H.V, H.P, H.W, H.spaces = self.V, self.P, self.W, self.spaces
        for k in ("L", "M1", "M2", "R", "x"):
            t = getattr(self, k)
            setattr(H, k, T.TensorMap(t.data.to(dtype), t.space))
        H.precision = self.precision
        H.outs = [None] * len(STEPS)
        H.graph = None
        return H

    def apply(self, x=None, precision=None):
        """y = H_eff(x); intermediates are reused across calls."""
        x = self.x if x is None else x
        prec = self.precision if precision is None else precision
        cur = None
        for i, (la, lb, lc, nout) in enumerate(STEPS):
            if i == 0:
                A, B = self.L, x
            else:
                A, B = cur, [None, self.M1, self.M2, self.R][i]
            out = self.outs[i]
            if out is None:
                out = T.contract_labels(A, la, B, lb, lc, nout, precision=prec)
                self.outs[i] = out
            else:
                T.contract_labels(A, la, B, lb, lc, nout, out=out, precision=prec)
            cur = out
        return cur

    def gemm_flops(self):
        """Σ_c 2·m·n·k over every grouped GEMM of the chain (plans cached)."""
        return sum(p.flops for p in self._plans())

    def _plans(self):
        return list(T._GEMM_CACHE.d.values())

    def mark_constant_operands(self, precision=None):
        """The environments L and R stay fixed while x changes (a Lanczos or Davidson run):
        mark them as constant operands of the grouped GEMMs that read them, so the 3×BF16
        form splits them into bf16 planes once (tkmi_gemm_plan_set_constant).  Call
        `invalidate_operands` after changing L or R in place."""
        saved = T._TRACE
        T.start_trace()
        try:
            self.apply(precision=precision)
        finally:
            tr = T.stop_trace()
            T._TRACE = saved
        lib = T._lib.load()
        envs = {self.L.data.data_ptr(), self.R.data.data_ptr()}
        self._const_plans = []
        for op in tr:
            if op[0] != "gemm":
                continue
            plan, (ta, _), (tb, _) = op[1], op[3], op[4]
            mask = (1 if ta.data_ptr() in envs else 0) | (2 if tb.data_ptr() in envs else 0)
            if mask:
                T._lib.check(lib.tkmi_gemm_plan_set_constant(plan.handle, mask), "set constant")
                self._const_plans.append(plan)

    def invalidate_operands(self):
        """L or R changed in place: the next 3×BF16 application splits them again."""
        lib = T._lib.load()
        for plan in getattr(self, "_const_plans", []):
            T._lib.check(lib.tkmi_gemm_plan_invalidate(plan.handle), "invalidate")

    def capture(self, precision=None):
        """Capture the whole chain in one HIP graph (intermediates live in the
        graph's private pool).  Returns the output tensor the graph writes.  In the
        3×BF16 form the constant environments are split once, before the capture, and the
        graph splits only the operands that change."""
        if precision == T._lib.PREC_BF16X3:
            self.mark_constant_operands(precision)
        self.apply(precision=precision)  # build plans / warm up (splits the constants)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            saved = self.outs
            self.outs = [None] * len(STEPS)
            # the graph keeps the device addresses of every plan it launches: pin them
            with T.pinning() as pins, torch.cuda.graph(g, stream=s):
                y = self.apply(precision=precision)
            self.graph_pins = pins
            self.graph_outs = self.outs
            self.outs = saved
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        self.graph_y = y
        return y

    def replay(self):
        self.graph.replay()
        return self.graph_y
