"""CPU: the small-operand contraction table (fusedcontract) equals the unfused
blas_contract! path — operand permutes (with the fermionic twists) followed by one GEMM
per coupled sector — evaluated independently with numpy from the permute tables and the
block layouts (src/tensors/tensoroperations.jl:392-462, src/tensors/linalg.jl:330-370)."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space
from tkmi import treetransformers as TT
from tkmi import fusedcontract as FC


class _T:
    """Minimal stand-in for a TensorMap (space only) for the host-side table code."""
    def __init__(self, space):
        self.space = space


def _apply_abelian(tab, x, ydim, n):
    y = np.zeros(ydim)
    for c, d, s in zip(tab["coeffs"], tab["dst"], tab["src"]):
        sz_d, st_d, off_d = d[:n], d[n:2 * n], d[2 * n]
        sz_s, st_s, off_s = s[:n], s[n:2 * n], s[2 * n]
        dv = np.lib.stride_tricks.as_strided(y[off_d:], shape=tuple(sz_d), strides=tuple(8 * v for v in st_d))
        dv[...] = c * _permuted(x, sz_s, st_s, off_s, tab["perm"])
    return y


def _permuted(x, sz, st, off, perm):
    v = np.lib.stride_tricks.as_strided(x[off:], shape=tuple(sz), strides=tuple(8 * s for s in st))
    return v.transpose(perm)


def _unfused(C, A, pA, B, pB, pC, twA, twB, a, b, alpha, beta, c0):
    WA2, WB2 = A.permute(*pA), B.permute(*pB)
    tA = TT.abelian_table(WA2, A, pA[0], pA[1], None, twA)
    tA["perm"] = pA[0] + pA[1]
    tB = TT.abelian_table(WB2, B, pB[0], pB[1], None, twB)
    tB["perm"] = pB[0] + pB[1]
    a2 = _apply_abelian(tA, a, WA2.dim(), A.numind)
    b2 = _apply_abelian(tB, b, WB2.dim(), B.numind)
    WC2 = WA2.compose(WB2)
    c2 = np.zeros(WC2.dim())
    LA = {c: (d, off) for c, d, off in WA2.degeneracystructure()[1]}
    LB = {c: (d, off) for c, d, off in WB2.degeneracystructure()[1]}
    for c, (m, n), off in WC2.degeneracystructure()[1]:
        (m_, k), oa = LA[c]
        (k_, n_), ob = LB[c]
        am = a2[oa:oa + m * k].reshape(k, m).T
        bm = b2[ob:ob + k * n].reshape(n, k).T
        c2[off:off + m * n] = (am @ bm).T.reshape(-1)
    tC = TT.abelian_table(C, WC2, pC[0], pC[1])
    tC["perm"] = pC[0] + pC[1]
    return beta * c0 + alpha * _apply_abelian(tC, c2, C.dim(), C.numind)


def _fused(C, A, pA, B, pB, pC, twA, twB, a, b, alpha, beta, c0):
    rows, (coef, idx) = FC.fused_table(_T(C), _T(A), pA, _T(B), pB, pC, twA, twB)
    # the plan's slice rank is that of A′'s codomain, and every slice stays inside its vector
    for R, Cn, szd, dsts, srcs in rows:
        assert len(szd) == len(pA[0])
        for sl, n in [(d, C.dim()) for d in dsts] + [(s_, A.dim()) for s_ in srcs]:
            assert len(sl) == len(szd) + 1
            assert 0 <= sl[-1] and sl[-1] + sum((z - 1) * st for z, st in zip(szd, sl[:-1])) < n
    U = np.where(idx >= 0, coef * b[np.maximum(idx, 0)], 0.0)
    out = beta * c0
    u0 = 0
    for R, Cn, szd, dsts, srcs in rows:
        Um = U[u0:u0 + R * Cn].reshape(Cn, R).T
        u0 += R * Cn
        sv = [np.lib.stride_tricks.as_strided(a[s[-1]:], shape=szd, strides=tuple(8 * v for v in s[:-1])) for s in srcs]
        for j, d in enumerate(dsts):
            dv = np.lib.stride_tricks.as_strided(out[d[-1]:], shape=szd, strides=tuple(8 * v for v in d[:-1]))
            dv += alpha * sum((Um[j, i] * sv[i] for i in range(Cn)), np.zeros(szd))
    return out


@pytest.mark.parametrize("case", ["u1", "fermion"])
def test_fused_equals_unfused(case):
    """The H_eff MPO step T1(a, w1, s1p, s2p, bp) · M1(w1, s1, s1p, w2)."""
    rng = np.random.default_rng(5)
    if case == "u1":
        V = U1Space((0, 5), (1, 3), (-1, 4), (2, 2))
        P = U1Space((0.5, 1), (-0.5, 1))
        W = U1Space((0, 3), (1, 1), (-1, 1))
        twA, twB = (), ()
    else:
        f = lambda pairs: GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})
        V = f([((0, 0), 4), ((1, 1), 3), ((1, -1), 2), ((0, 2), 2)])
        P = f([((0, 0), 1), ((1, 1), 1), ((1, -1), 1)])
        W = f([((0, 0), 2), ((1, 1), 1)])
        twA, twB = (), (0,)
    A = HomSpace(ProductSpace(V, W.dual(), P, P), ProductSpace(V))
    B = HomSpace(ProductSpace(W, P), ProductSpace(P, W))
    pA, pB = ((0, 3, 4), (1, 2)), ((0, 2), (1, 3))
    C2 = A.permute(*pA).compose(B.permute(*pB))       # (a, s2p, bp | s1, w2) GEMM layout
    for pC in (((0, 1, 2), (3, 4)), ((0, 3), (1, 2, 4))):  # as is, and as the H_eff step writes it
        C = C2.permute(*pC)
        a, b = rng.standard_normal(A.dim()), rng.standard_normal(B.dim())
        c0 = rng.standard_normal(C.dim())
        ref = _unfused(C, A, pA, B, pB, pC, twA, twB, a, b, 0.75, -1.25, c0)
        got = _fused(C, A, pA, B, pB, pC, twA, twB, a, b, 0.75, -1.25, c0)
        assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_mixed_element_types_are_refused():
    """A tiny float32 B with float64 A and C: the fused path would read B with C's
    element type, so it must not apply; the contraction raises before any launch."""
    import torch
    from tkmi import tensors as T
    V = U1Space((0, 3), (1, 2), (-1, 2))
    P = U1Space((0, 1), (1, 1))
    WA = HomSpace(ProductSpace(V, P), ProductSpace(V))
    WB = HomSpace(ProductSpace(P), ProductSpace(P))
    WC = HomSpace(ProductSpace(V, P), ProductSpace(V))

    def mk(W, dt):  # host tensors: nothing here reaches the device
        return T.TensorMap(torch.zeros(W.dim(), dtype=dt), W)

    pA, pB = ((0, 2), (1,)), ((0,), (1,))
    A, C = mk(WA, torch.float64), mk(WC, torch.float64)
    assert FC.eligible(C, A, pA, mk(WB, torch.float64), pB)
    for dA, dB, dC in [(torch.float64, torch.float32, torch.float64),
                       (torch.float32, torch.float64, torch.float64),
                       (torch.float32, torch.float32, torch.float64)]:
        A, B, C = mk(WA, dA), mk(WB, dB), mk(WC, dC)
        assert not FC.eligible(C, A, pA, B, pB)
        with pytest.raises(TypeError):
            T.blas_contract_(C, A, pA, B, pB, ((0, 2), (1,)), 1.0, 0.0)


def test_float32_gemm_arithmetic():
    """Float32 tensors multiply in FP32 (the sgemm of a Float32 mul!) unless 3×BF16 is asked
    for; FP64 arithmetic on float32 data is refused.  Float64 keeps the default form."""
    import torch
    from tkmi import tensors as T, _lib
    assert T.gemm_precision(torch.float32) == _lib.PREC_F32
    assert T.gemm_precision(torch.float32, _lib.PREC_BF16X3) == _lib.PREC_BF16X3
    with pytest.raises(TypeError):
        T.gemm_precision(torch.float32, _lib.PREC_F64)
    assert T.gemm_precision(torch.float64) == T.DEFAULT_PRECISION
    assert T.gemm_precision(torch.float64, _lib.PREC_BF16X3) == _lib.PREC_BF16X3
