"""GPU: complex element types (ComplexF64 = TKMI_C128, ComplexF32 = TKMI_C64).

* Abelian permutes (U1 and fZ2⊠U1) of complex tensors, bit for bit against the oracle's
  permute of the real and imaginary parts, with real and with complex α, β: the reference's
  tensoradd! forms (α·coeff)·A and β·C as Julia's complex products (four rounded real
  products, then their difference and sum) and adds them — restated here with separate
  numpy operations.  ComplexF32 data is promoted to FP64 and rounded once when stored.
* The adjoint of a complex tensor is read conjugated: permute(A′) equals the conjugate of
  the same permute of A's real and imaginary parts through the real path.
* SU(2) (several source trees per block): the complex permute equals the real path applied
  to the real and imaginary parts, bit for bit.
* The complex grouped GEMM (zgemm/cgemm of mul!): integer-valued data makes every product
  exact, so results equal numpy bit for bit — all transpose/conjugate combinations, complex
  α and β, split-K, k = 0 groups, α = 0.
* Contractions of complex tensors through the host mirror (tensorcontract! with conj
  operands) against the dense einsum.
"""
import numpy as np
import pytest
import torch
import ctypes as C

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, ProductSpace, HomSpace, U1Space, SU2Space

from oracle import abelian as O
from helpers import cfg1_spaces, oracle_legs, kind_of, random_u1_space, random_perm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _cplx_ref_permute(W, x, p1, p2, coeff_alpha=1.0):
    cod, dom = oracle_legs(W)
    re, _, _ = O.permute(np.ascontiguousarray(x.real), cod, dom, p1, p2, kind_of(W))
    im, _, _ = O.permute(np.ascontiguousarray(x.imag), cod, dom, p1, p2, kind_of(W))
    return re, im


def _cmul(ar, ai, xr, xi):
    """Julia's complex product: four separately rounded products, then difference and sum."""
    p1, p2, p3, p4 = ar * xr, ai * xi, ar * xi, ai * xr
    return p1 - p2, p3 + p4


def _fspace(pairs, dual=False):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs.items()}, dual)


_SCALARS = [(1.0, 0.0), (0.3, 0.0), (0.3, -1.7), (0.0, 2.0)]
_BSCALARS = [(0.0, 0.0), (0.7, 0.0), (0.7, 1.0 / 3.0), (1.0, 0.0)]


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("seed", range(8))
def test_abelian_complex_permutes(T, seed, dtype):
    rng = np.random.default_rng(3000 + seed)
    if seed % 2:
        W = HomSpace(ProductSpace(_fspace({(0, 0): 2, (1, 1): 3, (1, -1): 1}), _fspace({(0, 0): 1, (1, 1): 2})),
                     ProductSpace(_fspace({(0, 1): 2, (1, 0): 2}, True), _fspace({(0, 0): 1, (1, 1): 1, (0, -1): 2})))
        n1, n2 = 2, 2
    else:
        n1, n2 = int(rng.integers(1, 4)), int(rng.integers(0, 3))
        W = HomSpace(ProductSpace(*[random_u1_space(rng, maxdeg=6) for _ in range(n1)]),
                     ProductSpace(*[random_u1_space(rng, maxdeg=6) for _ in range(n2)]))
    if W.dim() == 0:
        pytest.skip("empty")
    p1, p2 = random_perm(rng, n1, n2)
    tdt = getattr(torch, dtype)
    A = T.TensorMap.randn(W, rng, dtype=tdt)
    x = A.numpy().astype(np.complex128)
    ar, ai = _SCALARS[seed % 4]
    br, bi = _BSCALARS[(seed // 4 + seed) % 4]
    C_ = T.TensorMap.randn(W.permute(p1, p2), rng, dtype=tdt)
    c0 = C_.numpy().astype(np.complex128)
    T.permute_(C_, A, p1, p2, complex(ar, ai), complex(br, bi))
    # per sub-block: the permuted source times the coefficient ±1 (exact), then the scalars
    sr, si = _cplx_ref_permute(W, x, p1, p2)
    pr, pi = _cmul(ar, ai, sr, si)
    if br == 0.0 and bi == 0.0:
        rr, ri = pr, pi
    else:
        qr, qi = _cmul(br, bi, c0.real, c0.imag)
        rr, ri = qr + pr, qi + pi
    ref = (rr + 1j * ri).astype(np.dtype(dtype))
    np.testing.assert_array_equal(C_.numpy(), ref)


def test_adjoint_is_conjugated(T):
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    rng = np.random.default_rng(11)
    A = T.TensorMap.randn(W, rng, dtype=torch.complex128)
    x = A.numpy()
    Are = T.TensorMap.from_numpy(x.real.copy(), W)
    Aim = T.TensorMap.from_numpy(x.imag.copy(), W)
    def materialized(t, p1, p2):
        # permute! into a plain tensor reads the lazy adjoint's data (permute() itself
        # returns the adjoint of a permuted parent)
        out = T.TensorMap.empty(t.space.permute(p1, p2), t.dtype)
        return T.permute_(out, t, p1, p2).numpy()

    for p1, p2 in [((0,), (1, 2)), ((2, 0), (1,)), ((1, 2, 0), ()), ((0, 1, 2), ())]:
        got = materialized(A.adjoint(), p1, p2)
        re = materialized(Are.adjoint(), p1, p2)
        im = materialized(Aim.adjoint(), p1, p2)
        np.testing.assert_array_equal(got, re - 1j * im)


def test_su2_complex_permute_matches_real_parts(T):
    V = SU2Space((0, 2), (0.5, 2), (1, 1))
    W = HomSpace(ProductSpace(V, V), ProductSpace(V, V))
    rng = np.random.default_rng(12)
    A = T.TensorMap.randn(W, rng, dtype=torch.complex128)
    x = A.numpy()
    Are = T.TensorMap.from_numpy(x.real.copy(), W)
    Aim = T.TensorMap.from_numpy(x.imag.copy(), W)
    for p1, p2 in [((3, 1), (2, 0)), ((0, 2), (1, 3)), ((1,), (3, 2, 0))]:
        got = T.permute(A, p1, p2, copy=True).numpy()
        ref = T.permute(Are, p1, p2, copy=True).numpy() + 1j * T.permute(Aim, p1, p2, copy=True).numpy()
        np.testing.assert_array_equal(got, ref)
        # complex α scales the recoupling sum
        C_ = T.TensorMap.zeros(W.permute(p1, p2), dtype=torch.complex128)
        T.permute_(C_, A, p1, p2, 0.5 - 0.25j, 0.0)
        assert np.linalg.norm(C_.numpy() - (0.5 - 0.25j) * ref) <= 1e-14 * np.linalg.norm(ref)


# ---------------------------------------------------------------------------
# complex grouped GEMM
# ---------------------------------------------------------------------------
def _groups(rng, shapes, transa, transb):
    oa = ob = oc = 0
    tab = []
    for (m, n, k), ta, tb in zip(shapes, transa, transb):
        lda = (k if ta else m) + int(rng.integers(0, 3))
        ldb = (n if tb else k) + int(rng.integers(0, 3))
        ldc = m + int(rng.integers(0, 3))
        lda, ldb, ldc = max(lda, 1), max(ldb, 1), max(ldc, 1)
        tab.append((m, n, k, oa, lda, ta, ob, ldb, tb, oc, ldc))
        oa += lda * (m if ta else k)
        ob += ldb * (k if tb else n)
        oc += ldc * n
    return tab, oa, ob, oc


def _zref(tab, A, B, Cm, alpha, beta, conja, conjb):
    out = Cm.copy()
    for (m, n, k, oa, lda, ta, ob, ldb, tb, oc, ldc) in tab:
        if m == 0 or n == 0:
            continue
        if k > 0:
            a = A[oa:oa + lda * m].reshape(m, lda)[:, :k] if ta else A[oa:oa + lda * k].reshape(k, lda).T[:m, :]
            b = B[ob:ob + ldb * k].reshape(k, ldb)[:, :n] if tb else B[ob:ob + ldb * n].reshape(n, ldb).T[:k, :]
            a = np.conj(a) if conja else a
            b = np.conj(b) if conjb else b
            prod = a @ b
        else:
            prod = np.zeros((m, n), dtype=A.dtype)
        c = out[oc:oc + ldc * n].reshape(n, ldc).T[:m, :]
        c[...] = alpha * prod + (beta * c if beta != 0 else 0)
    return out


def _zrun(lib, tab, A, B, Cm, prec, dtype, conja, conjb, alpha, beta):
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    try:
        npdt = np.complex128 if dtype == lib.TKMI_C128 else np.complex64
        dA, dB = torch.from_numpy(A.astype(npdt)).cuda(), torch.from_numpy(B.astype(npdt)).cuda()
        dC = torch.from_numpy(Cm.astype(npdt)).cuda()
        s = torch.cuda.current_stream().cuda_stream
        lib.check(L.tkmi_gemm_exec_z(h, prec, dtype, conja, conjb, alpha.real, alpha.imag, beta.real, beta.imag,
                                     dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec_z")
        out = dC.cpu().numpy()
        # a second call on the same plan reuses its complex tiles
        dC.copy_(torch.from_numpy(Cm.astype(npdt)))
        lib.check(L.tkmi_gemm_exec_z(h, prec, dtype, conja, conjb, alpha.real, alpha.imag, beta.real, beta.imag,
                                     dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec_z")
        assert np.array_equal(dC.cpu().numpy(), out)
    finally:
        L.tkmi_gemm_plan_destroy(h)
    return out


def _int_complex(rng, n, lo=-4, hi=5):
    return rng.integers(lo, hi, n).astype(np.float64) + 1j * rng.integers(lo, hi, n).astype(np.float64)


@pytest.mark.parametrize("dtype,prec", [("c128", 0), ("c64", 1), ("c64", 0), ("c128", 1)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_complex_gemm_exact(ta, tb, dtype, prec):
    from tkmi import _lib, tensors
    tensors._device()
    rng = np.random.default_rng(40 + 2 * ta + tb)
    shapes = [(1, 1, 1), (16, 16, 4), (29, 24, 5), (65, 63, 17), (130, 70, 33), (200, 129, 290), (5, 300, 7),
              (300, 4, 4), (17, 33, 0), (1, 1, 3000), (40, 8, 30)]
    tab, na, nb, nc = _groups(rng, shapes, [ta] * len(shapes), [tb] * len(shapes))
    A, B, Cm = _int_complex(rng, na), _int_complex(rng, nb), _int_complex(rng, nc)
    dt = _lib.TKMI_C128 if dtype == "c128" else _lib.TKMI_C64
    for conja, conjb, alpha, beta in [(0, 0, 1.0 + 0j, 0j), (ta, tb, 2.0 - 1.0j, -1.0 + 2.0j),
                                      (1, 0, 1j, 1.0 + 0j), (0, 1, 0j, 3.0 - 1j)]:
        out = _zrun(_lib, tab, A, B, Cm, prec, dt, conja, conjb, alpha, beta)
        ref = _zref(tab, A, B, Cm, alpha, beta, conja, conjb)
        np.testing.assert_array_equal(out, ref.astype(out.dtype), err_msg=str((conja, conjb, alpha, beta)))


def test_complex_gemm_split_k_and_refusals():
    from tkmi import _lib, tensors
    tensors._device()
    rng = np.random.default_rng(50)
    shapes = [(100, 90, 3000), (64, 64, 1000), (130, 33, 4097)]
    tab, na, nb, nc = _groups(rng, shapes, [0, 1, 0], [1, 0, 0])
    A, B, Cm = _int_complex(rng, na, -2, 3), _int_complex(rng, nb, -2, 3), _int_complex(rng, nc)
    out = _zrun(_lib, tab, A, B, Cm, 0, _lib.TKMI_C128, 1, 0, 0.5 + 1j, 1.0 - 1j)
    np.testing.assert_array_equal(out, _zref(tab, A, B, Cm, 0.5 + 1j, 1.0 - 1j, 1, 0))
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [_lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = _lib.i32(arr[:, 5]), _lib.i32(arr[:, 8])
    L = _lib.load()
    h = C.c_void_p()
    _lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                       cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
               "plan")
    try:
        d = torch.zeros(1, dtype=torch.complex128, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        assert L.tkmi_gemm_exec_z(h, 2, _lib.TKMI_C128, 0, 0, 1.0, 0.0, 0.0, 0.0, d.data_ptr(), d.data_ptr(),
                                  d.data_ptr(), s) == 6   # TKMI_ERR_UNSUPPORTED: no 3×BF16 for complex
        assert L.tkmi_gemm_exec_z(h, 0, _lib.TKMI_F64, 0, 0, 1.0, 0.0, 0.0, 0.0, d.data_ptr(), d.data_ptr(),
                                  d.data_ptr(), s) == 1   # TKMI_ERR_ARGUMENT: not a complex dtype
    finally:
        L.tkmi_gemm_plan_destroy(h)


def test_complex_contractions_dense(T):
    """tensorcontract! on complex U1 tensors (plain and conjugated operands) and mul! with
    an adjoint operand, against the dense einsum."""
    V, P = cfg1_spaces()
    rng = np.random.default_rng(13)
    WA = HomSpace(ProductSpace(V, P), ProductSpace(V))
    WL = HomSpace(ProductSpace(V), ProductSpace(V))
    A = T.TensorMap.randn(WA, rng, dtype=torch.complex128)
    L = T.TensorMap.randn(WL, rng, dtype=torch.complex128)
    kind = kind_of(WA)

    def dense(t):
        W = t.space
        x = t.numpy()
        return O.to_dense(np.ascontiguousarray(x.real), *oracle_legs(W), kind) + \
            1j * O.to_dense(np.ascontiguousarray(x.imag), *oracle_legs(W), kind)

    dA, dL = dense(A), dense(L)
    # C[a p; b] = L[a; y]·A[y p; b]
    Cc = T.contract_labels(L, "ay", A, "ypb", "apb", 2)
    ref = np.einsum("ay,ypb->apb", dL, dA)
    assert np.linalg.norm(dense(Cc) - ref) <= 1e-13 * np.linalg.norm(ref)
    # ρ[a; a'] = A[a p; b]·conj(A[a' p; b])
    R = T.contract_labels(A, "apb", A, "cpb", "ac", 1, conjB=True)
    refR = np.einsum("apb,cpb->ac", dA, np.conj(dA))
    assert np.linalg.norm(dense(R) - refR) <= 1e-13 * np.linalg.norm(refR)
    # mul!(C, L′, L) with the adjoint operand: conj-transposed blocks
    M = T.TensorMap.zeros(WL, dtype=torch.complex128)
    T.mul_(M, L.adjoint(), L, 1.0 - 0.5j, 0.0)
    refM = (1.0 - 0.5j) * (np.conj(dL).T @ dL)
    assert np.linalg.norm(dense(M) - refM) <= 1e-13 * np.linalg.norm(refM)
    # complex64 in FP32
    A32, L32 = (T.TensorMap.from_numpy(t.numpy(), t.space, torch.complex64) for t in (A, L))
    C32 = T.contract_labels(L32, "ay", A32, "ypb", "apb", 2)
    assert C32.dtype == torch.complex64
    assert np.linalg.norm(dense(C32) - ref) <= 1e-5 * np.linalg.norm(ref)
