"""GPU: edge cases of the kernels' calling conventions, as the reference inherits them from
BLAS / TensorOperations:
- β = 0: the destination is written without being read (NaN in it does not survive);
- α = 0: the sources are not read and C = β·C (`mul!` reduces to `rmul!`,
  LinearAlgebra's generic matmul; BLAS dgemm does not reference A and B);
- tensors without any block (dimension 0) go through permute and contract;
- C blocks with no matching A·B block get C_c = β·C_c (`mul!`'s rmul! branch,
  src/tensors/linalg.jl:330-370)."""
import numpy as np
import pytest
import torch

from tkmi.spaces import HomSpace, ProductSpace, U1Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _nan_like(T, W):
    return T.TensorMap(torch.full((W.dim(),), float("nan"), dtype=torch.float64, device="cuda"), W)


def _spaces():
    V = U1Space((0, 5), (1, 4), (-1, 3), (2, 2))
    P = U1Space((0, 1), (1, 1), (-1, 1))
    return V, P


def test_beta_zero_never_reads_destination(T):
    from tkmi import fusedcontract as FC
    rng = np.random.default_rng(1)
    V, P = _spaces()
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V, P)), rng)
    # permute
    C = _nan_like(T, A.space.permute((3, 0), (2, 1)))
    T.permute_(C, A, (3, 0), (2, 1), 1.0, 0.0)
    assert np.isfinite(C.numpy()).all()
    # grouped GEMM
    B = T.TensorMap.randn(HomSpace(ProductSpace(V, P), ProductSpace(V, P)), rng)
    C = _nan_like(T, A.space.compose(B.space))
    T.mul_(C, A, B, 1.0, 0.0)
    ref = C.numpy().copy()
    assert np.isfinite(ref).all()
    # trace
    Ct = _nan_like(T, A.space.select((0,), (2,)))
    T.trace_permute_(Ct, A, ((0,), (2,)), ((1,), (3,)), 1.0, 0.0)
    assert np.isfinite(Ct.numpy()).all()
    # small-operand contraction (fused transform)
    Wm = U1Space((0, 2), (1, 1))
    Ab = T.TensorMap.randn(HomSpace(ProductSpace(V, Wm.dual()), ProductSpace(P, V)), rng)
    Mb = T.TensorMap.randn(HomSpace(ProductSpace(Wm, P), ProductSpace(P, Wm)), rng)
    pA, pB, pAB = ((0, 3), (1, 2)), ((0, 1), (2, 3)), ((0, 2), (1, 3))
    VC = T.tensorcontract_space(Ab, pA, False, Mb, pB, False, pAB)
    assert FC.eligible(T.TensorMap.empty(VC), Ab, pA, Mb, pB)
    Cf = _nan_like(T, VC)
    T.tensorcontract_(Cf, Ab, pA, False, Mb, pB, False, pAB, 1.0, 0.0)
    assert np.isfinite(Cf.numpy()).all()


def test_alpha_zero_never_reads_sources(T):
    rng = np.random.default_rng(2)
    V, P = _spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V, P))
    A = _nan_like(T, W)
    B = _nan_like(T, W)
    # permute: C = β·C
    C0 = T.TensorMap.randn(W.permute((3, 0), (2, 1)), rng)
    C = C0.copy()
    T.permute_(C, A, (3, 0), (2, 1), 0.0, 0.5)
    assert np.array_equal(C.numpy(), 0.5 * C0.numpy())
    # grouped GEMM, every precision form: C = β·C, and β = 0 gives zeros
    from tkmi import _lib
    D0 = T.TensorMap.randn(W.compose(W), rng)
    for prec in (_lib.PREC_F64, _lib.PREC_F32, _lib.PREC_BF16X3):
        D = D0.copy()
        T.mul_(D, A, B, 0.0, -2.0, precision=prec)
        assert np.array_equal(D.numpy(), -2.0 * D0.numpy())
        T.mul_(D, A, B, 0.0, 0.0, precision=prec)
        assert not D.numpy().any()
    # small-operand contraction: the gathered coefficients are not applied
    Wm = U1Space((0, 2), (1, 1))
    Ab = _nan_like(T, HomSpace(ProductSpace(V, Wm.dual()), ProductSpace(P, V)))
    Mb = _nan_like(T, HomSpace(ProductSpace(Wm, P), ProductSpace(P, Wm)))
    pA, pB, pAB = ((0, 3), (1, 2)), ((0, 1), (2, 3)), ((0, 2), (1, 3))
    VC = T.tensorcontract_space(Ab, pA, False, Mb, pB, False, pAB)
    F0 = T.TensorMap.randn(VC, rng)
    F = F0.copy()
    T.tensorcontract_(F, Ab, pA, False, Mb, pB, False, pAB, 0.0, 1.5)
    assert np.array_equal(F.numpy(), 1.5 * F0.numpy())


def test_empty_tensors(T):
    rng = np.random.default_rng(3)
    V = U1Space((0, 3), (1, 2))
    Vodd = U1Space((5, 2))  # no fusion channel back to the charges of V
    W = HomSpace(ProductSpace(V, Vodd), ProductSpace(V))
    assert W.dim() == 0
    A = T.TensorMap.empty(W)
    B = T.permute(A, (2, 0), (1,), copy=True)
    assert B.dim() == 0
    L = T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(V)), rng)
    C = T.contract_labels(A, "apb", L, "bc", "apc", 2)
    assert C.dim() == 0


def test_unmatched_blocks_scale_by_beta(T):
    """C has coupled sectors that A·B lacks: those blocks become β·C_c."""
    rng = np.random.default_rng(4)
    V = U1Space((0, 3), (1, 2), (2, 2))
    Vs = U1Space((0, 3), (1, 2))
    A = T.TensorMap.randn(HomSpace(ProductSpace(V), ProductSpace(Vs)), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(Vs), ProductSpace(V)), rng)
    WC = HomSpace(ProductSpace(V), ProductSpace(V))
    C0 = T.TensorMap.randn(WC, rng)
    C = C0.copy()
    T.mul_(C, A, B, 1.25, -0.5)
    got, c0 = C.blocks_host(), C0.blocks_host()
    a, b = A.blocks_host(), B.blocks_host()
    for c, blk in got.items():
        ref = -0.5 * c0[c] + (1.25 * a[c] @ b[c] if c in a and c in b else 0.0)
        assert np.allclose(blk, ref, rtol=1e-13, atol=1e-13), c
    assert any(c not in a for c in got)  # the charge-2 block has no A·B partner


def test_empty_operand_into_nonempty_result(T):
    """A contraction whose operand has no blocks still writes C = β·C (zeros for β = 0):
    the fused small-operand transform then has destinations without sources and never
    reads A, so an empty A (no data pointer) is valid — fused and unfused paths agree."""
    from tkmi.sectors import U1Irrep, FermionParity, ProductSector
    from tkmi.spaces import GradedSpace
    f = lambda pairs: GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})
    V1, V2, V3 = f([((0, 0), 2), ((1, 1), 1)]), f([((1, -1), 2), ((0, 0), 1)]), f([((0, 1), 1), ((1, 0), 2)])
    V4, V5 = f([((0, 0), 1), ((1, 1), 1), ((1, -1), 1)]), f([((1, 0), 1), ((0, -1), 2), ((0, 0), 1)])
    rng = np.random.default_rng(5)
    t1 = T.TensorMap.randn(HomSpace(ProductSpace(V1, V2, V3), ProductSpace(V4)), rng)
    t2 = T.TensorMap.randn(HomSpace(ProductSpace(V2.dual(), V5), ProductSpace(V4.dual(), V1)), rng)
    assert t1.dim() == 0 and t2.dim() > 0
    from tkmi import fusedcontract as FC
    pA, pB, pAB = T.contraction_indices("xyaz", "ybzx", "ab", 1)
    VC = T.tensorcontract_space(t1, pA, False, t2, pB, False, pAB)
    assert FC.eligible(T.TensorMap.empty(VC), t1, pA, t2, pB)
    saved = T.FUSE_SMALL_OPERAND
    try:
        for fuse in (True, False):
            T.FUSE_SMALL_OPERAND = fuse
            C = T.contract_labels(t1, "xyaz", t2, "ybzx", "ab", 1)
            assert C.dim() > 0 and not C.numpy().any()
            C0 = T.TensorMap.randn(C.space, rng)
            C2 = C0.copy()
            T.contract_labels(t1, "xyaz", t2, "ybzx", "ab", 1, out=C2, beta=-0.5)
            assert np.array_equal(C2.numpy(), -0.5 * C0.numpy())
    finally:
        T.FUSE_SMALL_OPERAND = saved
