"""GPU: small-operand contractions run as one fused transform (operand permutes and the
skinny per-sector GEMMs folded together); results agree with the unfused
permute + mul! path of blas_contract! (src/tensors/tensoroperations.jl:392-462)."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def _both(T, fn):
    saved = T.FUSE_SMALL_OPERAND
    try:
        T.FUSE_SMALL_OPERAND = True
        T.start_trace()
        fused = fn()
        trace = T.stop_trace()
        T.FUSE_SMALL_OPERAND = False
        plain = fn()
    finally:
        T.FUSE_SMALL_OPERAND = saved
    return fused, plain, trace


def test_heff_chain_fused_vs_unfused(T):
    from tkmi.heff import TwoSiteHeff, two_site_spaces
    V, P, W = two_site_spaces(D=160, qmax=4)
    H = TwoSiteHeff(V, P, W, np.random.default_rng(3))
    fused, plain, trace = _both(T, lambda: H.apply().numpy())
    assert sum(1 for op in trace if op[0] == "fused") == 2          # the two MPO steps
    assert np.linalg.norm(fused - plain) <= 1e-13 * np.linalg.norm(plain)


@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_fermionic_small_operand(T, beta):
    rng = np.random.default_rng(11)
    f = lambda pairs: GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): n for (p, q), n in pairs})
    V = f([((0, 0), 6), ((1, 1), 4), ((1, -1), 3), ((0, 2), 2)])
    P = f([((0, 0), 1), ((1, 1), 1), ((1, -1), 1)])
    W = f([((0, 0), 2), ((1, 1), 1)])
    A = T.TensorMap.randn(HomSpace(ProductSpace(V, W.dual(), P, P), ProductSpace(V)), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(W, P), ProductSpace(P, W)), rng)
    C0 = None

    def run():
        out = T.contract_labels(A, "awsxb", B, "wtsv", "axbtv", 3, alpha=0.75)
        if beta != 0.0:
            nonlocal C0
            if C0 is None:
                C0 = T.TensorMap.randn(out.space, np.random.default_rng(12))
            out2 = C0.copy()
            T.contract_labels(A, "awsxb", B, "wtsv", "axbtv", 3, out=out2, alpha=0.75, beta=beta)
            return out2.numpy()
        return out.numpy()

    fused, plain, trace = _both(T, run)
    assert any(op[0] == "fused" for op in trace)
    assert np.linalg.norm(fused - plain) <= 1e-13 * np.linalg.norm(plain)
