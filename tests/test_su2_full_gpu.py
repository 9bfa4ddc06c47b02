"""GPU: the non-abelian permute (K2) at the full SU(2) rank-4 size
V = SU2Space(0=>64, 1/2=>96, 1=>96, 3/2=>64, 2=>32), W = V⊗V ← V⊗V
(10 355 736 576 elements = 82.8 GB per tensor, input and output both resident).

Size-independent checks: the dim(c)-weighted norm is preserved
(test/tensors/indexmanipulations.jl:55-75), and sampled fusion blocks agree
with an independent numpy evaluation of the same recoupling table
(dst_j = Σ_i U[j,i]·permutedims(src_i, p)) to 1e-13.  The table itself is pinned
by the dense Clebsch-Gordan check at small degeneracies (test_su2_cpu.py)."""
import numpy as np
import pytest
import torch

from tkmi.spaces import HomSpace, ProductSpace, SU2Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def test_cfg3_full_size_permute(T):
    from tkmi import chainrules as CR
    free = torch.cuda.mem_get_info()[0]
    V = SU2Space((0, 64), (0.5, 96), (1, 96), (1.5, 64), (2, 32))
    W = HomSpace(ProductSpace(V, V), ProductSpace(V, V))
    n = W.dim()
    assert n == 10355736576
    if free < 2 * n * 8 + (8 << 30):
        pytest.skip(f"needs {2 * n * 8 / 2**30:.0f} GiB of device memory, {free / 2**30:.0f} GiB free")
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    A = T.TensorMap(torch.randn(n, dtype=torch.float64, device="cuda", generator=g), W)
    p1, p2 = (3, 1), (2, 0)
    B = T.permute(A, p1, p2)
    torch.cuda.synchronize()
    na, nb = CR.inner(A, A), CR.inner(B, B)
    assert abs(na - nb) <= 1e-12 * na
    plan = T.treebraider(B.space, W, p1, p2)
    perm = p1 + p2
    rows = sorted([r for r in plan.table if r[1].shape[0] >= 2], key=lambda r: r[0])[:12]
    rows += [r for r in plan.table if r[1].shape[0] == 1][:2]
    for w, U, sz_dst, so_dst, sz_src, so_src, fs_src, fs_dst in rows:
        srcs = [torch.as_strided(A.data, tuple(sz_src), tuple(so[:-1]), so[-1]).cpu().numpy() for so in so_src]
        for j, so in enumerate(so_dst):
            got = torch.as_strided(B.data, tuple(sz_dst), tuple(so[:-1]), so[-1]).cpu().numpy()
            ref = sum(U[j, i] * srcs[i].transpose(perm) for i in range(len(srcs)))
            assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
    del A, B
    torch.cuda.empty_cache()
