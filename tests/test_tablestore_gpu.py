"""GPU: plans installed from a table file reproduce the freshly built plans bit for bit,
without rebuilding any fusion-tree table on the host."""
import numpy as np
import pytest

from tkmi.spaces import HomSpace, ProductSpace, U1Space, SU2Space

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from tkmi import tensors
    tensors._device()
    return tensors


def test_tables_from_file(T, tmp_path, monkeypatch):
    from tkmi import treetransformers as TT
    from tkmi import tablestore as TS
    rng = np.random.default_rng(8)
    S = SU2Space((0, 3), (0.5, 2), (1, 2))
    V = U1Space((0, 4), (1, 3), (-1, 2))
    A = T.TensorMap.randn(HomSpace(ProductSpace(S, S), ProductSpace(S, S)), rng)
    B = T.TensorMap.randn(HomSpace(ProductSpace(V, V.dual()), ProductSpace(V)), rng)

    def run():
        return [T.permute(A, (3, 1), (2, 0)).numpy(), T.permute(B, (2, 0), (1,)).numpy(),
                T.tensortrace(A, ((1,), (3,)), ((0,), (2,))).numpy(), np.array([T.tr(A)])]

    TT.empty_globalcaches()
    ref = run()
    path = tmp_path / "plans.npz"
    n = TS.save_tables(path)
    assert n >= 4
    TT.empty_globalcaches()

    def no_tables(*a, **k):
        raise AssertionError("table rebuilt on the host")

    monkeypatch.setattr(TT, "generic_table", no_tables)
    monkeypatch.setattr(TT, "abelian_table", no_tables)
    monkeypatch.setattr(TT, "trace_table", no_tables)
    assert TS.load_tables(path) == n
    got = run()
    for r, g in zip(ref, got):
        np.testing.assert_array_equal(g, r)
