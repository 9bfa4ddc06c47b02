"""CPU: transform-table files round-trip cache keys (spaces, index tuples, levels,
flags) and the packed plan arrays exactly; files load without pickle."""
import types

import numpy as np

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, HomSpace, ProductSpace, U1Space, SU2Space
from tkmi import treetransformers as TT
from tkmi import tablestore as TS


def _entry(key, kind, perm, lens, packed):
    return key, types.SimpleNamespace(kind=kind, perm=perm, lens=lens, packed=packed)


def test_round_trip(tmp_path):
    V = U1Space((0, 3), (1, 2), (-1, 1))
    F = GradedSpace({ProductSector(FermionParity(1), U1Irrep(1)): 2, ProductSector(FermionParity(0), U1Irrep(0)): 1})
    S = SU2Space((0, 2), (0.5, 1), (1, 1))
    entries = []
    # abelian permute with a dual leg and braid levels
    W = HomSpace(ProductSpace(V, V.dual()), ProductSpace(V))
    p1, p2 = (2, 0), (1,)
    Wd = W.permute(p1, p2)
    tab = TT.abelian_table(Wd, W, p1, p2)
    entries.append(_entry(("braid", Wd, W, p1, p2, ((0, 1), (2,)), False, ()), "abelian", p1 + p2,
                          (Wd.dim(), W.dim()), TT.pack_abelian(tab)))
    # fermionic permute with a destination twist
    Wf = HomSpace(ProductSpace(F, F), ProductSpace(F))
    Wfd = Wf.permute((1,), (2, 0))
    entries.append(_entry(("braid", Wfd, Wf, (1,), (2, 0), None, True, (0,)), "abelian", (1, 2, 0),
                          (Wfd.dim(), Wf.dim()), TT.pack_abelian(TT.abelian_table(Wfd, Wf, (1,), (2, 0)))))
    # SU(2) generic permute
    Ws = HomSpace(ProductSpace(S, S), ProductSpace(S, S))
    Wsd = Ws.permute((3, 1), (2, 0))
    rows = TT.generic_table(Wsd, Ws, (3, 1), (2, 0))
    entries.append(_entry(("braid", Wsd, Ws, (3, 1), (2, 0), None, False, ()), "generic", (3, 1, 2, 0),
                          (Wsd.dim(), Ws.dim()), TT.pack_generic(rows)))
    # SU(2) partial trace and a scalar (empty) destination
    Wt = Ws.select((1,), (3,))
    trows = TT.trace_table(Wt, Ws, (1,), (3,), (0,), (2,))
    entries.append(_entry(("trace", Wt, Ws, (1,), (3,), (0,), (2,), False), "trace", (0, 1), (Wt.dim(), Ws.dim()),
                          TT.pack_trace(trows)))
    W0 = Ws.select((), ())
    entries.append(_entry(("trace", W0, Ws, (), (), (0, 1), (2, 3), True), "trace", (), (1, Ws.dim()),
                          TT.pack_trace(TT.trace_table(W0, Ws, (), (), (0, 1), (2, 3)))))
    # planar transpose of a fermionic space, and an adjoint-trace key (no planar flag)
    Wfp = Wf.permute((0,), (2, 1))
    entries.append(_entry(("transpose", Wfp, Wf, (0,), (2, 1), True), "abelian", (0, 2, 1),
                          (Wfp.dim(), Wf.dim()), TT.pack_abelian(TT.abelian_table(Wfp, Wf, (0,), (2, 1)))))
    entries.append(_entry(("trace_adj", Ws, Wt, (1,), (3,), (0,), (2,)), "abelian", (0, 1, 2, 3),
                          (Ws.dim(), Wt.dim()), TT.pack_abelian(tab)))
    entries.append(_entry(("flat", 17), "abelian", (0,), (17, 17), TT.pack_abelian(
        {"coeffs": np.ones(1), "dst": np.array([[17, 1, 0]]), "src": np.array([[17, 1, 0]])})))
    path = tmp_path / "tables.npz"
    assert TS.save_tables(path, entries) == len(entries)
    back = TS.read_tables(path)
    assert len(back) == len(entries)
    for (key, plan), (k2, kind, perm, lens, packed) in zip(entries, back):
        assert k2 == key and hash(k2) == hash(key)
        assert kind == plan.kind and perm == tuple(plan.perm) and lens == tuple(plan.lens)
        assert sorted(packed) == sorted(plan.packed)
        for nm in packed:
            np.testing.assert_array_equal(packed[nm], plan.packed[nm])
            assert packed[nm].dtype == np.asarray(plan.packed[nm]).dtype
    # the scalar destination keeps its sector type
    assert back[4][0][1].unit == W0.unit


def test_rejects_foreign_file(tmp_path):
    path = tmp_path / "x.npz"
    blob = np.frombuffer(b'{"format": "other"}', dtype=np.uint8)
    with open(path, "wb") as f:
        np.savez(f, header=blob)
    try:
        TS.read_tables(path)
    except ValueError:
        return
    raise AssertionError("foreign file accepted")
