"""Transform tables on disk.

A table file holds entries of the transformer cache (`treebraider` permutes and braids,
`tracer` partial traces, `tr`, flat copies) under the same keys the cache uses, in the
spirit of the reference's `treebraider`/`treetransposer` caches
(`src/tensors/treetransformers.jl:156-180`, `src/auxiliary/caches.jl:135-173`):
- keys are stored as JSON (spaces as sector labels and degeneracies, index tuples, braid
  levels, flags);
- tables are stored as the packed integer/float arrays the C ABI plan constructors take.

Loading installs device plans directly from the arrays, so a later run skips the host
work of building fusion-tree tables (the dominant host cost for non-abelian tensors).
Files are `.npz` archives read with `allow_pickle=False`; nothing in them is executed.
"""
from __future__ import annotations

import json

import numpy as np

from . import treetransformers as TT
from .sectors import U1Irrep, SU2Irrep, FermionParity, ProductSector
from .spaces import GradedSpace, ProductSpace, HomSpace

FORMAT = "tkmi-transform-tables"
VERSION = 1


# --------------------------------------------------------------------------- keys
def sector_to_json(c):
    if isinstance(c, U1Irrep):
        return ["U1", c.two_q]
    if isinstance(c, SU2Irrep):
        return ["SU2", c.two_j]
    if isinstance(c, FermionParity):
        return ["fZ2", int(c.isodd)]
    if isinstance(c, ProductSector):
        return ["prod", [sector_to_json(x) for x in c.sectors]]
    raise TypeError(f"no table-file encoding for sector {c!r}")


def sector_from_json(d):
    t, v = d
    if t == "U1":
        return U1Irrep(two_q=int(v))
    if t == "SU2":
        return SU2Irrep(two_j=int(v))
    if t == "fZ2":
        return FermionParity(bool(v))
    if t == "prod":
        return ProductSector(*[sector_from_json(x) for x in v])
    raise ValueError(f"unknown sector tag {t!r}")


def space_to_json(V):
    return {"dual": bool(V.isdual), "dims": [[sector_to_json(c), int(n)] for c, n in V._dims.items()]}


def space_from_json(d):
    return GradedSpace({sector_from_json(c): n for c, n in d["dims"]}, d["dual"])


def homspace_to_json(W):
    return {"cod": [space_to_json(V) for V in W.codomain], "dom": [space_to_json(V) for V in W.domain],
            "unit": sector_to_json(W.unit)}


def homspace_from_json(d):
    return HomSpace(ProductSpace(*[space_from_json(x) for x in d["cod"]]),
                    ProductSpace(*[space_from_json(x) for x in d["dom"]]), unit=sector_from_json(d["unit"]))


def _tup(x):
    return None if x is None else tuple(_tup(y) if isinstance(y, (list, tuple)) else y for y in x)


def key_to_json(key):
    kind = key[0]
    if kind == "braid":
        _, Vd, Vs, p1, p2, levels, adj, twist = key
        return {"kind": kind, "dst": homspace_to_json(Vd), "src": homspace_to_json(Vs), "p1": list(p1),
                "p2": list(p2), "levels": None if levels is None else [list(levels[0]), list(levels[1])],
                "adjoint_src": bool(adj), "dst_twist": list(twist)}
    if kind == "transpose":
        _, Vd, Vs, p1, p2, adj = key
        return {"kind": kind, "dst": homspace_to_json(Vd), "src": homspace_to_json(Vs), "p1": list(p1),
                "p2": list(p2), "adjoint_src": bool(adj)}
    if kind in ("trace", "trace_adj"):
        # trace keys carry a trailing planar flag (planartrace!); adjoint trace keys do not
        _, Vd, Vs, p1, p2, q1, q2 = key[:7]
        d = {"kind": kind, "dst": homspace_to_json(Vd), "src": homspace_to_json(Vs), "p1": list(p1),
             "p2": list(p2), "q1": list(q1), "q2": list(q2)}
        if kind == "trace":
            d["planar"] = bool(key[7])
        return d
    if kind in ("tr", "tr_adj"):
        return {"kind": kind, "space": homspace_to_json(key[1])}
    if kind == "flip":
        _, Vd, Vs, inds, inv = key
        return {"kind": kind, "dst": homspace_to_json(Vd), "src": homspace_to_json(Vs), "inds": list(inds),
                "inv": bool(inv)}
    if kind == "flat":
        return {"kind": kind, "length": int(key[1])}
    raise TypeError(f"no table-file encoding for cache key kind {kind!r}")


def key_from_json(d):
    kind = d["kind"]
    if kind == "braid":
        lv = None if d["levels"] is None else (tuple(d["levels"][0]), tuple(d["levels"][1]))
        return ("braid", homspace_from_json(d["dst"]), homspace_from_json(d["src"]), tuple(d["p1"]),
                tuple(d["p2"]), lv, bool(d["adjoint_src"]), tuple(d["dst_twist"]))
    if kind == "transpose":
        return ("transpose", homspace_from_json(d["dst"]), homspace_from_json(d["src"]), tuple(d["p1"]),
                tuple(d["p2"]), bool(d["adjoint_src"]))
    if kind in ("trace", "trace_adj"):
        key = (kind, homspace_from_json(d["dst"]), homspace_from_json(d["src"]), tuple(d["p1"]),
               tuple(d["p2"]), tuple(d["q1"]), tuple(d["q2"]))
        return key + (bool(d["planar"]),) if kind == "trace" else key
    if kind in ("tr", "tr_adj"):
        return (kind, homspace_from_json(d["space"]))
    if kind == "flip":
        return ("flip", homspace_from_json(d["dst"]), homspace_from_json(d["src"]), tuple(d["inds"]), bool(d["inv"]))
    if kind == "flat":
        return ("flat", int(d["length"]))
    raise ValueError(f"unknown cache key kind {kind!r}")


# --------------------------------------------------------------------------- files
def save_tables(path, entries=None):
    """Write cache entries (default: every entry of the transformer cache) to `path`.
    `entries`: iterable of (key, plan) where plan carries `kind`, `perm`, `lens` and
    `packed` (as plans made by this package do)."""
    if entries is None:
        entries = list(TT._TRANSFORMER_CACHE.d.items())
    header = {"format": FORMAT, "version": VERSION, "entries": []}
    arrays = {}
    for i, (key, plan) in enumerate(entries):
        names = sorted(plan.packed)
        header["entries"].append({"key": key_to_json(key), "kind": plan.kind, "perm": list(plan.perm),
                                  "lens": list(plan.lens), "arrays": names})
        for nm in names:
            arrays[f"e{i}_{nm}"] = np.asarray(plan.packed[nm])
    blob = np.frombuffer(json.dumps(header).encode("utf-8"), dtype=np.uint8)
    with open(path, "wb") as f:
        np.savez(f, header=blob, **arrays)
    return len(header["entries"])


def read_tables(path):
    """Entries of a table file: a list of (key, kind, perm, lens, packed arrays)."""
    with np.load(path, allow_pickle=False) as z:
        header = json.loads(bytes(z["header"]).decode("utf-8"))
        if header.get("format") != FORMAT or header.get("version") != VERSION:
            raise ValueError(f"{path}: not a {FORMAT} v{VERSION} file")
        out = []
        for i, e in enumerate(header["entries"]):
            packed = {nm: z[f"e{i}_{nm}"] for nm in e["arrays"]}
            out.append((key_from_json(e["key"]), e["kind"], tuple(e["perm"]), tuple(e["lens"]), packed))
    return out


def load_tables(path):
    """Install every entry of a table file as a device plan in the transformer cache;
    returns the number of plans installed."""
    n = 0
    for key, kind, perm, lens, packed in read_tables(path):
        plan = TT.plan_from_packed(kind, perm, packed, lens[0], lens[1])
        TT._TRANSFORMER_CACHE.put(key, plan)
        n += 1
    return n
