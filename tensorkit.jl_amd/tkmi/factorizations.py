"""svd_compact / svd_trunc for block-sparse tensors, FP64 on the device.

- svd_compact: per coupled block A_c = U_c S_c Vᴴ_c (`MAK.svd_compact!` per block,
  src/factorizations/matrixalgebrakit.jl:34-51), one `tkmi_svd_exec` call.
- svd_trunc: `truncrank(howmany) & trunctol(atol)` selection over all blocks with
  quantum-dimension weights and the truncation error ε (src/factorizations/
  truncation.jl:68-83,171-322), computed by `tkmi_truncate_select` on the
  device; the kept counts per block come back to the host to define the
  truncated space, then the leading columns/rows are copied on the device.

Singular values are returned as a `SectorVector`-like dict {c: device vector}.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import tensors as T
from .spaces import GradedSpace, HomSpace, ProductSpace
from .treetransformers import _make_abelian_plan, _LRU


class SVDPlan:
    def __init__(self, W):
        blocks = [(c, d1, d2, off) for c, (d1, d2), off in W.degeneracystructure()[1]]
        self.blocks = blocks
        arr = np.array([[d1, d2, off, max(d1, 1)] for c, d1, d2, off in blocks], dtype=np.int64).reshape(-1, 4)
        lib = _lib.ensure_device()
        h = C.c_void_p()
        cols = [_lib.i64(arr[:, i]) for i in range(4)]
        _lib.check(lib.tkmi_svd_plan_create(C.byref(h), len(blocks), cols[0][1], cols[1][1], cols[2][1],
                                            cols[3][1]), "tkmi_svd_plan_create")
        self.handle = h
        self._destroy = lib.tkmi_svd_plan_destroy
        self.r = [min(d1, d2) for _, d1, d2, _ in blocks]

    def __del__(self):
        # the destroy function is bound at creation: module globals may already be
        # torn down when the interpreter exits
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self.handle = None


# Factorisation plans own streams, rocBLAS handles and device workspace, and a sweep
# meets a new space at every truncation: the caches are bounded (least recently used
# plan evicted and destroyed) and emptied by empty_globalcaches().
FACTORIZATION_CACHE_SIZE = 16
_SVD_CACHE = _LRU(FACTORIZATION_CACHE_SIZE)

# The `alg` of svd_compact / svd_trunc: MatrixAlgebraKit's rocSOLVER SVD algorithms, and
# this library's block one-sided Jacobi (every block of the call in the same launches)
SVD_ALGORITHMS = {"qr_iteration": _lib.SVD_QRITERATION, "divide_and_conquer": _lib.SVD_DIVIDEANDCONQUER,
                  "jacobi": _lib.SVD_JACOBI, "block_jacobi": _lib.SVD_BLOCKJACOBI}
# block Jacobi: every block of the call in the same launches (the two-site H_eff output's 18
# blocks, up to 562², take about a third of the time of rocSOLVER's divide and conquer on
# eight concurrent lanes, σ within 1e-12 of LAPACK)
DEFAULT_SVD_ALGORITHM = "block_jacobi"
# the `alg` of eigh_full / eigh_trunc: the rocSOLVER symmetric eigensolvers
EIGH_ALGORITHMS = {"qr_iteration": _lib.SVD_QRITERATION, "divide_and_conquer": _lib.SVD_DIVIDEANDCONQUER,
                   "jacobi": _lib.SVD_JACOBI}
DEFAULT_EIGH_ALGORITHM = "divide_and_conquer"


def _mid_space(W, counts):
    """V_r: one leg carrying r_c copies of every coupled sector c."""
    return GradedSpace({c: n for c, n in counts.items() if n > 0})


def svd_compact(t, alg=None):
    """(U, S, Vᴴ) with U: codomain ← V_r, S: {c: σ_c} (descending), Vᴴ: V_r ← domain.
    `alg`: one of SVD_ALGORITHMS (default DEFAULT_SVD_ALGORITHM)."""
    W = t.space
    code = SVD_ALGORITHMS.get(alg or DEFAULT_SVD_ALGORITHM)
    if code is None:
        raise ValueError(f"unknown SVD algorithm {alg!r}; expected one of {sorted(SVD_ALGORITHMS)}")
    plan = _SVD_CACHE.get(W)
    if plan is None:
        plan = SVDPlan(W)
        _SVD_CACHE.put(W, plan)
    counts = {c: r for (c, _, _, _), r in zip(plan.blocks, plan.r)}
    Vr = _mid_space(W, counts)
    U = T.TensorMap.empty(HomSpace(W.codomain, ProductSpace(Vr)))
    Vt = T.TensorMap.empty(HomSpace(ProductSpace(Vr), W.domain))
    S = torch.empty(sum(plan.r), dtype=torch.float64, device=T._device())
    lib = _lib.load()
    _lib.check(lib.tkmi_svd_plan_set_algorithm(plan.handle, code), "svd algorithm")
    _lib.check(lib.tkmi_svd_exec(plan.handle, t.data.data_ptr(), U.data.data_ptr(), S.data_ptr(),
                                 Vt.data.data_ptr(), T._stream()), "svd")
    offs, o = {}, 0
    for (c, _, _, _), r in zip(plan.blocks, plan.r):
        offs[c] = (o, r)
        o += r
    return U, (S, offs), Vt


def _copy_leading(dst, src, sizes_src_dst):
    """Copy strided leading sub-matrices block by block with one transform plan."""
    rows_d, rows_s = [], []
    for (m, n, ld_s, off_s, ld_d, off_d) in sizes_src_dst:
        if m == 0 or n == 0:
            continue
        rows_d.append([m, n, 1, ld_d, off_d])
        rows_s.append([m, n, 1, ld_s, off_s])
    if not rows_d:
        return
    tab = {"coeffs": np.ones(len(rows_d)), "dst": np.array(rows_d, dtype=np.int64),
           "src": np.array(rows_s, dtype=np.int64)}
    plan = _make_abelian_plan(tab, (0, 1), dst.data.numel(), src.data.numel())
    lib = _lib.load()
    _lib.check(lib.tkmi_transform_exec(plan.handle, _lib.TKMI_F64, src.data.data_ptr(), dst.data.data_ptr(),
                                       1.0, 0.0, T._stream()), "truncate copy")


def svd_trunc(t, howmany=None, atol=0.0, alg=None):
    """`svd_trunc(t; trunc = truncrank(howmany) & trunctol(; atol))`.

    Returns (Ũ, S̃, Ṽᴴ, ε) with S̃ = {c: kept σ_c (device)} and ε the
    dim(c)-weighted 2-norm of the discarded singular values (a device scalar
    read back as float)."""
    U, (S, offs), Vt = svd_compact(t, alg)
    R = S.numel()
    weights = torch.empty(R, dtype=torch.float64, device=S.device)
    w_host = np.zeros(R)
    for c, (o, r) in offs.items():
        w_host[o:o + r] = c.dim()
    weights.copy_(torch.from_numpy(w_host))
    keep = torch.empty(max(R, 1), dtype=torch.uint8, device=S.device)
    disc = torch.empty(max(R, 1), dtype=torch.float64, device=S.device)
    err = torch.empty(1, dtype=torch.float64, device=S.device)
    hm = float(howmany) if howmany is not None else float("inf")
    lib = _lib.load()
    _lib.check(lib.tkmi_truncate_select(R, S.data_ptr(), weights.data_ptr(), hm, float(atol), keep.data_ptr(),
                                        disc.data_ptr(), err.data_ptr(), T._stream()), "truncate")
    keep_h = keep[:R].cpu().numpy().astype(bool)
    kept = {}
    for c, (o, r) in offs.items():
        k = int(keep_h[o:o + r].sum())
        # σ_c are sorted descending, so the kept set of a block is its leading k values
        if k and not keep_h[o:o + k].all():
            raise RuntimeError("truncation kept a non-leading singular value (ties)")
        kept[c] = k
    W = t.space
    Vk = _mid_space(W, kept)
    Ut = T.TensorMap.empty(HomSpace(W.codomain, ProductSpace(Vk)))
    Vtt = T.TensorMap.empty(HomSpace(ProductSpace(Vk), W.domain))
    uS = U.space.degeneracystructure()[1]
    vS = Vt.space.degeneracystructure()[1]
    uD = {c: (d, off) for c, d, off in Ut.space.degeneracystructure()[1]}
    vD = {c: (d, off) for c, d, off in Vtt.space.degeneracystructure()[1]}
    ucopy, vcopy = [], []
    for c, (d1, r), off in uS:
        if kept[c]:
            (dd1, k), doff = uD[c]
            ucopy.append((d1, kept[c], d1, off, dd1, doff))
    for c, (r, d2), off in vS:
        if kept[c]:
            (k, dd2), doff = vD[c]
            vcopy.append((kept[c], d2, r, off, k, doff))
    _copy_leading(Ut, U, ucopy)
    _copy_leading(Vtt, Vt, vcopy)
    Sk = {c: S[o:o + kept[c]] for c, (o, r) in offs.items() if kept[c]}
    return Ut, Sk, Vtt, float(err.item())


# ---------------------------------------------------------------------------
# eigh_full / eigh_trunc (symmetric blocks)
# ---------------------------------------------------------------------------
class EighPlan:
    def __init__(self, W):
        blocks = [(c, d1, d2, off) for c, (d1, d2), off in W.degeneracystructure()[1]]
        if any(d1 != d2 for _, d1, d2, _ in blocks):
            from .spaces import SpaceMismatch
            raise SpaceMismatch("eigh: every block must be square (domain ≅ codomain)")
        self.blocks = blocks
        arr = np.array([[d1, off, max(d1, 1)] for c, d1, d2, off in blocks], dtype=np.int64).reshape(-1, 3)
        lib = _lib.ensure_device()
        h = C.c_void_p()
        cols = [_lib.i64(arr[:, i]) for i in range(3)]
        _lib.check(lib.tkmi_eigh_plan_create(C.byref(h), len(blocks), cols[0][1], cols[1][1], cols[2][1]),
                   "tkmi_eigh_plan_create")
        self.handle = h
        self._destroy = lib.tkmi_svd_plan_destroy
        self.r = [d1 for _, d1, _, _ in blocks]

    def __del__(self):
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)
            self.handle = None


_EIGH_CACHE = _LRU(FACTORIZATION_CACHE_SIZE)


def eigh_full(t, alg=None):
    """`eigh_full(t)` for a symmetric t (`MAK.eigh_full!` per block,
    `src/factorizations/matrixalgebrakit.jl:34-51`): (D, V) with D = (λ, offsets),
    λ_c ascending per block, and V: codomain ← V_r orthogonal with t = V·D·Vᵀ.
    Only the lower triangle of each block is read."""
    W = t.space
    code = EIGH_ALGORITHMS.get(alg or DEFAULT_EIGH_ALGORITHM)
    if code is None:
        raise ValueError(f"unknown eigh algorithm {alg!r}; expected one of {sorted(EIGH_ALGORITHMS)}")
    plan = _EIGH_CACHE.get(W)
    if plan is None:
        plan = EighPlan(W)
        _EIGH_CACHE.put(W, plan)
    counts = {c: r for (c, _, _, _), r in zip(plan.blocks, plan.r)}
    Vr = _mid_space(W, counts)
    V = T.TensorMap.empty(HomSpace(W.codomain, ProductSpace(Vr)))
    lam = torch.empty(max(sum(plan.r), 1), dtype=torch.float64, device=T._device())
    lib = _lib.load()
    _lib.check(lib.tkmi_svd_plan_set_algorithm(plan.handle, code), "eigh algorithm")
    _lib.check(lib.tkmi_eigh_exec(plan.handle, t.data.data_ptr(), lam.data_ptr(), V.data.data_ptr(), T._stream()),
               "eigh")
    offs, o = {}, 0
    for (c, _, _, _), r in zip(plan.blocks, plan.r):
        offs[c] = (o, r)
        o += r
    return (lam, offs), V


def eigh_trunc(t, howmany=None, atol=0.0, alg=None):
    """`eigh_trunc(t; trunc = truncrank(howmany) & trunctol(; atol))`: keeps the
    eigenvalues of largest |λ| (dim(c)-weighted count, `truncation.jl:171-203`) that are
    not below atol in magnitude; returns (D̃ = {c: kept λ_c, ascending}, Ṽ, ε) with
    ε = sqrt(Σ_discarded dim(c)·λ²)."""
    (lam, offs), V = eigh_full(t, alg)
    R = sum(r for _, r in offs.values())
    weights = torch.empty(max(R, 1), dtype=torch.float64, device=lam.device)
    w_host = np.zeros(max(R, 1))
    for c, (o, r) in offs.items():
        w_host[o:o + r] = c.dim()
    weights.copy_(torch.from_numpy(w_host))
    keep = torch.empty(max(R, 1), dtype=torch.uint8, device=lam.device)
    disc = torch.empty(max(R, 1), dtype=torch.float64, device=lam.device)
    err = torch.empty(1, dtype=torch.float64, device=lam.device)
    hm = float(howmany) if howmany is not None else float("inf")
    lib = _lib.load()
    _lib.check(lib.tkmi_truncate_select(R, lam.data_ptr(), weights.data_ptr(), hm, float(atol), keep.data_ptr(),
                                        disc.data_ptr(), err.data_ptr(), T._stream()), "truncate")
    keep_h = keep[:R].cpu().numpy().astype(bool)
    kept = {c: int(keep_h[o:o + r].sum()) for c, (o, r) in offs.items()}
    W = t.space
    Vk = _mid_space(W, kept)
    Vt_ = T.TensorMap.empty(HomSpace(W.codomain, ProductSpace(Vk)))
    vS = {c: (d, off) for c, d, off in V.space.degeneracystructure()[1]}
    vD = {c: (d, off) for c, d, off in Vt_.space.degeneracystructure()[1]}
    # the kept columns of a block are runs of the ascending spectrum (largest |λ| sit at
    # both ends): copy each run as one strided sub-matrix
    copies, Dk = [], {}
    for c, (o, r) in offs.items():
        if not kept[c]:
            continue
        (n, _), soff = vS[c]
        (_, k), doff = vD[c]
        mask = keep_h[o:o + r]
        j, dst = 0, 0
        while j < r:
            if mask[j]:
                j1 = j
                while j1 < r and mask[j1]:
                    j1 += 1
                copies.append((n, j1 - j, n, soff + j * n, n, doff + dst * n))
                dst += j1 - j
                j = j1
            else:
                j += 1
        idx = torch.from_numpy(np.nonzero(mask)[0] + o).to(lam.device)
        Dk[c] = lam.index_select(0, idx)
    _copy_leading(Vt_, V, copies)
    return Dk, Vt_, float(err.item())
