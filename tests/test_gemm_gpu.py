"""GPU parity: grouped per-sector GEMM (K3 FP64, K4 FP32, K5 3×BF16) through
the C ABI vs numpy.  Integer-valued inputs make FP64 results exact (any
summation order), which pins the MFMA fragment layouts bit-for-bit."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from tkmi import _lib, tensors
    tensors._device()
    return _lib


def _make_groups(rng, shapes, transa, transb):
    """Pack groups into flat A, B, C host arrays; returns (arrays, tables)."""
    offs_a = offs_b = offs_c = 0
    tab = []
    for (m, n, k), ta, tb in zip(shapes, transa, transb):
        lda = (k if ta else m) + int(rng.integers(0, 3))
        ldb = (n if tb else k) + int(rng.integers(0, 3))
        ldc = m + int(rng.integers(0, 3))
        lda, ldb, ldc = max(lda, 1), max(ldb, 1), max(ldc, 1)
        asz = lda * (m if ta else k)
        bsz = ldb * (k if tb else n)
        csz = ldc * n
        tab.append((m, n, k, offs_a, lda, ta, offs_b, ldb, tb, offs_c, ldc))
        offs_a += asz
        offs_b += bsz
        offs_c += csz
    return tab, offs_a, offs_b, offs_c


def _run(lib, tab, A, B, Cm, prec, alpha, beta, tile=0):
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta = lib.i32(arr[:, 5])
    tb = lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    lib.check(L.tkmi_gemm_plan_set_tile(h, prec, tile), "tile")
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    dC = torch.from_numpy(Cm).cuda()
    s = torch.cuda.current_stream().cuda_stream
    if A.dtype == np.float32:
        # the blocks of Float32 tensors: A, B and C all float32
        assert B.dtype == Cm.dtype == np.float32
        lib.check(L.tkmi_gemm_exec_dtype(h, prec, lib.TKMI_F32, 3, alpha, beta, dA.data_ptr(), dB.data_ptr(),
                                         dC.data_ptr(), s), "exec f32")
    else:
        lib.check(L.tkmi_gemm_exec(h, prec, alpha, beta, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec")
    out = dC.cpu().numpy()
    L.tkmi_gemm_plan_destroy(h)
    return out


def _ref(tab, A, B, Cm, alpha, beta):
    out = Cm.copy()
    for (m, n, k, oa, lda, ta, ob, ldb, tb, oc, ldc) in tab:
        if m == 0 or n == 0:
            continue
        if k > 0:
            # op(A) is m×k: column-major (ld=lda) or its transpose
            a = A[oa:oa + lda * m].reshape(m, lda)[:, :k] if ta else A[oa:oa + lda * k].reshape(k, lda).T[:m, :]
            # op(B) is k×n
            b = B[ob:ob + ldb * k].reshape(k, ldb)[:, :n] if tb else B[ob:ob + ldb * n].reshape(n, ldb).T[:k, :]
            prod = a @ b
        else:
            prod = np.zeros((m, n))
        c = out[oc:oc + ldc * n].reshape(n, ldc).T[:m, :]
        c[...] = alpha * prod + (beta * c if beta != 0 else 0)
    return out


SHAPES = [(1, 1, 1), (16, 16, 4), (29, 24, 5), (64, 64, 16), (65, 63, 17), (130, 70, 33), (200, 129, 290),
          (5, 300, 7), (300, 4, 4), (17, 33, 0), (0, 5, 3)]


ALL_TILES = [64, 128, 1002, 1003, 1004, 1005, 1006, 1007, 1008, 1009, 1010, 1011, 1012, 1013, 1014, 1015, 1016,
             1017, 1018, 1019]


@pytest.mark.parametrize("tile", ALL_TILES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_f64_exact_integer(lib, ta, tb, tile):
    rng = np.random.default_rng(10 * ta + tb)
    tab, na, nb, nc = _make_groups(rng, SHAPES, [ta] * len(SHAPES), [tb] * len(SHAPES))
    A = rng.integers(-8, 9, na).astype(np.float64)
    B = rng.integers(-8, 9, nb).astype(np.float64)
    Cm = rng.integers(-8, 9, nc).astype(np.float64)
    for alpha, beta in ((1.0, 0.0), (2.0, -1.0), (1.0, 1.0)):
        out = _run(lib, tab, A, B, Cm, 0, alpha, beta, tile)
        ref = _ref(tab, A, B, Cm, alpha, beta)
        np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("tile", [1013, 1014, 1015, 1016, 1017, 1018, 1019])
def test_direct_load_kernel_exact(lib, tile):
    """The direct global → LDS kernel (configurations 13–16 on 64×64 tiles, 17–19 on
    128×128, 128×64 and 64×128; every group untransposed): whole tiles with one, two and
    many slabs, partial last slabs, ragged edge tiles, long K split (13, 17–19: the
    model's factor, 15, 16: two and three), odd leading dimensions
    and offsets (16-byte chunks only 8-byte aligned) — exact on integer data for β = 0, −1
    and 1, and equal bit for bit on random data to the one-buffer kernel's result at the
    same split factor (configurations 10–12); the same for the direct-load 3×BF16 kernel,
    which these configurations run in that form (transposes included: its panels are
    always untransposed)."""
    rng = np.random.default_rng(40 + tile)
    shapes = [(64, 64, 16), (128, 64, 32), (192, 130, 290), (64, 64, 1000), (70, 200, 45), (65, 65, 17),
              (256, 128, 2050), (33, 300, 600)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0] * len(shapes), [0] * len(shapes))
    assert any(r[4] % 2 for r in tab) or any(r[3] % 2 for r in tab)
    A = rng.integers(-8, 9, na).astype(np.float64)
    B = rng.integers(-8, 9, nb).astype(np.float64)
    Cm = rng.integers(-8, 9, nc).astype(np.float64)
    for alpha, beta in ((1.0, 0.0), (2.0, -1.0), (1.0, 1.0)):
        np.testing.assert_array_equal(_run(lib, tab, A, B, Cm, 0, alpha, beta, tile), _ref(tab, A, B, Cm, alpha, beta))
    A, B, Cm = rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc)
    out = _run(lib, tab, A, B, Cm, 0, 0.75, 0.5, tile)
    ref = _ref(tab, A, B, Cm, 0.75, 0.5)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    # the one-buffer kernel with the same split factor forms the same sums in the same k order
    if tile in (1014, 1015, 1016):  # 13 and 17–19 take the model's split factor, 14–16 fix it at 1–3
        same = {1014: 1010, 1015: 1011, 1016: 1012}[tile]
        np.testing.assert_array_equal(out, _run(lib, tab, A, B, Cm, 0, 0.75, 0.5, same))
        # 3×BF16: the direct-load kernel (configurations 13–16) forms the same six products
        # per slab in the same order as the register-staged one (5, 10–12)
        np.testing.assert_array_equal(_run(lib, tab, A, B, Cm, 2, 0.75, 0.5, tile),
                                      _run(lib, tab, A, B, Cm, 2, 0.75, 0.5, same))
    out = _run(lib, tab, A, B, Cm, 2, 0.75, 0.5, tile)
    assert np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref)


def test_default_config_large_products(lib):
    """Without autotuning, FP64 plans of untransposed products run the direct-load kernel on
    64×64 tiles (configuration 13), large uniform products and short ragged ones alike;
    exact on integer data."""
    _lib = lib
    L = lib.load()
    rng = np.random.default_rng(61)
    for shapes, want in (([(2048, 2048, 40)] * 8, 13), ([(300, 280, 290), (130, 70, 33)], 13)):
        tab, na, nb, nc = _make_groups(rng, shapes, [0] * len(shapes), [0] * len(shapes))
        A = rng.integers(-8, 9, na).astype(np.float64)
        B = rng.integers(-8, 9, nb).astype(np.float64)
        Cm = rng.integers(-8, 9, nc).astype(np.float64)
        arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
        cols = [_lib.i64(arr[:, i]) for i in range(11)]
        h = C.c_void_p()
        _lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                           cols[4][1], _lib.i32(arr[:, 5])[1], cols[6][1], cols[7][1],
                                           _lib.i32(arr[:, 8])[1], cols[9][1], cols[10][1]), "plan")
        try:
            cfg, nt = C.c_int32(), C.c_int64()
            _lib.check(L.tkmi_gemm_plan_query(h, _lib.PREC_F64, C.byref(cfg), C.byref(nt)), "query")
            assert cfg.value == want, (shapes, cfg.value)
            dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, Cm))
            s = torch.cuda.current_stream().cuda_stream
            _lib.check(L.tkmi_gemm_exec(h, _lib.PREC_F64, 2.0, -1.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s),
                       "exec")
            np.testing.assert_array_equal(dC.cpu().numpy(), _ref(tab, A, B, Cm, 2.0, -1.0))
        finally:
            L.tkmi_gemm_plan_destroy(h)


def test_f64_random(lib):
    rng = np.random.default_rng(7)
    shapes = [(int(rng.integers(1, 400)), int(rng.integers(1, 400)), int(rng.integers(1, 300))) for _ in range(12)]
    tab, na, nb, nc = _make_groups(rng, shapes, list(rng.integers(0, 2, 12)), list(rng.integers(0, 2, 12)))
    A, B, Cm = rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc)
    out = _run(lib, tab, A, B, Cm, 0, 1.0, 0.5)
    ref = _ref(tab, A, B, Cm, 1.0, 0.5)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec,tol", [(1, 2e-5), (2, 1e-6)])
def test_reduced_precision_forms(lib, prec, tol):
    """FP32 MFMA and 3×BF16 split forms: relative Frobenius error vs FP64."""
    rng = np.random.default_rng(11)
    shapes = [(300, 280, 290), (64, 64, 64), (33, 17, 1000), (128, 129, 5)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 1], [0, 0, 1, 1])
    A, B, Cm = rng.standard_normal(na), rng.standard_normal(nb), np.zeros(nc)
    ref = _ref(tab, A, B, Cm, 1.0, 0.0)
    for tile in (64, 128):
        out = _run(lib, tab, A, B, Cm, prec, 1.0, 0.0, tile)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        assert err <= tol, (prec, tile, err)


@pytest.mark.parametrize("tile", ALL_TILES)
def test_split_k_exact(lib, tile):
    """Few output tiles and long K engage split-K (deterministic workspace
    reduction); integer data keeps the check exact."""
    rng = np.random.default_rng(21)
    shapes = [(100, 90, 3000), (64, 64, 1000), (130, 33, 4097), (40, 700, 513)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 1], [1, 0, 0, 1])
    A = rng.integers(-4, 5, na).astype(np.float64)
    B = rng.integers(-4, 5, nb).astype(np.float64)
    Cm = rng.integers(-4, 5, nc).astype(np.float64)
    for alpha, beta in ((1.0, 0.0), (0.5, 2.0)):
        for prec in (0, 2):
            out = _run(lib, tab, A, B, Cm, prec, alpha, beta, tile)
            ref = _ref(tab, A, B, Cm, alpha, beta)
            if prec == 0:
                np.testing.assert_array_equal(out, ref)
            else:
                assert np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref)


def test_split_k_model_engages_split(lib):
    """The schedule model splits a lone long-K group (one output tile, 4096-deep K) and
    keeps a plan with many output tiles and short K whole; the split result stays exact."""
    rng = np.random.default_rng(22)
    L = lib.load()
    for shapes, expect_split in (([(64, 64, 4096)], True), ([(2048, 2048, 64)], False)):
        tab, na, nb, nc = _make_groups(rng, shapes, [0], [0])
        arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
        cols = [lib.i64(arr[:, i]) for i in range(11)]
        ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
        h = C.c_void_p()
        lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                          cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1],
                                          cols[10][1]), "plan")
        sk, slots = C.c_int64(), C.c_int64()
        lib.check(L.tkmi_gemm_plan_splitk(h, 5, C.byref(sk), C.byref(slots)), "splitk")
        L.tkmi_gemm_plan_destroy(h)
        assert slots.value >= 256  # at least one workgroup per CU
        assert (sk.value > 1) == expect_split, (shapes, sk.value)
        if expect_split:
            A = rng.integers(-4, 5, na).astype(np.float64)
            B = rng.integers(-4, 5, nb).astype(np.float64)
            Cm = rng.integers(-4, 5, nc).astype(np.float64)
            np.testing.assert_array_equal(_run(lib, tab, A, B, Cm, 0, 1.0, -1.0, 1005),
                                          _ref(tab, A, B, Cm, 1.0, -1.0))


def test_split_k_repeated_calls_and_graph_replays(lib):
    """A split-K plan reused across calls (β = 0 and β ≠ 0, every precision form) and
    replayed from a captured graph gives the full result every time: FP64 with integer
    data is exact, and every form is bitwise repeatable (the partial products are
    summed in a fixed order)."""
    rng = np.random.default_rng(23)
    shapes = [(64, 64, 4096), (200, 150, 3000), (130, 33, 2049)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0], [1, 0, 0])
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    try:
        sk, slots = C.c_int64(), C.c_int64()
        lib.check(L.tkmi_gemm_plan_splitk(h, 5, C.byref(sk), C.byref(slots)), "splitk")
        assert sk.value > 1
        A = rng.integers(-4, 5, na).astype(np.float64)
        B = rng.integers(-4, 5, nb).astype(np.float64)
        C0 = rng.integers(-4, 5, nc).astype(np.float64)
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        dC = torch.from_numpy(C0).cuda()
        for prec in (0, 1, 2):
            lib.check(L.tkmi_gemm_plan_set_tile(h, prec, 1005), "tile")
        s = torch.cuda.current_stream().cuda_stream
        exact = {}
        for prec in (0, 1, 2):
            for rep, beta in enumerate((0.0, -1.0, 0.0, 2.0)):
                dC.copy_(torch.from_numpy(C0))
                lib.check(L.tkmi_gemm_exec(h, prec, 1.0, beta, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec")
                out = dC.cpu().numpy()
                ref = _ref(tab, A, B, C0, 1.0, beta)
                if prec == 0:
                    np.testing.assert_array_equal(out, ref)
                else:
                    assert np.linalg.norm(out - ref) <= 1e-5 * np.linalg.norm(ref), (prec, rep)
                if (prec, beta) in exact:
                    assert np.array_equal(out, exact[(prec, beta)]), (prec, rep)
                exact[(prec, beta)] = out
        # graph replays of one FP64 call (β = 0), with eager work between the replays
        dC.zero_()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        cs = torch.cuda.Stream()
        cs.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(cs), torch.cuda.graph(g, stream=cs):
            lib.check(L.tkmi_gemm_exec(h, 0, 1.0, 0.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(),
                                       cs.cuda_stream), "exec (capture)")
        torch.cuda.current_stream().wait_stream(cs)
        ref = _ref(tab, A, B, C0, 1.0, 0.0)
        dC0 = torch.from_numpy(C0).cuda()
        for _ in range(4):
            dC.copy_(dC0)  # β = 0 overwrites every group's block; the padding keeps C0
            g.replay()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(dC.cpu().numpy(), ref)
            dC.copy_(dC0)
            lib.check(L.tkmi_gemm_exec(h, 0, 1.0, 0.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec")
            torch.cuda.synchronize()
            np.testing.assert_array_equal(dC.cpu().numpy(), ref)
    finally:
        torch.cuda.synchronize()
        L.tkmi_gemm_plan_destroy(h)


def test_exec_parts_split_reduction_separately(lib):
    """tkmi_gemm_exec_parts timing bits: the MFMA launch without its split-K reduction
    (bits 0 | 2) followed by the reduction alone (bit 3) and the skinny part (bit 1)
    gives exactly what tkmi_gemm_exec gives (the same kernels in the same order)."""
    rng = np.random.default_rng(24)
    shapes = [(64, 64, 4096), (200, 150, 3000), (130, 33, 2049), (300, 8, 20), (40, 70, 300)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 0, 1], [1, 0, 0, 1, 0])
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    try:
        A, B, C0 = rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc)
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        s = torch.cuda.current_stream().cuda_stream
        for prec in (0, 1, 2):
            lib.check(L.tkmi_gemm_plan_set_tile(h, prec, 1005), "tile")
            sk, slots = C.c_int64(), C.c_int64()
            lib.check(L.tkmi_gemm_plan_splitk(h, 5, C.byref(sk), C.byref(slots)), "splitk")
            assert sk.value > 1
            for beta in (0.0, -0.5):
                full = torch.from_numpy(C0).cuda()
                lib.check(L.tkmi_gemm_exec(h, prec, 1.25, beta, dA.data_ptr(), dB.data_ptr(), full.data_ptr(), s),
                          "exec")
                parts = torch.from_numpy(C0).cuda()
                for bits in (5, 8, 2):
                    lib.check(L.tkmi_gemm_exec_parts(h, prec, bits, 1.25, beta, dA.data_ptr(), dB.data_ptr(),
                                                     parts.data_ptr(), s), f"parts {bits}")
                assert torch.equal(parts, full), (prec, beta)
    finally:
        torch.cuda.synchronize()
        L.tkmi_gemm_plan_destroy(h)


def test_time_next_records_the_mfma_launch(lib):
    """tkmi_gemm_plan_time_next: the next MFMA-tile launch records its own start and end
    in the two events (elapsed time positive, within the time of a loop around it), the
    result is the untimed launch's bit for bit in every precision form, the hook is
    one-shot, and a single event is refused."""
    rng = np.random.default_rng(31)
    shapes = [(512, 384, 290), (200, 150, 3000), (300, 8, 20)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0], [0, 0, 1])
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    try:
        A, B, C0 = rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc)
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        st = torch.cuda.current_stream()
        s = st.cuda_stream
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0, w1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        e1.record(st)
        torch.cuda.synchronize()
        for prec in (0, 1, 2):
            plain = torch.from_numpy(C0).cuda()
            lib.check(L.tkmi_gemm_exec(h, prec, 0.75, 0.5, dA.data_ptr(), dB.data_ptr(), plain.data_ptr(), s), "exec")
            timed = torch.from_numpy(C0).cuda()
            lib.check(L.tkmi_gemm_plan_time_next(h, C.c_void_p(e0.cuda_event), C.c_void_p(e1.cuda_event)), "time")
            w0.record(st)
            lib.check(L.tkmi_gemm_exec(h, prec, 0.75, 0.5, dA.data_ptr(), dB.data_ptr(), timed.data_ptr(), s), "exec")
            w1.record(st)
            torch.cuda.synchronize()
            assert torch.equal(timed, plain), prec
            ms, outer = e0.elapsed_time(e1), w0.elapsed_time(w1)
            assert 0.0 < ms <= outer, (prec, ms, outer)
            # one-shot: a later launch leaves the events as they are
            lib.check(L.tkmi_gemm_exec(h, prec, 0.75, 0.5, dA.data_ptr(), dB.data_ptr(), timed.data_ptr(), s), "exec")
            torch.cuda.synchronize()
            assert e0.elapsed_time(e1) == ms, prec
        assert L.tkmi_gemm_plan_time_next(h, C.c_void_p(e0.cuda_event), None) != 0
        lib.check(L.tkmi_gemm_plan_time_next(h, None, None), "clear")
    finally:
        torch.cuda.synchronize()
        L.tkmi_gemm_plan_destroy(h)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_dot_groups(lib, ta, tb):
    """1×1 results with a long K (full contractions to a scalar) run as two-pass dot
    products in FP64 for every precision form: exact for integer data (any order), beside
    skinny groups in the same plan; α = 0 scales C; a repeated call is bit-identical."""
    rng = np.random.default_rng(40 + 2 * ta + tb)
    shapes = [(1, 1, 65), (1, 1, 8192), (1, 1, 8193), (1, 1, 100003), (1, 1, 2_500_000), (40, 8, 30), (1, 1, 7)]
    tab, na, nb, nc = _make_groups(rng, shapes, [ta] * len(shapes), [tb] * len(shapes))
    A = rng.integers(-3, 4, na).astype(np.float64)
    B = rng.integers(-3, 4, nb).astype(np.float64)
    Cm = rng.integers(-3, 4, nc).astype(np.float64)
    for prec in (0, 1, 2):
        for alpha, beta in ((1.0, 0.0), (2.0, -1.0), (0.0, 3.0)):
            out = _run(lib, tab, A, B, Cm, prec, alpha, beta)
            np.testing.assert_array_equal(out, _ref(tab, A, B, Cm, alpha, beta))
    # a k×1 operand with ld = k beyond the MFMA kernel's 32-bit slab offsets: fine here
    big = [(1, 1, 5_000_000)]
    tb_, na_, nb_, nc_ = _make_groups(rng, big, [0], [0])
    Ab, Bb = rng.integers(-2, 3, na_).astype(np.float64), rng.integers(-2, 3, nb_).astype(np.float64)
    Cb = np.zeros(nc_)
    np.testing.assert_array_equal(_run(lib, tb_, Ab, Bb, Cb, 0, 1.0, 0.0), _ref(tb_, Ab, Bb, Cb, 1.0, 0.0))
    Af, Bf = rng.standard_normal(na), rng.standard_normal(nb)
    o1 = _run(lib, tab, Af, Bf, Cm, 0, 1.0, 0.0)
    o2 = _run(lib, tab, Af, Bf, Cm, 0, 1.0, 0.0)
    assert np.array_equal(o1, o2)
    ref = _ref(tab, Af, Bf, Cm, 1.0, 0.0)
    assert np.linalg.norm(o1 - ref) <= 1e-12 * np.linalg.norm(ref)


# ---------------------------------------------------------------------------
# float32 tensors (tkmi_gemm_exec_dtype with TKMI_F32): A, B, C float32, results formed
# in FP64 and rounded once.  Integer data keeps every partial sum below 2^24, so the FP32
# MFMA, the exact bf16 split of a float and the final rounding are all exact.
# ---------------------------------------------------------------------------
def _f32(*xs):
    return [x.astype(np.float32) for x in xs]


@pytest.mark.parametrize("tile", ALL_TILES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_f32_data_exact_integer(lib, ta, tb, tile):
    rng = np.random.default_rng(50 + 10 * ta + tb)
    tab, na, nb, nc = _make_groups(rng, SHAPES, [ta] * len(SHAPES), [tb] * len(SHAPES))
    A = rng.integers(-8, 9, na).astype(np.float64)
    B = rng.integers(-8, 9, nb).astype(np.float64)
    Cm = rng.integers(-8, 9, nc).astype(np.float64)
    for prec in (1, 2):
        for alpha, beta in ((1.0, 0.0), (2.0, -1.0), (1.0, 1.0)):
            out = _run(lib, tab, *_f32(A, B, Cm), prec, alpha, beta, tile)
            assert out.dtype == np.float32
            np.testing.assert_array_equal(out, _ref(tab, A, B, Cm, alpha, beta).astype(np.float32))


@pytest.mark.parametrize("tile", [1013, 1014, 1015, 1016])
def test_f32_direct_load_kernel_exact(lib, tile):
    """The direct-load FP32 kernel (configurations 13–16 on FP32 data, every group
    untransposed): whole tiles, tiles past m and n, partial last slabs, long K split,
    odd leading dimensions and offsets (16-byte chunks only 4-byte aligned) — exact on
    integer data, and bit for bit the one-buffer FP32 kernel's result at the same split
    factor on random data."""
    rng = np.random.default_rng(70 + tile)
    shapes = [(64, 64, 16), (128, 64, 32), (192, 130, 290), (64, 64, 1000), (70, 200, 45), (65, 65, 17),
              (256, 128, 2050), (33, 300, 600), (3, 70, 100)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0] * len(shapes), [0] * len(shapes))
    A = rng.integers(-8, 9, na).astype(np.float64)
    B = rng.integers(-8, 9, nb).astype(np.float64)
    Cm = rng.integers(-8, 9, nc).astype(np.float64)
    for alpha, beta in ((1.0, 0.0), (2.0, -1.0), (1.0, 1.0)):
        out = _run(lib, tab, *_f32(A, B, Cm), 1, alpha, beta, tile)
        np.testing.assert_array_equal(out, _ref(tab, A, B, Cm, alpha, beta).astype(np.float32))
    A, B, Cm = _f32(rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc))
    out = _run(lib, tab, A, B, Cm, 1, 0.75, 0.5, tile)
    ref = _ref(tab, A.astype(np.float64), B.astype(np.float64), Cm.astype(np.float64), 0.75, 0.5)
    assert np.linalg.norm(out - ref) <= 2e-6 * np.linalg.norm(ref)
    if tile != 1013:
        same = {1014: 1010, 1015: 1011, 1016: 1012}[tile]
        np.testing.assert_array_equal(out, _run(lib, tab, A, B, Cm, 1, 0.75, 0.5, same))


@pytest.mark.parametrize("tile", [1002, 1005, 1008, 1009])
def test_f32_data_split_k_and_dots_exact(lib, tile):
    """Split-K groups (FP64 partials, reduced into float32 C) and dot groups (FP64 sums
    rounded once: a sum above 2^24 rounds like the exact sum cast to float32)."""
    rng = np.random.default_rng(61)
    shapes = [(100, 90, 3000), (64, 64, 1000), (130, 33, 4097), (40, 700, 513), (1, 1, 8193), (1, 1, 2_500_000),
              (40, 8, 30)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 1, 0, 1, 0], [1, 0, 0, 1, 1, 0, 0])
    A = rng.integers(-4, 5, na).astype(np.float64)
    B = rng.integers(-4, 5, nb).astype(np.float64)
    Cm = rng.integers(-4, 5, nc).astype(np.float64)
    for prec in (1, 2):
        for alpha, beta in ((1.0, 0.0), (0.5, 2.0), (0.0, 3.0)):
            out = _run(lib, tab, *_f32(A, B, Cm), prec, alpha, beta, tile)
            np.testing.assert_array_equal(out, _ref(tab, A, B, Cm, alpha, beta).astype(np.float32))


@pytest.mark.parametrize("prec,tol", [(1, 2e-5), (2, 1e-6)])
def test_f32_data_random(lib, prec, tol):
    """Random float32 data: relative Frobenius error against the FP64 product of the same
    float32 inputs (FP32 MFMA: FP32 accumulation; 3×BF16: an exact split, the three
    smallest of the nine products dropped)."""
    rng = np.random.default_rng(62)
    shapes = [(300, 280, 290), (64, 64, 64), (33, 17, 1000), (128, 129, 5), (12, 300, 20)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0, 1, 0], [0, 0, 1, 1, 1])
    A, B, Cm = _f32(rng.standard_normal(na), rng.standard_normal(nb), rng.standard_normal(nc))
    ref = _ref(tab, A.astype(np.float64), B.astype(np.float64), Cm.astype(np.float64), 1.5, -0.5)
    for tile in (1005, 1002):
        out = _run(lib, tab, A, B, Cm, prec, 1.5, -0.5, tile)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        assert err <= tol, (prec, tile, err)


def test_f32_data_refuses_fp64_arithmetic(lib):
    rng = np.random.default_rng(63)
    tab, na, nb, nc = _make_groups(rng, [(8, 8, 8)], [0], [0])
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), 1, cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    try:
        d = torch.zeros(max(na, nb, nc), dtype=torch.float32, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        assert L.tkmi_gemm_exec_dtype(h, 0, lib.TKMI_F32, 3, 1.0, 0.0, d.data_ptr(), d.data_ptr(), d.data_ptr(),
                                      s) == 6  # TKMI_ERR_UNSUPPORTED
        assert L.tkmi_gemm_exec_dtype(h, 1, 7, 3, 1.0, 0.0, d.data_ptr(), d.data_ptr(), d.data_ptr(), s) == 1
    finally:
        L.tkmi_gemm_plan_destroy(h)


def test_bf16x3_constant_operands(lib):
    """tkmi_gemm_plan_set_constant: a constant operand is split into bf16 planes once and
    reused while later calls pass the same pointer — a change made in place without
    tkmi_gemm_plan_invalidate is not seen (the reuse is real), invalidate or another
    pointer splits again — and the other operand is split at every call."""
    rng = np.random.default_rng(31)
    shapes = [(300, 280, 290), (130, 70, 33), (64, 64, 64)]
    tab, na, nb, nc = _make_groups(rng, shapes, [0, 1, 0], [0, 0, 1])
    arr = np.array(tab, dtype=np.int64).reshape(-1, 11)
    cols = [lib.i64(arr[:, i]) for i in range(11)]
    ta, tb = lib.i32(arr[:, 5]), lib.i32(arr[:, 8])
    L = lib.load()
    h = C.c_void_p()
    lib.check(L.tkmi_gemm_plan_create(C.byref(h), len(tab), cols[0][1], cols[1][1], cols[2][1], cols[3][1],
                                      cols[4][1], ta[1], cols[6][1], cols[7][1], tb[1], cols[9][1], cols[10][1]),
              "plan")
    s = torch.cuda.current_stream().cuda_stream
    Z = np.zeros(nc)

    def run(dA, dB):
        dC = torch.zeros(nc, dtype=torch.float64, device="cuda")
        lib.check(L.tkmi_gemm_exec(h, 2, 1.0, 0.0, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), s), "exec")
        return dC.cpu().numpy()

    def close(out, A, B):
        ref = _ref(tab, A, B, Z, 1.0, 0.0)
        return np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref)

    try:
        for mask, which in ((1, "A"), (2, "B")):
            A0, B0 = rng.standard_normal(na), rng.standard_normal(nb)
            dA, dB = torch.from_numpy(A0).cuda(), torch.from_numpy(B0).cuda()
            lib.check(L.tkmi_gemm_plan_set_constant(h, mask), "constant")
            first = run(dA, dB)
            assert close(first, A0, B0)
            # change the constant operand in place: the old split is reused, bit for bit
            X1 = rng.standard_normal(na if which == "A" else nb)
            (dA if which == "A" else dB).copy_(torch.from_numpy(X1))
            stale = run(dA, dB)
            np.testing.assert_array_equal(stale, first)
            A1, B1 = (X1, B0) if which == "A" else (A0, X1)
            lib.check(L.tkmi_gemm_plan_invalidate(h), "invalidate")
            assert close(run(dA, dB), A1, B1)
            # the other operand is split at every call
            Y2 = rng.standard_normal(nb if which == "A" else na)
            (dB if which == "A" else dA).copy_(torch.from_numpy(Y2))
            A2, B2 = (A1, Y2) if which == "A" else (Y2, B1)
            assert close(run(dA, dB), A2, B2)
            # another pointer for the constant operand: split again
            X3 = rng.standard_normal(na if which == "A" else nb)
            d3 = torch.from_numpy(X3).cuda()
            out3 = run(d3, dB) if which == "A" else run(dA, d3)
            assert close(out3, *((X3, B2) if which == "A" else (A2, X3)))
        # nothing constant: an in-place change is seen at once
        lib.check(L.tkmi_gemm_plan_set_constant(h, 0), "constant")
        A4 = rng.standard_normal(na)
        dA.copy_(torch.from_numpy(A4))
        assert close(run(dA, dB), A4, dB.cpu().numpy())
        assert L.tkmi_gemm_plan_set_constant(h, 4) == 1
    finally:
        L.tkmi_gemm_plan_destroy(h)
