// Bounds-checked CPU run of libtkmi's transform and grouped-GEMM kernels (built from the
// unchanged .hip sources against the CPU stand-in of hip/hip_runtime.h, under
// AddressSanitizer).  Every kernel class is launched through the C ABI on small random
// problems and checked against a direct loop over the same tables:
//   * abelian permutes (copy and transpose tiles, long and short rows, ranks 1–4),
//   * generic blocks with several source and destination trees (tree classes 1–8, source
//     chunks beyond 8), their adjoint, and gathered coefficients,
//   * every element type (F64, F32, C128, C64; complex α, β and conjugated sources),
//   * the grouped GEMM for every tile configuration and precision form (FP64, FP32,
//     3×BF16 on FP64 and FP32 data), complex groups, skinny and dot groups, split-K and
//     α = 0 — integer-valued data, so every result is exact.
// Prints "EMU OK" and exits 0 when every check passes; an out-of-bounds access anywhere
// ends the run with AddressSanitizer's report.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/tkmi.h"

typedef std::complex<double> cd;

static int g_fail = 0;
static std::mt19937_64 rng(12345);

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int irand(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); }
static double drand() { return std::uniform_real_distribution<double>(-1.0, 1.0)(rng); }

#define CHECK(cond, ...)                                                     \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);        \
            std::fprintf(stderr, __VA_ARGS__);                               \
            std::fprintf(stderr, " (%s)\n", tkmi_last_error());              \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

// ---- element types ------------------------------------------------------------------
static size_t esize(int dt) { return dt == TKMI_F64 ? 8 : dt == TKMI_F32 ? 4 : dt == TKMI_C128 ? 16 : 8; }
static bool cplx(int dt) { return dt == TKMI_C128 || dt == TKMI_C64; }

static cd get(const void *p, int dt, int64_t i) {
    switch (dt) {
        case TKMI_F64: return ((const double *)p)[i];
        case TKMI_F32: return ((const float *)p)[i];
        case TKMI_C128: return cd(((const double *)p)[2 * i], ((const double *)p)[2 * i + 1]);
        default: return cd(((const float *)p)[2 * i], ((const float *)p)[2 * i + 1]);
    }
}
static void set(void *p, int dt, int64_t i, cd v) {
    switch (dt) {
        case TKMI_F64: ((double *)p)[i] = v.real(); break;
        case TKMI_F32: ((float *)p)[i] = (float)v.real(); break;
        case TKMI_C128: ((double *)p)[2 * i] = v.real(); ((double *)p)[2 * i + 1] = v.imag(); break;
        default: ((float *)p)[2 * i] = (float)v.real(); ((float *)p)[2 * i + 1] = (float)v.imag();
    }
}
struct Buf {
    int dt;
    int64_t n;
    std::vector<unsigned char> host;
    void *dev = nullptr;
    Buf(int dt_, int64_t n_, bool integer = false) : dt(dt_), n(n_), host(esize(dt_) * (size_t)std::max<int64_t>(n_, 1)) {
        for (int64_t i = 0; i < n; ++i)
            set(host.data(), dt, i, integer ? cd(irand(-3, 3), cplx(dt) ? irand(-3, 3) : 0)
                                            : cd(drand(), cplx(dt) ? drand() : 0.0));
        // exactly the bytes of the vector: a read past its end is a heap overflow
        tkmi_alloc(&dev, esize(dt) * (size_t)std::max<int64_t>(n, 1));
        std::memcpy(dev, host.data(), host.size());
    }
    ~Buf() { tkmi_free(dev); }
    cd at_dev(int64_t i) const { return get(dev, dt, i); }
    cd at_host(int64_t i) const { return get(host.data(), dt, i); }
};
static double tol(int dt) { return (dt == TKMI_F32 || dt == TKMI_C64) ? 1e-5 : 1e-12; }

// ---- strided sub-blocks -------------------------------------------------------------
struct View {
    std::vector<int64_t> sz, st;
    int64_t off;
};
// a sub-block of the given sizes placed in a column-major parent block with padded
// leading dimensions, at the end of a buffer of current length `len`
static View place(const std::vector<int64_t> &sz, int64_t &len) {
    View v;
    v.sz = sz;
    int64_t s = 1;
    for (size_t d = 0; d < sz.size(); ++d) {
        v.st.push_back(s);
        s *= sz[d] + (irand(0, 3) == 0 ? irand(1, 2) : 0);
    }
    v.off = len + irand(0, 2);
    len = v.off + s;
    return v;
}
static int64_t at(const View &v, const std::vector<int64_t> &idx) {
    int64_t o = v.off;
    for (size_t d = 0; d < idx.size(); ++d) o += idx[d] * v.st[d];
    return o;
}
template <typename F>
static void each_index(const std::vector<int64_t> &sz, F f) {
    std::vector<int64_t> idx(sz.size(), 0);
    int64_t total = 1;
    for (auto s : sz) total *= s;
    for (int64_t c = 0; c < total; ++c) {
        f(idx);
        for (size_t d = 0; d < sz.size(); ++d) {
            if (++idx[d] < sz[d]) break;
            idx[d] = 0;
        }
    }
}

static std::vector<int64_t> rand_sizes(int N) {
    std::vector<int64_t> s(N);
    for (int d = 0; d < N; ++d) {
        const int r = irand(0, 9);
        s[d] = r == 0 ? 1 : (r < 7 ? irand(2, 9) : irand(20, 300));
    }
    int64_t tot = 1;
    for (auto x : s) tot *= x;
    while (tot > 40000) {  // keep the emulated runs short
        int d = irand(0, N - 1);
        tot /= s[d];
        s[d] = std::max<int64_t>(1, s[d] / 3);
        tot *= s[d];
    }
    return s;
}

// ---- transforms ---------------------------------------------------------------------
struct TBlockSpec {
    int R, C;
    std::vector<int64_t> sz_src;  // sizes in source order
    std::vector<View> dst, src;
    std::vector<double> U;  // R×C column-major
};

static void run_transform(int N, const std::vector<int32_t> &perm, std::vector<TBlockSpec> &blocks,
                          int64_t dst_len, int64_t src_len, bool abelian) {
    // tables
    std::vector<int64_t> rows, cols, szd, szs, sod, sos, ds, ss;
    std::vector<double> U, coeffs;
    for (auto &b : blocks) {
        rows.push_back(b.R);
        cols.push_back(b.C);
        for (int d = 0; d < N; ++d) szd.push_back(b.sz_src[perm[d]]);
        for (int d = 0; d < N; ++d) szs.push_back(b.sz_src[d]);
        for (auto &v : b.dst) {
            for (int d = 0; d < N; ++d) sod.push_back(v.st[d]);
            sod.push_back(v.off);
        }
        for (auto &v : b.src) {
            for (int d = 0; d < N; ++d) sos.push_back(v.st[d]);
            sos.push_back(v.off);
        }
        U.insert(U.end(), b.U.begin(), b.U.end());
        if (abelian) {
            coeffs.push_back(b.U[0]);
            for (int d = 0; d < N; ++d) ds.push_back(b.dst[0].sz[d]);
            for (int d = 0; d < N; ++d) ds.push_back(b.dst[0].st[d]);
            ds.push_back(b.dst[0].off);
            for (int d = 0; d < N; ++d) ss.push_back(b.src[0].sz[d]);
            for (int d = 0; d < N; ++d) ss.push_back(b.src[0].st[d]);
            ss.push_back(b.src[0].off);
        }
    }
    tkmi_transform_plan plan = nullptr;
    int rc;
    if (abelian)
        rc = tkmi_abelian_plan_create(&plan, N, perm.data(), (int64_t)blocks.size(), coeffs.data(), ds.data(),
                                      ss.data(), dst_len, src_len);
    else
        rc = tkmi_generic_plan_create(&plan, N, perm.data(), (int64_t)blocks.size(), rows.data(), cols.data(),
                                      U.data(), szd.data(), szs.data(), sod.data(), sos.data(), nullptr, nullptr,
                                      dst_len, src_len);
    CHECK(rc == TKMI_OK, "plan create rc=%d", rc);
    if (rc) return;
    const int dts[] = {TKMI_F64, TKMI_F32, TKMI_C128, TKMI_C64};
    for (int dt : dts) {
        const cd alpha = cplx(dt) ? cd(0.75, -0.5) : cd(0.75, 0), beta = cplx(dt) ? cd(0.5, 0.25) : cd(-0.5, 0);
        const int conj = cplx(dt) ? irand(0, 1) : 0;
        Buf src(dt, src_len), dst(dt, dst_len);
        rc = tkmi_transform_exec_ex(plan, dt, conj, src.dev, dst.dev, alpha.real(), alpha.imag(), beta.real(),
                                    beta.imag(), nullptr);
        CHECK(rc == TKMI_OK, "exec rc=%d", rc);
        // reference: every destination element written by a block
        std::vector<cd> ref(dst_len);
        std::vector<char> written(dst_len, 0);
        for (int64_t i = 0; i < dst_len; ++i) ref[i] = dst.at_host(i);
        for (auto &b : blocks) {
            std::vector<int64_t> szdst(N);
            for (int d = 0; d < N; ++d) szdst[d] = b.sz_src[perm[d]];
            for (int j = 0; j < b.R; ++j)
                each_index(szdst, [&](const std::vector<int64_t> &idx) {
                    std::vector<int64_t> sidx(N);
                    for (int d = 0; d < N; ++d) sidx[perm[d]] = idx[d];
                    cd s = 0;
                    for (int i = 0; i < b.C; ++i) {
                        cd x = src.at_host(at(b.src[i], sidx));
                        if (conj) x = std::conj(x);
                        s += b.U[j + (int64_t)i * b.R] * x;
                    }
                    const int64_t o = at(b.dst[j], idx);
                    ref[o] = beta * ref[o] + alpha * s;
                    written[o] = 1;
                });
        }
        double err = 0, nrm = 0;
        for (int64_t i = 0; i < dst_len; ++i) {
            err = std::max(err, std::abs(dst.at_dev(i) - ref[i]));
            nrm = std::max(nrm, std::abs(ref[i]));
        }
        CHECK(err <= tol(dt) * std::max(1.0, nrm), "transform dt=%d abelian=%d err=%g", dt, (int)abelian, err);
    }
    // adjoint (real data): dsrc = β dsrc + α Uᵀ·(inverse permutation of ddst)
    if (!abelian) {
        Buf ddst(TKMI_F64, dst_len), dsrc(TKMI_F64, src_len);
        rc = tkmi_transform_adjoint_exec(plan, TKMI_F64, ddst.dev, dsrc.dev, 1.25, 1.0, nullptr);
        CHECK(rc == TKMI_OK, "adjoint rc=%d", rc);
        std::vector<double> ref(src_len);
        for (int64_t i = 0; i < src_len; ++i) ref[i] = dsrc.at_host(i).real();
        for (auto &b : blocks) {
            std::vector<int64_t> szdst(N);
            for (int d = 0; d < N; ++d) szdst[d] = b.sz_src[perm[d]];
            each_index(szdst, [&](const std::vector<int64_t> &idx) {
                std::vector<int64_t> sidx(N);
                for (int d = 0; d < N; ++d) sidx[perm[d]] = idx[d];
                for (int i = 0; i < b.C; ++i) {
                    double s = 0;
                    for (int j = 0; j < b.R; ++j) s += b.U[j + (int64_t)i * b.R] * ddst.at_host(at(b.dst[j], idx)).real();
                    ref[at(b.src[i], sidx)] += 1.25 * s;
                }
            });
        }
        double err = 0;
        for (int64_t i = 0; i < src_len; ++i) err = std::max(err, std::abs(dsrc.at_dev(i).real() - ref[i]));
        CHECK(err <= 1e-11, "adjoint err=%g", err);
        // gathered coefficients: U[e] = coef[e]·g[idx[e]]
        const int64_t nu = (int64_t)U.size();
        std::vector<double> coef(nu), g(7);
        std::vector<int64_t> gi(nu);
        for (auto &x : g) x = drand();
        for (int64_t e = 0; e < nu; ++e) {
            coef[e] = drand();
            gi[e] = irand(-1, 6);
        }
        rc = tkmi_transform_plan_set_gather(plan, nu, coef.data(), gi.data());
        CHECK(rc == TKMI_OK, "set_gather rc=%d", rc);
        Buf gdev(TKMI_F64, 7);
        std::memcpy(gdev.dev, g.data(), 7 * sizeof(double));
        Buf src(TKMI_F64, src_len), dst(TKMI_F64, dst_len);
        rc = tkmi_transform_exec_gather(plan, TKMI_F64, src.dev, dst.dev, gdev.dev, 1.0, 0.0, nullptr);
        CHECK(rc == TKMI_OK, "exec_gather rc=%d", rc);
        std::vector<double> refg(dst_len);
        for (int64_t i = 0; i < dst_len; ++i) refg[i] = 0.0;
        int64_t u0 = 0;
        for (auto &b : blocks) {
            std::vector<int64_t> szdst(N);
            for (int d = 0; d < N; ++d) szdst[d] = b.sz_src[perm[d]];
            for (int j = 0; j < b.R; ++j)
                each_index(szdst, [&](const std::vector<int64_t> &idx) {
                    std::vector<int64_t> sidx(N);
                    for (int d = 0; d < N; ++d) sidx[perm[d]] = idx[d];
                    double s = 0;
                    for (int i = 0; i < b.C; ++i) {
                        const int64_t e = u0 + j + (int64_t)i * b.R;
                        const double u = gi[e] < 0 ? 0.0 : coef[e] * g[gi[e]];
                        s += u * src.at_host(at(b.src[i], sidx)).real();
                    }
                    refg[at(b.dst[j], idx)] = s;
                });
            u0 += (int64_t)b.R * b.C;
        }
        for (auto &b : blocks) (void)b;
        double errg = 0;
        for (auto &b : blocks) {
            std::vector<int64_t> szdst(N);
            for (int d = 0; d < N; ++d) szdst[d] = b.sz_src[perm[d]];
            for (int j = 0; j < b.R; ++j)
                each_index(szdst, [&](const std::vector<int64_t> &idx) {
                    const int64_t o = at(b.dst[j], idx);
                    errg = std::max(errg, std::abs(dst.at_dev(o).real() - refg[o]));
                });
        }
        CHECK(errg <= 1e-11, "gather err=%g", errg);
    }
    tkmi_transform_plan_destroy(plan);
}

static void transform_cases() {
    // fixed cases first: short innermost rows kept innermost while the next dims swap
    // (vector-transpose tiles), then random ranks, sizes and permutations
    const std::vector<std::pair<std::vector<int64_t>, std::vector<int32_t>>> fixed = {
        {{8, 5, 7}, {0, 2, 1}}, {{12, 9, 10, 3}, {0, 2, 1, 3}}, {{24, 17, 3, 20}, {0, 3, 2, 1}},
        {{3, 40, 33}, {0, 2, 1}}};
    for (int rep = 0; rep < 48 + 2 * (int)fixed.size(); ++rep) {
        const int fx = rep / 2;
        const bool is_fixed = fx < (int)fixed.size();
        const int N = is_fixed ? (int)fixed[fx].first.size() : irand(1, 4);
        std::vector<int32_t> perm(N);
        for (int d = 0; d < N; ++d) perm[d] = d;
        if (is_fixed)
            perm = fixed[fx].second;
        else
            std::shuffle(perm.begin(), perm.end(), rng);
        const bool abelian = rep % 2 == 0;
        std::vector<TBlockSpec> blocks;
        int64_t dst_len = 0, src_len = 0;
        const int nb = irand(1, 4);
        for (int b = 0; b < nb; ++b) {
            TBlockSpec B;
            B.R = abelian ? 1 : irand(1, 9);
            B.C = abelian ? 1 : (irand(0, 4) == 0 ? irand(9, 12) : irand(1, 4));
            B.sz_src = is_fixed ? fixed[fx].first : rand_sizes(N);
            if (!abelian) {
                int64_t tot = 1;
                for (auto x : B.sz_src) tot *= x;
                while (tot * (B.R + B.C) > 60000) {
                    int d = irand(0, N - 1);
                    tot /= B.sz_src[d];
                    B.sz_src[d] = std::max<int64_t>(1, B.sz_src[d] / 2);
                    tot *= B.sz_src[d];
                }
            }
            std::vector<int64_t> szd(N);
            for (int d = 0; d < N; ++d) szd[d] = B.sz_src[perm[d]];
            for (int j = 0; j < B.R; ++j) B.dst.push_back(place(szd, dst_len));
            for (int i = 0; i < B.C; ++i) B.src.push_back(place(B.sz_src, src_len));
            for (int e = 0; e < B.R * B.C; ++e) B.U.push_back(abelian ? (irand(0, 1) ? 1.0 : -1.0) : drand());
            blocks.push_back(B);
        }
        run_transform(N, perm, blocks, dst_len, src_len, abelian);
    }
}

// ---- grouped GEMM -----------------------------------------------------------------
struct GSpec {
    int64_t m, n, k, a_off, lda, b_off, ldb, c_off, ldc;
    int ta, tb;
};

static void gemm_cases() {
    // every kind of group: ragged MFMA tiles, skinny (both orientations), k = 0, a dot
    // product, and one long-K group that the schedule model splits
    // (the last four, untransposed only, for the direct-load kernels: whole 64×64 tiles with
    // a partial last slab; long enough to be split; whole 128×128 and 128×64 tiles with
    // whole slabs only; and, last in A with lda = m, a tile past m whose K is whole slabs,
    // so that a read of row m in the group's last column would leave the buffer)
    const int64_t shapes[][3] = {{1, 1, 1},   {16, 16, 4},  {29, 24, 5},   {65, 63, 17},  {5, 70, 7},
                                 {70, 4, 4},  {17, 33, 0},  {1, 1, 300},   {40, 8, 30},   {20, 20, 520},
                                 {130, 70, 45}, {64, 64, 600}, {136, 130, 32}, {70, 40, 32}};
    for (int variant = 0; variant < 4; ++variant) {
        const int nshapes = (int)(sizeof(shapes) / sizeof(shapes[0]));
        const int ng = nshapes - (variant == 0 ? 0 : 4);
        std::vector<GSpec> gs;
        int64_t na = 0, nb = 0, nc = 0;
        for (int g = 0; g < ng; ++g) {
            GSpec G;
            G.m = shapes[g][0];
            G.n = shapes[g][1];
            G.k = shapes[g][2];
            G.ta = variant & 1;
            G.tb = (variant >> 1) & 1;
            G.lda = (G.ta ? G.k : G.m) + (g == nshapes - 1 ? 0 : irand(0, 2));
            G.ldb = (G.tb ? G.n : G.k) + irand(0, 2);
            G.ldc = G.m + irand(0, 2);
            G.lda = std::max<int64_t>(G.lda, 1);
            G.ldb = std::max<int64_t>(G.ldb, 1);
            G.a_off = na;
            G.b_off = nb;
            G.c_off = nc;
            na += G.lda * (G.ta ? G.m : G.k);
            nb += G.ldb * (G.tb ? G.k : G.n);
            nc += G.ldc * G.n;
            gs.push_back(G);
        }
        std::vector<int64_t> m, n, k, ao, la, bo, lb, co, lc;
        std::vector<int32_t> ta, tb;
        for (auto &G : gs) {
            m.push_back(G.m); n.push_back(G.n); k.push_back(G.k);
            ao.push_back(G.a_off); la.push_back(G.lda); ta.push_back(G.ta);
            bo.push_back(G.b_off); lb.push_back(G.ldb); tb.push_back(G.tb);
            co.push_back(G.c_off); lc.push_back(G.ldc);
        }
        tkmi_gemm_plan plan = nullptr;
        int rc = tkmi_gemm_plan_create(&plan, ng, m.data(), n.data(), k.data(), ao.data(), la.data(), ta.data(),
                                       bo.data(), lb.data(), tb.data(), co.data(), lc.data());
        CHECK(rc == TKMI_OK, "gemm plan rc=%d", rc);
        if (rc) continue;
        auto reference = [&](const Buf &A, const Buf &B, const Buf &C, cd alpha, cd beta, int ca, int cb) {
            std::vector<cd> ref(nc);
            for (int64_t i = 0; i < nc; ++i) ref[i] = C.at_host(i);
            for (auto &G : gs)
                for (int64_t j = 0; j < G.n; ++j)
                    for (int64_t i = 0; i < G.m; ++i) {
                        cd s = 0;
                        for (int64_t l = 0; l < G.k; ++l) {
                            cd a = A.at_host(G.a_off + (G.ta ? l + i * G.lda : i + l * G.lda));
                            cd b = B.at_host(G.b_off + (G.tb ? j + l * G.ldb : l + j * G.ldb));
                            if (ca) a = std::conj(a);
                            if (cb) b = std::conj(b);
                            s += a * b;
                        }
                        cd &c = ref[G.c_off + i + j * G.ldc];
                        c = alpha * s + (beta != cd(0) ? beta * c : cd(0));
                    }
            return ref;
        };
        auto compare = [&](const Buf &C, const std::vector<cd> &ref, const char *what, int cfg, int prec) {
            double err = 0;
            for (int64_t i = 0; i < nc; ++i) err = std::max(err, std::abs(C.at_dev(i) - ref[i]));
            CHECK(err == 0.0, "gemm %s variant=%d cfg=%d prec=%d err=%g", what, variant, cfg, prec, err);
        };
        // real data: every configuration in FP64 (each kernel template once per transpose
        // case), the FP32 and 3×BF16 forms on the configurations with distinct kernels
        int ncfg = 0;
        for (int32_t bm, bn, bk, nbuf; tkmi_gemm_config_info(ncfg, &bm, &bn, &bk, &nbuf) == TKMI_OK;) ++ncfg;
        for (int cfg = 0; cfg < ncfg; ++cfg) {
            for (int prec = 0; prec < 3; ++prec) {
                // untransposed: every configuration in FP64, the distinct FP32 / 3×BF16 kernels
                // (one-buffer, double-buffered 128-wide, two-stage, persistent, direct-load
                // unsplit and split); transposed: the kernels those forms share
                const bool distinct = cfg <= 2 || cfg == 5 || cfg == 8 || cfg == 9 || cfg == 13 || cfg == 15;
                const bool run = (variant == 0 && (prec == 0 || distinct)) || cfg == 5 || cfg == 8 || cfg == 9 ||
                                 (cfg == 1 && prec == 2);
                if (!run) continue;
                const double t0 = now();
                rc = tkmi_gemm_plan_set_tile(plan, prec, 1000 + cfg);
                CHECK(rc == TKMI_OK, "set_tile rc=%d", rc);
                Buf A(TKMI_F64, na, true), B(TKMI_F64, nb, true), C(TKMI_F64, nc, true);
                rc = tkmi_gemm_exec(plan, prec, 2.0, -1.0, (const double *)A.dev, (const double *)B.dev,
                                    (double *)C.dev, nullptr);
                CHECK(rc == TKMI_OK, "exec rc=%d", rc);
                compare(C, reference(A, B, C, 2.0, -1.0, 0, 0), "f64 data", cfg, prec);
                if (prec > 0 && (cfg == 5 || cfg == 8 || cfg == 2 || cfg >= 13)) {
                    Buf A32(TKMI_F32, na, true), B32(TKMI_F32, nb, true), C32(TKMI_F32, nc, true);
                    rc = tkmi_gemm_exec_dtype(plan, prec, TKMI_F32, 3, 1.0, 1.0, A32.dev, B32.dev, C32.dev, nullptr);
                    CHECK(rc == TKMI_OK, "exec_dtype rc=%d", rc);
                    compare(C32, reference(A32, B32, C32, 1.0, 1.0, 0, 0), "f32 data", cfg, prec);
                }
                if (getenv("EMU_VERBOSE"))
                    std::printf("gemm variant %d cfg %d prec %d: %.2f s\n", variant, cfg, prec, now() - t0);
            }
        }
        // α = 0: C = β·C without reading A or B
        {
            Buf C(TKMI_F64, nc, true);
            rc = tkmi_gemm_exec(plan, 0, 0.0, 3.0, nullptr, nullptr, (double *)C.dev, nullptr);
            CHECK(rc == TKMI_OK, "alpha=0 rc=%d", rc);
            std::vector<cd> ref(nc);
            for (int64_t i = 0; i < nc; ++i) ref[i] = C.at_host(i);
            for (auto &G : gs)
                for (int64_t j = 0; j < G.n; ++j)
                    for (int64_t i = 0; i < G.m; ++i) ref[G.c_off + i + j * G.ldc] *= 3.0;
            compare(C, ref, "alpha=0", -1, 0);
        }
        // complex data (ComplexF64 on FP64 MFMA, ComplexF32 on FP32 MFMA)
        for (int dt : {TKMI_C128, TKMI_C64})
            for (int prec = 0; prec < 2; ++prec) {
                const double t0 = now();
                const int ca = irand(0, 1), cb = irand(0, 1);
                Buf A(dt, na, true), B(dt, nb, true), C(dt, nc, true);
                rc = tkmi_gemm_exec_z(plan, prec, dt, ca, cb, 2.0, -1.0, 1.0, 1.0, A.dev, B.dev, C.dev, nullptr);
                CHECK(rc == TKMI_OK, "exec_z rc=%d", rc);
                compare(C, reference(A, B, C, cd(2, -1), cd(1, 1), ca, cb), "complex", dt, prec);
                if (getenv("EMU_VERBOSE")) std::printf("complex variant %d dt %d: %.2f s\n", variant, dt, now() - t0);
            }
        tkmi_gemm_plan_destroy(plan);
    }
}

// Block Jacobi SVD (tkmi_jacobi.hip), called through its internal entry points: several
// blocks of one call (square, tall with a partial 16-column block, wide (transposed),
// tiny, one with an exactly zero column, one with two equal columns), each checked by
// U·diag(S)·Vt = A, orthonormal U and Vt, and S non-negative and descending.
struct emu_stream;
namespace tkmi {
struct JacobiWork;
int jacobi_create(JacobiWork **out, int64_t nblocks, const int64_t *m, const int64_t *n, const int64_t *a_off,
                  const int64_t *lda, const int64_t *u_off, const int64_t *s_off, const int64_t *vt_off);
void jacobi_destroy(JacobiWork *w);
int jacobi_svd(JacobiWork *w, const double *A, double *U, double *S, double *Vt, emu_stream *s, int *sweeps);
}  // namespace tkmi

static void jacobi_cases() {
    const std::vector<std::pair<int64_t, int64_t>> shapes = {{40, 40}, {70, 33}, {20, 57}, {5, 5}, {50, 20}, {34, 34}};
    std::vector<int64_t> m, n, aoff, lda, uoff, soff, voff;
    int64_t alen = 0, ulen = 0, slen = 0, vlen = 0;
    for (auto [mm, nn] : shapes) {
        const int64_t r = std::min(mm, nn);
        m.push_back(mm);
        n.push_back(nn);
        lda.push_back(mm + 3);
        aoff.push_back(alen);
        alen += (mm + 3) * nn;
        uoff.push_back(ulen);
        ulen += mm * r;
        soff.push_back(slen);
        slen += r;
        voff.push_back(vlen);
        vlen += r * nn;
    }
    std::vector<double> A(alen);
    for (auto &x : A) x = drand();
    for (int64_t i = 0; i < m[4]; ++i) A[aoff[4] + i + 7 * lda[4]] = 0.0;                      // a zero column
    for (int64_t i = 0; i < m[5]; ++i) A[aoff[5] + i + 9 * lda[5]] = A[aoff[5] + i + 2 * lda[5]];  // equal columns
    std::vector<double> U(ulen), S(slen), Vt(vlen);
    tkmi::JacobiWork *w = nullptr;
    int rc = tkmi::jacobi_create(&w, (int64_t)shapes.size(), m.data(), n.data(), aoff.data(), lda.data(), uoff.data(),
                                 soff.data(), voff.data());
    CHECK(rc == TKMI_OK, "jacobi_create rc=%d", rc);
    int sweeps = 0;
    rc = tkmi::jacobi_svd(w, A.data(), U.data(), S.data(), Vt.data(), nullptr, &sweeps);
    CHECK(rc == TKMI_OK, "jacobi_svd rc=%d", rc);
    tkmi::jacobi_destroy(w);
    std::printf("block Jacobi: %d sweeps\n", sweeps);
    for (size_t b = 0; b < shapes.size(); ++b) {
        const int64_t mm = m[b], nn = n[b], r = std::min(mm, nn);
        auto u = [&](int64_t i, int64_t k) { return U[uoff[b] + i + k * mm]; };
        auto vt = [&](int64_t k, int64_t j) { return Vt[voff[b] + k + j * r]; };
        double anorm = 0, rec = 0, ou = 0, ov = 0;
        for (int64_t i = 0; i < mm; ++i)
            for (int64_t j = 0; j < nn; ++j) {
                double x = 0;
                for (int64_t k = 0; k < r; ++k) x += u(i, k) * S[soff[b] + k] * vt(k, j);
                const double a = A[aoff[b] + i + j * lda[b]];
                anorm += a * a;
                rec += (x - a) * (x - a);
            }
        for (int64_t k = 0; k < r; ++k)
            for (int64_t l = 0; l < r; ++l) {
                double du = 0, dv = 0;
                for (int64_t i = 0; i < mm; ++i) du += u(i, k) * u(i, l);
                for (int64_t j = 0; j < nn; ++j) dv += vt(k, j) * vt(l, j);
                ou += (du - (k == l)) * (du - (k == l));
                ov += (dv - (k == l)) * (dv - (k == l));
            }
        CHECK(std::sqrt(rec) <= 1e-12 * std::sqrt(anorm), "jacobi block %zu: reconstruction %g", b, std::sqrt(rec));
        CHECK(std::sqrt(ou) <= 1e-12 * r && std::sqrt(ov) <= 1e-12 * r, "jacobi block %zu: orthonormality %g %g", b,
              std::sqrt(ou), std::sqrt(ov));
        for (int64_t k = 0; k < r; ++k) {
            CHECK(S[soff[b] + k] >= 0.0, "jacobi block %zu: negative sigma", b);
            if (k) CHECK(S[soff[b] + k] <= S[soff[b] + k - 1], "jacobi block %zu: sigma not descending", b);
        }
    }
    CHECK(S[soff[4] + 19] == 0.0, "the zero column gives an exact zero sigma (%g)", S[soff[4] + 19]);
}

int main() {
    const double t0 = now();
    jacobi_cases();
    std::printf("jacobi cases done (%d failures)\n", g_fail);
    std::fflush(stdout);
    transform_cases();
    std::printf("transform cases: %.1f s\n", now() - t0);
    std::printf("transform cases done (%d failures)\n", g_fail);
    std::fflush(stdout);
    gemm_cases();
    if (g_fail) {
        std::printf("EMU FAILED: %d checks\n", g_fail);
        return 1;
    }
    std::printf("EMU OK\n");
    return 0;
}
