// Grouped per-coupled-sector GEMM on CDNA4 matrix cores (K3 FP64, K4 FP32, K5 3×BF16).
//
// One launch computes C_g = α·op(A_g)·op(B_g) + β·C_g for every coupled sector g of a
// contraction (the merge-walk of mul!, src/tensors/linalg.jl:330-370).  Tiles of all
// groups are numbered heaviest first (per-tile K, descending: the long tiles start
// early and the short ones fill the tail); workgroup → tile goes through an XCD-aware
// remap (chunks of consecutive tiles dealt round-robin to the 8 XCDs) and a binary
// search over the groups' tile_start prefix sums.
//
// Workgroup = 256 threads = 4 wave64s in a 2×2 arrangement; a BM×BN output tile is
// computed as (BM/2)×(BN/2) per wave from 16×16 MFMA fragments.  K advances in BK
// slabs staged through double-buffered LDS; the next slab's global loads are issued
// before the current slab's MFMAs.
//
// MFMA lane layouts used (wave64, i = row, j = col, l = lane):
//   v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32:
//       A[i][k]: i = l%16, k = l/16;  B[k][j]: j = l%16, k = l/16
//       D[i][j], 4 values v per lane, j = l%16 and
//         f32: i = 4*(l/16) + v      (4 consecutive rows per lane)
//         f64: i = (l/16) + 4*v      (rows interleaved by 4)
//   v_mfma_f32_16x16x32_bf16: A[i][k], 8 values e per lane: i = l%16, k = 8*(l/16) + e
//       (B likewise with j), D as above.
#include "tkmi_internal.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

namespace {

constexpr int NT = 256;

struct GGroup {
    int64_t m, n, k;
    int64_t a_off, lda, b_off, ldb, c_off, ldc;
    int32_t ta, tb;
    int64_t tiles_m, tiles_n, tile_start;
    int32_t skinny_mode;  // skinny kernel only: 0 = long along m, 1 = long along n
    int32_t pad_;
    int64_t splits;  // split-K factor (MFMA kernels): tiles_m·tiles_n·splits workgroups
    int64_t kchunk;  // K elements per split (multiple of 32)
    int64_t ws_off;  // split > 1: partial products at ws[ws_off + z·m·n], z < splits
    // 3×BF16 split panels of this group: A as 3 planes [mp][kp] at pa (plane stride sa),
    // B as 3 planes [np][kp] at pb (stride sb); rows padded to 128, K to 32, zero filled
    int64_t pa, pb, kp, sa, sb;
};

// Work of the split/pack pre-pass for one MFMA group: nA tiles of A rows, then the
// tiles of B columns (each PACK_T rows × PACK_T of K, K fastest), numbered from start.
// startA / startB number the A tiles and the B tiles on their own, for a pass that splits
// only one operand (the other one's panels are still valid: see tkmi_gemm_plan_set_constant).
struct PackGroup {
    int64_t m, n, k, a_off, lda, b_off, ldb;
    int32_t ta, tb;
    int64_t mp, np, kp, pa, pb, sa, sb;
    int64_t start, nA, startA, startB, nB;
};

// Split-K reduction record: C = α·Σ_z ws[ws_off + z·m·n + ·] + β·C, fixed z order.
struct RGroup {
    int64_t m, n, c_off, ldc, ws_off, splits, elem_start;
};

// Groups whose smaller output dimension is at most SKINNY (and K at most SKINNY_K) go
// to the bandwidth-bound skinny kernel (they would leave most of an MFMA tile idle).
constexpr int64_t SKINNY = 32;
constexpr int64_t SKINNY_K = 64;  // and a short K: otherwise one thread's serial K loop dominates

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

__device__ inline int64_t xcd_remap(int64_t bid, int64_t ntiles) {
    // Hardware dispatches workgroup ids round-robin over the 8 XCDs.  Tiles are dealt
    // to XCDs in chunks of CH consecutive tiles (chunk c → XCD c % 8): the tiles of a
    // chunk share A panels in one XCD's L2, and every XCD sees the same mix of heavy
    // (long-K) and light groups, so no XCD finishes last by a wide margin.  The tail
    // that does not fill a round of 8 chunks keeps the identity mapping.
    constexpr int64_t NX = 8, CH = 16;
    const int64_t full = ntiles / (NX * CH) * (NX * CH);
    if (bid >= full) return bid;
    const int64_t x = bid % NX, pos = bid / NX;
    return ((pos / CH) * NX + x) * CH + pos % CH;
}

__device__ inline int find_group(const GGroup *__restrict__ g, int64_t ng, int64_t t) {
    int64_t lo = 0, hi = ng - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (g[mid].tile_start <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return (int)lo;
}

// One element of C: r (the FP64 result α·(A·B)) plus β·C, rounded once to C's element type.
template <typename TD>
__device__ __forceinline__ void put_c(TD *p, double r, double beta, bool has_beta) {
    *p = (TD)(has_beta ? r + beta * (double)*p : r);
}

// --------------------------------------------------------------------------------
// FP64 / FP32 kernel (same data flow; T_LDS = compute type of the MFMA)
// --------------------------------------------------------------------------------
template <typename TC>
struct MfmaK4;
template <>
struct MfmaK4<double> {
    typedef d4_t acc_t;
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct MfmaK4<float> {
    typedef f4_t acc_t;
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// NBUF = 2: double-buffered LDS, one barrier per slab.  NBUF = 1: one LDS buffer (half
// the LDS, so more workgroups per CU), the next slab waits in registers while the
// current one is multiplied; two barriers per slab.  NBUF = 3: one LDS buffer and two
// register stages, loads two slabs ahead.  TD is the element type of A, B and C in
// memory (double, or float for FP32 tensors); split-K partial products stay FP64.
template <typename TC, typename TD, int BM, int BN, int BK, int NBUF>
__device__ __forceinline__ void gemm_k4_body(const GGroup *__restrict__ groups, int64_t ngroups, int64_t ntiles,
                                             const TD *__restrict__ A, const TD *__restrict__ Bm,
                                             TD *__restrict__ C, double *__restrict__ ws, double alpha,
                                             double beta) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int AE = BM * BK / NT;  // A elements per thread per slab
    constexpr int BE = BN * BK / NT;
    // LDS rows hold the K slab of one output row/column contiguously; the pad makes the
    // row pitch an odd multiple of 16 bytes, so the 16-byte reads of 16 lanes (16 rows)
    // fall in distinct banks.
    constexpr int KP = BK + 16 / (int)sizeof(TC);
    constexpr int NLDS = NBUF == 2 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) TC As[NLDS][BM][KP];
    __shared__ __attribute__((aligned(16))) TC Bs[NLDS][BN][KP];
    using acc_t = typename MfmaK4<TC>::acc_t;

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);  // end of this split's K range
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t lda = G.lda, ldb = G.ldb;
    const int ta = G.ta, tb = G.tb;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // Per-thread element coordinates of the A and B slabs (fixed over K): global order is
    // coalesced (contiguous index fastest), the LDS position is [row][k].  Addresses are a
    // wave-uniform slab base pointer (advanced on the scalar side) plus a 32-bit byte
    // offset per element, so the loads take the scalar-base form; interior tiles and
    // slabs skip every bounds test.
    // element e of this thread: (row i, k) of the A slab, (k, column j) of the B slab;
    // cheap shifts of tid, recomputed where needed rather than held in registers
    auto a_el = [&](int e, int &i, int &k) {
        const int idx = e * NT + tid;
        if (!ta) { i = idx % BM; k = idx / BM; }
        else     { k = idx % BK; i = idx / BK; }
    };
    auto b_el = [&](int e, int &j, int &k) {
        const int idx = e * NT + tid;
        if (!tb) { k = idx % BK; j = idx / BK; }
        else     { j = idx % BN; k = idx / BN; }
    };
    uint32_t aoff[AE], boff[BE];  // byte offsets from the slab base
#pragma unroll
    for (int e = 0; e < AE; ++e) {
        int i, k;
        a_el(e, i, k);
        aoff[e] = (uint32_t)(sizeof(TD) * (ta ? (k + i * lda) : (i + k * lda)));
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
        int j, k;
        b_el(e, j, k);
        boff[e] = (uint32_t)(sizeof(TD) * (tb ? (j + k * ldb) : (k + j * ldb)));
    }
    auto ld_at = [](const TD *base, uint32_t off) {
        return *reinterpret_cast<const TD *>(reinterpret_cast<const char *>(base) + off);
    };
    const int64_t astep = ta ? BK : BK * lda;  // pointer advance per K slab
    const int64_t bstep = tb ? BK * ldb : BK;
    const TD *__restrict__ Aslab = A + G.a_off + (ta ? row0 * lda + kbeg : row0 + kbeg * lda);
    const TD *__restrict__ Bslab = Bm + G.b_off + (tb ? col0 + kbeg * ldb : col0 * ldb + kbeg);
    const bool tile_full = (row0 + BM <= m) && (col0 + BN <= n);

    // slabs are loaded strictly in K order: every load advances the slab pointers
    TD ra[AE], rb[BE];
    auto load_into = [&](TD (&xa)[AE], TD (&xb)[BE], int64_t k0) {
        if (tile_full && k0 + BK <= K) {
#pragma unroll
            for (int e = 0; e < AE; ++e) xa[e] = ld_at(Aslab, aoff[e]);
#pragma unroll
            for (int e = 0; e < BE; ++e) xb[e] = ld_at(Bslab, boff[e]);
        } else {
            // edge tile or last slab: 32-bit limits, uniform over the workgroup
            const int ilim = (int)min((int64_t)BM, m - row0), jlim = (int)min((int64_t)BN, n - col0);
            const int klim = (int)min((int64_t)BK, K - k0);
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                int i, k;
                a_el(e, i, k);
                xa[e] = (i < ilim && k < klim) ? ld_at(Aslab, aoff[e]) : TD(0);
            }
#pragma unroll
            for (int e = 0; e < BE; ++e) {
                int j, k;
                b_el(e, j, k);
                xb[e] = (j < jlim && k < klim) ? ld_at(Bslab, boff[e]) : TD(0);
            }
        }
        Aslab += astep;
        Bslab += bstep;
    };
    auto store_from = [&](int buf, const TD (&xa)[AE], const TD (&xb)[BE]) {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            int i, k;
            a_el(e, i, k);
            As[buf][i][k] = (TC)xa[e];
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            int j, k;
            b_el(e, j, k);
            Bs[buf][j][k] = (TC)xb[e];
        }
    };
    auto load_slab = [&](int64_t k0) { load_into(ra, rb, k0); };
    auto store_slab = [&](int buf) { store_from(buf, ra, rb); };

    const int64_t nk = (K - kbeg + BK - 1) / BK;
    // Fragments are read from LDS as one 16-byte vector per lane and operand: KV = 2
    // (FP64) or 4 (FP32) consecutive k of one row.  Lane group g = lane/16 holds
    // k = kk + KV·g + s for MFMA step s < KV — the same map for A and B, so each step's
    // 4-deep MFMA sums 4 distinct k and a round of KV steps covers KC = 4·KV k.  Reading
    // 16 rather than 32 bytes per lane halves the live FP64 fragment registers.
    constexpr int KV = 16 / (int)sizeof(TC), KC = 4 * KV;
    static_assert(BK % KC == 0, "K slab must hold whole fragment rounds");
    const int kq = KV * (lane >> 4);
    // Edge tiles skip the 16×16 fragments that lie wholly outside C (zero padding), so the
    // padding of a ragged tile costs whole 16-row/column fragments rather than whole
    // tiles; the test is uniform over the wave (FP64 interior tiles run without it).
    bool frag_ok[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
            frag_ok[i][j] = row0 + wm * WM + i * 16 < m && col0 + wn * WN + j * 16 < n;
    auto mma_slab = [&](int buf, auto edge) {
        constexpr bool EDGE = decltype(edge)::value;
#pragma unroll
        for (int kk = 0; kk < BK; kk += KC) {
            typedef TC vk_t __attribute__((ext_vector_type(KV)));
            vk_t af[FM], bfr[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                af[i] = *reinterpret_cast<const vk_t *>(&As[buf][wm * WM + i * 16 + (lane & 15)][kk + kq]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                bfr[j] = *reinterpret_cast<const vk_t *>(&Bs[buf][wn * WN + j * 16 + (lane & 15)][kk + kq]);
#pragma unroll
            for (int s = 0; s < KV; ++s)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        if (!EDGE || frag_ok[i][j]) acc[i][j] = MfmaK4<TC>::mma(af[i][s], bfr[j][s], acc[i][j]);
        }
    };
    auto run = [&](auto edge) {
    if constexpr (NBUF == 2) {
        if (nk > 0) {
            load_slab(kbeg);
            store_slab(0);
        }
        __syncthreads();
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int buf = (int)(kt & 1);
            if (kt + 1 < nk) load_slab(kbeg + (kt + 1) * BK);
            mma_slab(buf, edge);
            if (kt + 1 < nk) store_slab(buf ^ 1);
            __syncthreads();
        }
    } else if constexpr (NBUF == 1) {
        if (nk > 0) load_slab(kbeg);
        for (int64_t kt = 0; kt < nk; ++kt) {
            if (kt > 0) __syncthreads();  // every wave is done reading the previous slab
            store_slab(0);
            __syncthreads();
            if (kt + 1 < nk) load_slab(kbeg + (kt + 1) * BK);  // lands while the MFMAs run
            mma_slab(0, edge);
        }
    } else {
        // one LDS buffer, two register stages: loads run two slabs ahead of the MFMAs
        TD ra1[AE], rb1[BE];
        if (nk > 0) load_into(ra, rb, kbeg);
        if (nk > 1) load_into(ra1, rb1, kbeg + BK);
        for (int64_t kt = 0; kt < nk; kt += 2) {
            if (kt > 0) __syncthreads();
            store_from(0, ra, rb);
            __syncthreads();
            if (kt + 2 < nk) load_into(ra, rb, kbeg + (kt + 2) * BK);
            mma_slab(0, edge);
            if (kt + 1 < nk) {
                __syncthreads();
                store_from(0, ra1, rb1);
                __syncthreads();
                if (kt + 3 < nk) load_into(ra1, rb1, kbeg + (kt + 3) * BK);
                mma_slab(0, edge);
            }
        }
    }
    };
    // FP64: interior tiles run a copy of the loop without the fragment test; FP32 keeps
    // one loop with the test (a second copy costs its kernels a wave per SIMD)
    if (sizeof(TC) == 4 || !tile_full)
        run(std::true_type{});
    else
        run(std::false_type{});

    // epilogue: C = α·acc + β·C, or this split's raw partial product into the workspace
    auto emit = [&](auto *Cg, int64_t ldc, double a_, bool has_beta) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int64_t col = col0 + wn * WN + j * 16 + (lane & 15);
                if (col >= n) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t row = row0 + wm * WM + i * 16 + MfmaK4<TC>::row(lane, v);
                    if (row < m) put_c(Cg + row + col * ldc, a_ * (double)acc[i][j][v], beta, has_beta);
                }
            }
    };
    const bool partial = G.splits > 1;
    if constexpr (std::is_same<TD, double>::value)
        emit(partial ? ws + G.ws_off + z * m * n : C + G.c_off, partial ? m : G.ldc, partial ? 1.0 : alpha,
             !partial && beta != 0.0);
    else if (partial)
        emit(ws + G.ws_off + z * m * n, m, 1.0, false);
    else
        emit(C + G.c_off, G.ldc, alpha, beta != 0.0);
}

template <typename TC, typename TD, int BM, int BN, int BK>
__global__ __launch_bounds__(NT) void gemm_k4_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                     int64_t ntiles, const TD *__restrict__ A,
                                                     const TD *__restrict__ Bm, TD *__restrict__ C,
                                                     double *__restrict__ ws, double alpha, double beta) {
    gemm_k4_body<TC, TD, BM, BN, BK, 2>(groups, ngroups, ntiles, A, Bm, C, ws, alpha, beta);
}

// 64×64×16 tiles fit 4 waves per SIMD (≤ 128 VGPR + AGPR) without spills when asked to
template <typename TC, typename TD, int BM, int BN, int BK>
__global__ __launch_bounds__(NT, (BM * BN <= 64 * 64 && BK <= 16) ? 4 : 1) void gemm_k4s_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                      int64_t ntiles, const TD *__restrict__ A,
                                                      const TD *__restrict__ Bm, TD *__restrict__ C,
                                                      double *__restrict__ ws, double alpha, double beta) {
    gemm_k4_body<TC, TD, BM, BN, BK, 1>(groups, ngroups, ntiles, A, Bm, C, ws, alpha, beta);
}

template <typename TC, typename TD, int BM, int BN, int BK>
__global__ __launch_bounds__(NT) void gemm_k4p_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                      int64_t ntiles, const TD *__restrict__ A,
                                                      const TD *__restrict__ Bm, TD *__restrict__ C,
                                                      double *__restrict__ ws, double alpha, double beta) {
    gemm_k4_body<TC, TD, BM, BN, BK, 3>(groups, ngroups, ntiles, A, Bm, C, ws, alpha, beta);
}

// --------------------------------------------------------------------------------
// FP64 kernel with direct global → LDS loads (A and B both untransposed: A m×k and B k×n
// column-major, the blocks of plain TensorMaps).  Slabs move with global_load_lds_dwordx4:
// each lane fetches 16 bytes that land at LDS base + 16·lane, so no VGPR staging and no LDS
// store instructions; two LDS stages let the loads of slab kt + 1 run while slab kt is
// multiplied, with one barrier per slab.  A 64×64 tile, 4 waves (2×2), 16-deep slabs;
// 16 KB of LDS per stage and 80 VGPRs, five workgroups per CU.  (A third stage measured
// slower: it leaves room for three workgroups per CU.)
//
// LDS layouts of a stage (16-byte chunks):
//  A: chunk (k, p) = rows 2p, 2p+1 of column k (contiguous in memory), at slot 32k + p.  A
//     wave's two 16-row fragments take the even and the odd rows of its 32 rows, so one
//     chunk read gives a lane its A value for both fragments.
//  B: chunk (q, j) = rows 2q, 2q+1 (k) of column j, at slot 8j + (q ^ (j & 7)): a column's
//     16 k stay contiguous (coalesced fetch) and the 16 lanes that read one k pair of 16
//     columns hit distinct banks.
// Edge tiles load directly too, from clamped addresses: a B column past n reads column
// n − 1, an A row pair at or past m − 1 reads rows (m − 1, m); what lands in rows or
// columns past C only reaches outputs that are not stored.  Row m of column k lies inside
// A's block for every k but the group's last (lda ≥ m), so on a tile past m the slab that
// holds the group's last column goes through registers, as does a split's partial last
// slab (zero filled past K).  Byte offsets inside a slab are 32-bit (the host picks this
// kernel only when they fit).
constexpr int DL_CH = 64 * 16 / 2;  // 16-byte chunks per operand slab (512)

__global__ __launch_bounds__(NT, 4) void gemm_k4d_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                         int64_t ntiles, const double *__restrict__ A,
                                                         const double *__restrict__ Bm, double *__restrict__ C,
                                                         double *__restrict__ ws, double alpha, double beta) {
    constexpr int BM = 64, BN = 64, BK = 16;
    typedef double d2v __attribute__((ext_vector_type(2)));
    __shared__ __attribute__((aligned(16))) double Ls[2][2][2 * DL_CH];

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t lda = G.lda, ldb = G.ldb;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int r16 = lane & 15, g4 = lane >> 4;
    const bool m_edge = row0 + BM > m, n_edge = col0 + BN > n;
    const int nfull = (int)((K - kbeg) / BK);  // whole slabs
    const bool tail = (K - kbeg) % BK != 0;    // a last partial slab
    const int nslab = nfull + (tail ? 1 : 0);
    const int ndirect = (m_edge && !tail && K == G.k) ? max(nfull - 1, 0) : nfull;
    const double *__restrict__ Aslab = A + G.a_off + row0 + kbeg * lda;
    const double *__restrict__ Bslab = Bm + G.b_off + col0 * ldb + kbeg;

    d4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = d4_t{0, 0, 0, 0};

    // direct path: wave w issues chunks [64(2w + e), 64(2w + e) + 64) of A and of B, e < 2;
    // chunk c is A (k = c / 32, p = c % 32) and B (j = c / 8, q = (c % 8) ^ (j & 7))
    const int rlast = (int)min((int64_t)(BM - 1), m - 1 - row0), clast = (int)min((int64_t)(BN - 1), n - 1 - col0);
    uint32_t aoff[2], boff[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int c = 64 * (2 * wave + e) + lane;
        aoff[e] = (uint32_t)(8 * (min(2 * (c & 31), rlast) + (c >> 5) * lda));
        const int j = c >> 3, q = (c & 7) ^ (j & 7);
        boff[e] = (uint32_t)(8 * (2 * q + min(j, clast) * ldb));
    }
    auto issue = [&](int kt, int st) {
        const char *ab = reinterpret_cast<const char *>(Aslab + (int64_t)kt * BK * lda);
        const char *bb = reinterpret_cast<const char *>(Bslab + (int64_t)kt * BK);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            __builtin_amdgcn_global_load_lds(
                (const void *)(ab + aoff[e]),
                (__attribute__((address_space(3))) void *)&Ls[st][0][2 * 64 * (2 * wave + e)], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                (const void *)(bb + boff[e]),
                (__attribute__((address_space(3))) void *)&Ls[st][1][2 * 64 * (2 * wave + e)], 16, 0, 0);
        }
    };
    // register path: thread t fills chunks t and t + 256 of A and of B, zero past C's
    // rows/columns and past K
    auto fill = [&](int kt, int st) {
        const int64_t k0 = kbeg + (int64_t)kt * BK;
        const int ilim = rlast + 1, jlim = clast + 1;
        const int klim = (int)min((int64_t)BK, K - k0);
        const double *ab = Aslab + (int64_t)kt * BK * lda;
        const double *bb = Bslab + (int64_t)kt * BK;
        d2v va[2], vb[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int c = tid + e * NT;
            const int k = c >> 5, i = 2 * (c & 31);
            va[e].x = (i < ilim && k < klim) ? ab[i + k * lda] : 0.0;
            va[e].y = (i + 1 < ilim && k < klim) ? ab[i + 1 + k * lda] : 0.0;
            const int j = c >> 3, q = (c & 7) ^ (j & 7);
            vb[e].x = (j < jlim && 2 * q < klim) ? bb[2 * q + j * ldb] : 0.0;
            vb[e].y = (j < jlim && 2 * q + 1 < klim) ? bb[2 * q + 1 + j * ldb] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int c = tid + e * NT;
            *reinterpret_cast<d2v *>(&Ls[st][0][2 * c]) = va[e];
            *reinterpret_cast<d2v *>(&Ls[st][1][2 * c]) = vb[e];
        }
    };
    // fragments that hold some row and column of C (edge tiles skip the others, a test
    // uniform over the wave): fragment fi holds rows wm·32 + 2i + fi, fj columns
    // wn·32 + 16fj + j
    bool frag_ok[2][2];
#pragma unroll
    for (int fi = 0; fi < 2; ++fi)
#pragma unroll
        for (int fj = 0; fj < 2; ++fj) frag_ok[fi][fj] = wm * 32 + fi <= rlast && wn * 32 + fj * 16 <= clast;
    // one slab from stage st: two rounds of 8 k; in step s of a round, lane group g4 holds
    // k = 8·round + 2·g4 + s for A and B alike
    auto mma = [&](int st, auto edge) {
        constexpr bool EDGE = decltype(edge)::value;
        const double *As = Ls[st][0], *Bs = Ls[st][1];
#pragma unroll
        for (int rd = 0; rd < 2; ++rd) {
            d2v a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
                a[s] = *reinterpret_cast<const d2v *>(&As[2 * ((8 * rd + 2 * g4 + s) * 32 + wm * 16 + r16)]);
#pragma unroll
            for (int fj = 0; fj < 2; ++fj) {
                const int j = wn * 32 + fj * 16 + r16, q = 4 * rd + g4;
                b[fj] = *reinterpret_cast<const d2v *>(&Bs[2 * (8 * j + (q ^ (j & 7)))]);
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int fi = 0; fi < 2; ++fi)
#pragma unroll
                    for (int fj = 0; fj < 2; ++fj)
                        if (!EDGE || frag_ok[fi][fj])
                            acc[fi][fj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][fi], b[fj][s], acc[fi][fj], 0, 0, 0);
        }
    };
    auto run = [&](auto edge) {
        if (ndirect > 0) issue(0, 0);
        for (int kt = 0; kt < ndirect; ++kt) {
            // this wave's loads of slab kt have landed; the barrier then covers every
            // wave's, and every wave is done reading the stage that slab kt + 1 refills.
            // A plain barrier: __syncthreads' release fence would also wait for loads
            // issued later; the waves' LDS reads of the previous slab are complete (their
            // MFMAs consumed them), and the asm statements keep the compiler from moving
            // LDS accesses across.
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (kt + 1 < ndirect) issue(kt + 1, (kt + 1) & 1);
            mma(kt & 1, edge);
        }
        for (int kt = ndirect; kt < nslab; ++kt) {
            __syncthreads();
            fill(kt, kt & 1);
            __syncthreads();
            mma(kt & 1, edge);
        }
    };
    if (m_edge || n_edge)
        run(std::true_type{});
    else
        run(std::false_type{});

    // epilogue: fragment fi holds rows wm·32 + 2i + fi, i = fragment row, so a lane holds
    // the two adjacent rows (2i, 2i + 1) of its column in acc[0] and acc[1]: one 16-byte
    // store writes both, and the 16 lanes of a row group cover 8 consecutive rows of 16
    // columns (whole 64-byte segments rather than every other row)
    const bool partial = G.splits > 1;
    double *Cg = partial ? ws + G.ws_off + z * m * n : C + G.c_off;
    const int64_t ldc = partial ? m : G.ldc;
    const double a_ = partial ? 1.0 : alpha;
    const bool has_beta = !partial && beta != 0.0;
#pragma unroll
    for (int fj = 0; fj < 2; ++fj) {
        const int64_t col = col0 + wn * 32 + fj * 16 + r16;
        if (col >= n) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = row0 + wm * 32 + 2 * (g4 + 4 * v);
            double *p = Cg + row + col * ldc;
            const double r0 = a_ * acc[0][fj][v], r1 = a_ * acc[1][fj][v];
            if (row + 1 < m) {
                d2v o;
                if (has_beta) {
                    d2v c;
                    __builtin_memcpy(&c, p, sizeof(c));
                    o = d2v{r0 + beta * c.x, r1 + beta * c.y};
                } else {
                    o = d2v{r0, r1};
                }
                __builtin_memcpy(p, &o, sizeof(o));
            } else if (row < m) {
                put_c(p, r0, beta, has_beta);
            }
        }
    }
}

// The direct-load kernel on 128×128, 128×64 and 64×128 tiles: the same slab layouts
// and edge rules as gemm_k4d_kernel, with 4 waves (2×2) each taking a (BM/2)×(BN/2) block
// of 16×16 fragments.  A's slab chunk (k, p) sits at slot (BM/2)k + p, and a wave's rows
// split into halves of 32, each read as the even and odd fragment pair; each LDS read
// and each loaded slab element feeds more MFMAs (LDS and L2 traffic per flop down by a
// quarter at 128×64, a half at 128×128), and the per-tile prologue and epilogue spread
// over more work, at three (128×64, 64×128: 48 KB of LDS) or two (128×128: 64 KB)
// workgroups per CU.  (The 64×64 tile keeps its own kernel: built from this template it
// measured 8–9% slower.)
template <int BM, int BN>
__global__ __launch_bounds__(NT, BM * BN > 8192 ? 2 : (BM * BN > 4096 ? 3 : 4)) void gemm_k4dw_kernel(
    const GGroup *__restrict__ groups, int64_t ngroups, int64_t ntiles, const double *__restrict__ A,
    const double *__restrict__ Bm, double *__restrict__ C, double *__restrict__ ws, double alpha, double beta) {
    static_assert((BM == 64 || BM == 128) && (BN == 64 || BN == 128) && BM * BN > 4096, "128-wide tiles");
    constexpr int BK = 16;
    constexpr int H = BM / 64;              // 32-row halves of a wave's rows
    constexpr int FN = BN / 32;             // 16-column fragments of a wave's columns
    constexpr int ACH = 8 * BM, BCH = 8 * BN;  // 16-byte chunks per slab of A and of B
    constexpr int AE = ACH / NT, BE = BCH / NT;  // chunks per lane (and 64-chunk blocks per wave)
    constexpr int AP = BM / 2;                 // A chunks per column k
    typedef double d2v __attribute__((ext_vector_type(2)));
    __shared__ __attribute__((aligned(16))) double La[2][2 * ACH];
    __shared__ __attribute__((aligned(16))) double Lb[2][2 * BCH];

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t lda = G.lda, ldb = G.ldb;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int r16 = lane & 15, g4 = lane >> 4;
    const bool m_edge = row0 + BM > m, n_edge = col0 + BN > n;
    const int nfull = (int)((K - kbeg) / BK);  // whole slabs
    const bool tail = (K - kbeg) % BK != 0;    // a last partial slab
    const int nslab = nfull + (tail ? 1 : 0);
    const int ndirect = (m_edge && !tail && K == G.k) ? max(nfull - 1, 0) : nfull;
    const double *__restrict__ Aslab = A + G.a_off + row0 + kbeg * lda;
    const double *__restrict__ Bslab = Bm + G.b_off + col0 * ldb + kbeg;

    d4_t acc[H][2][FN];
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[h][i][j] = d4_t{0, 0, 0, 0};

    // direct path: wave w issues the 64-chunk blocks AE·w + e of A (e < AE) and BE·w + e of
    // B; chunk c is A (k = c / AP, p = c % AP) and B (j = c / 8, q = (c % 8) ^ (j & 7))
    const int rlast = (int)min((int64_t)(BM - 1), m - 1 - row0), clast = (int)min((int64_t)(BN - 1), n - 1 - col0);
    uint32_t aoff[AE], boff[BE];
#pragma unroll
    for (int e = 0; e < AE; ++e) {
        const int c = 64 * (AE * wave + e) + lane;
        aoff[e] = (uint32_t)(8 * (min(2 * (c % AP), rlast) + (c / AP) * lda));
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
        const int c = 64 * (BE * wave + e) + lane;
        const int j = c >> 3, q = (c & 7) ^ (j & 7);
        boff[e] = (uint32_t)(8 * (2 * q + min(j, clast) * ldb));
    }
    auto issue = [&](int kt, int st) {
        const char *ab = reinterpret_cast<const char *>(Aslab + (int64_t)kt * BK * lda);
        const char *bb = reinterpret_cast<const char *>(Bslab + (int64_t)kt * BK);
#pragma unroll
        for (int e = 0; e < AE; ++e)
            __builtin_amdgcn_global_load_lds((const void *)(ab + aoff[e]),
                                             (__attribute__((address_space(3))) void *)&La[st][2 * 64 * (AE * wave + e)],
                                             16, 0, 0);
#pragma unroll
        for (int e = 0; e < BE; ++e)
            __builtin_amdgcn_global_load_lds((const void *)(bb + boff[e]),
                                             (__attribute__((address_space(3))) void *)&Lb[st][2 * 64 * (BE * wave + e)],
                                             16, 0, 0);
    };
    // register path: thread t fills chunks t + e·NT of A (e < AE) and of B (e < BE), zero
    // past C's rows/columns and past K
    auto fill = [&](int kt, int st) {
        const int64_t k0 = kbeg + (int64_t)kt * BK;
        const int ilim = rlast + 1, jlim = clast + 1;
        const int klim = (int)min((int64_t)BK, K - k0);
        const double *ab = Aslab + (int64_t)kt * BK * lda;
        const double *bb = Bslab + (int64_t)kt * BK;
        d2v va[AE], vb[BE];
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            const int c = tid + e * NT;
            const int k = c / AP, i = 2 * (c % AP);
            va[e].x = (i < ilim && k < klim) ? ab[i + k * lda] : 0.0;
            va[e].y = (i + 1 < ilim && k < klim) ? ab[i + 1 + k * lda] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int c = tid + e * NT;
            const int j = c >> 3, q = (c & 7) ^ (j & 7);
            vb[e].x = (j < jlim && 2 * q < klim) ? bb[2 * q + j * ldb] : 0.0;
            vb[e].y = (j < jlim && 2 * q + 1 < klim) ? bb[2 * q + 1 + j * ldb] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < AE; ++e) *reinterpret_cast<d2v *>(&La[st][2 * (tid + e * NT)]) = va[e];
#pragma unroll
        for (int e = 0; e < BE; ++e) *reinterpret_cast<d2v *>(&Lb[st][2 * (tid + e * NT)]) = vb[e];
    };
    // fragments that hold some row and column of C (edge tiles skip the others, a test
    // uniform over the wave): fragment (h, fi) holds rows wm·BM/2 + 32h + 2i + fi, fj
    // columns wn·BN/2 + 16fj + j
    bool frag_ok[H][2][FN];
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int fi = 0; fi < 2; ++fi)
#pragma unroll
            for (int fj = 0; fj < FN; ++fj)
                frag_ok[h][fi][fj] = wm * (BM / 2) + 32 * h + fi <= rlast && wn * (BN / 2) + fj * 16 <= clast;
    // one slab from stage st: two rounds of 8 k; in step s of a round, lane group g4 holds
    // k = 8·round + 2·g4 + s for A and B alike
    auto mma = [&](int st, auto edge) {
        constexpr bool EDGE = decltype(edge)::value;
        const double *As = La[st], *Bs = Lb[st];
#pragma unroll
        for (int rd = 0; rd < 2; ++rd) {
            d2v a[H][2], b[FN];
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    a[h][s] = *reinterpret_cast<const d2v *>(
                        &As[2 * ((8 * rd + 2 * g4 + s) * AP + wm * (BM / 4) + 16 * h + r16)]);
#pragma unroll
            for (int fj = 0; fj < FN; ++fj) {
                const int j = wn * (BN / 2) + fj * 16 + r16, q = 4 * rd + g4;
                b[fj] = *reinterpret_cast<const d2v *>(&Bs[2 * (8 * j + (q ^ (j & 7)))]);
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int h = 0; h < H; ++h)
#pragma unroll
                    for (int fi = 0; fi < 2; ++fi)
#pragma unroll
                        for (int fj = 0; fj < FN; ++fj)
                            if (!EDGE || frag_ok[h][fi][fj])
                                acc[h][fi][fj] =
                                    __builtin_amdgcn_mfma_f64_16x16x4f64(a[h][s][fi], b[fj][s], acc[h][fi][fj], 0, 0, 0);
        }
    };
    auto run = [&](auto edge) {
        if (ndirect > 0) issue(0, 0);
        for (int kt = 0; kt < ndirect; ++kt) {
            // this wave's loads of slab kt have landed; the barrier then covers every
            // wave's, and every wave is done reading the stage that slab kt + 1 refills.
            // A plain barrier: __syncthreads' release fence would also wait for loads
            // issued later; the waves' LDS reads of the previous slab are complete (their
            // MFMAs consumed them), and the asm statements keep the compiler from moving
            // LDS accesses across.
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (kt + 1 < ndirect) issue(kt + 1, (kt + 1) & 1);
            mma(kt & 1, edge);
        }
        for (int kt = ndirect; kt < nslab; ++kt) {
            __syncthreads();
            fill(kt, kt & 1);
            __syncthreads();
            mma(kt & 1, edge);
        }
    };
    if (m_edge || n_edge)
        run(std::true_type{});
    else
        run(std::false_type{});

    // epilogue: fragment (h, fi) holds rows wm·BM/2 + 32h + 2i + fi, i = fragment row, so a
    // lane holds the two adjacent rows (2i, 2i + 1) of its column in acc[h][0] and
    // acc[h][1]: one 16-byte store writes both, and the 16 lanes of a row group cover 8
    // consecutive rows of 16 columns (whole 64-byte segments rather than every other row)
    const bool partial = G.splits > 1;
    double *Cg = partial ? ws + G.ws_off + z * m * n : C + G.c_off;
    const int64_t ldc = partial ? m : G.ldc;
    const double a_ = partial ? 1.0 : alpha;
    const bool has_beta = !partial && beta != 0.0;
#pragma unroll
    for (int fj = 0; fj < FN; ++fj) {
        const int64_t col = col0 + wn * (BN / 2) + fj * 16 + r16;
        if (col >= n) continue;
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = row0 + wm * (BM / 2) + 32 * h + 2 * (g4 + 4 * v);
                double *p = Cg + row + col * ldc;
                const double r0 = a_ * acc[h][0][fj][v], r1 = a_ * acc[h][1][fj][v];
                if (row + 1 < m) {
                    d2v o;
                    if (has_beta) {
                        d2v c;
                        __builtin_memcpy(&c, p, sizeof(c));
                        o = d2v{r0 + beta * c.x, r1 + beta * c.y};
                    } else {
                        o = d2v{r0, r1};
                    }
                    __builtin_memcpy(p, &o, sizeof(o));
                } else if (row < m) {
                    put_c(p, r0, beta, has_beta);
                }
            }
    }
}

// Calibration of the FP64 matrix-core rate: `iters` rounds of 16 v_mfma_f64_16x16x4_f64
// per wave on four independent accumulators, no memory traffic but one store per lane
// at the end (out[blockIdx·NT + thread], which keeps the work live).
__global__ __launch_bounds__(NT) void mfma_f64_calib_kernel(double *__restrict__ out, int iters, double x) {
    d4_t acc[4] = {};
    const double a = x + threadIdx.x, b = x - threadIdx.x;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    double t = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) t += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[(int64_t)blockIdx.x * NT + threadIdx.x] = t;
}

// --------------------------------------------------------------------------------
// FP32 kernel with direct global → LDS loads (FP32 data, A and B untransposed): the
// FP32 counterpart of gemm_k4d_kernel.  A 16-byte chunk holds 4 floats, so a wave takes a
// 64×16 strip of the 64×64 tile (the four waves side by side along the columns): its four
// 16-row fragments take the rows ≡ 0, 1, 2, 3 mod 4, and one chunk read (rows 4i … 4i + 3
// of one k) gives a lane its A value for all four; one B chunk (k … k + 3 of one column)
// gives it the four MFMA steps of a 16-deep slab.
//  A: chunk (k, p) = rows 4p … 4p+3 of column k, at slot 16k + p;
//  B: chunk (q, j) = k 4q … 4q+3 of column j, at slot 4j + (q ^ ((j / 4) mod 4)).
// Two LDS stages of 8 KB, one barrier per slab.  Tiles past n load clamped columns;
// tiles past m, and a split's partial last slab, go through registers with zero fill.
// The MFMA steps are the one-buffer FP32 kernel's (the same sums in the same k order).
__global__ __launch_bounds__(NT) void gemm_k4d32_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                        int64_t ntiles, const float *__restrict__ A,
                                                        const float *__restrict__ Bm, float *__restrict__ C,
                                                        double *__restrict__ ws, double alpha, double beta) {
    constexpr int BM = 64, BN = 64, BK = 16, CH = 256;  // CH: 16-byte chunks per operand slab
    __shared__ __attribute__((aligned(16))) f4_t Ls[2][2][CH];

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t lda = G.lda, ldb = G.ldb;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    const bool m_edge = row0 + BM > m;
    const int nfull = (int)((K - kbeg) / BK);
    const int nslab = nfull + ((K - kbeg) % BK != 0 ? 1 : 0);
    const int ndirect = m_edge ? 0 : nfull;
    const float *__restrict__ Aslab = A + G.a_off + row0 + kbeg * lda;
    const float *__restrict__ Bslab = Bm + G.b_off + col0 * ldb + kbeg;
    const int rlast = (int)min((int64_t)(BM - 1), m - 1 - row0), clast = (int)min((int64_t)(BN - 1), n - 1 - col0);

    f4_t acc[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = f4_t{0, 0, 0, 0};

    // direct path: wave w issues chunks [64w, 64w + 64) of A and of B
    const int c = 64 * wave + lane;
    const uint32_t aoff = (uint32_t)(4 * (4 * (c & 15) + (c >> 4) * lda));
    const int bj = c >> 2, bq = (c & 3) ^ ((bj >> 2) & 3);
    const uint32_t boff = (uint32_t)(4 * (4 * bq + min(bj, clast) * ldb));
    auto issue = [&](int kt, int st) {
        const char *ab = reinterpret_cast<const char *>(Aslab + (int64_t)kt * BK * lda);
        const char *bb = reinterpret_cast<const char *>(Bslab + (int64_t)kt * BK);
        __builtin_amdgcn_global_load_lds((const void *)(ab + aoff),
                                         (__attribute__((address_space(3))) void *)&Ls[st][0][64 * wave], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void *)(bb + boff),
                                         (__attribute__((address_space(3))) void *)&Ls[st][1][64 * wave], 16, 0, 0);
    };
    // register path: thread t fills chunk t of A and of B, zero past C and past K
    auto fill = [&](int kt, int st) {
        const int64_t k0 = kbeg + (int64_t)kt * BK;
        const int klim = (int)min((int64_t)BK, K - k0);
        const float *ab = Aslab + (int64_t)kt * BK * lda;
        const float *bb = Bslab + (int64_t)kt * BK;
        const int k = tid >> 4, i = 4 * (tid & 15);
        f4_t va, vb;
#pragma unroll
        for (int e = 0; e < 4; ++e) va[e] = (i + e <= rlast && k < klim) ? ab[i + e + k * lda] : 0.0f;
        const int j = tid >> 2, q = (tid & 3) ^ ((j >> 2) & 3);
#pragma unroll
        for (int e = 0; e < 4; ++e) vb[e] = (j <= clast && 4 * q + e < klim) ? bb[4 * q + e + j * ldb] : 0.0f;
        Ls[st][0][tid] = va;
        Ls[st][1][tid] = vb;
    };
    // one slab: step s of the MFMA sums k = 4·g4 + s for A and B alike
    auto mma = [&](int st) {
        const int j = 16 * wave + r16;
        const f4_t b = Ls[st][1][4 * j + (g4 ^ ((j >> 2) & 3))];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const f4_t a = Ls[st][0][16 * (4 * g4 + s) + r16];
#pragma unroll
            for (int f = 0; f < 4; ++f) acc[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[f], b[s], acc[f], 0, 0, 0);
        }
    };
    if (ndirect > 0) issue(0, 0);
    for (int kt = 0; kt < ndirect; ++kt) {
        // as in gemm_k4d_kernel: this wave's loads landed, then a plain barrier
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (kt + 1 < ndirect) issue(kt + 1, (kt + 1) & 1);
        mma(kt & 1);
    }
    for (int kt = ndirect; kt < nslab; ++kt) {
        __syncthreads();
        fill(kt, kt & 1);
        __syncthreads();
        mma(kt & 1);
    }

    // epilogue: fragment f holds rows 4i + f, and the lane's value v of every fragment is
    // fragment row 4·g4 + v, so the four fragments give rows 16·g4 + 4v … + 3 of its
    // column: one 16-byte store (FP32 C) or four partial sums (FP64 workspace)
    const bool partial = G.splits > 1;
    const int64_t col = col0 + 16 * wave + r16;
    if (col >= n) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t row = row0 + 16 * g4 + 4 * v;
        if (row >= m) continue;
        if (partial) {
            double *p = ws + G.ws_off + z * m * n + row + col * m;
#pragma unroll
            for (int f = 0; f < 4; ++f)
                if (row + f < m) p[f] = (double)acc[f][v];
        } else {
            float *p = C + G.c_off + row + col * G.ldc;
            const bool has_beta = beta != 0.0;
            if (row + 3 < m) {
                f4_t o;
                f4_t cv = f4_t{0, 0, 0, 0};
                if (has_beta) __builtin_memcpy(&cv, p, sizeof(cv));
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const double r = alpha * (double)acc[f][v];
                    o[f] = (float)(has_beta ? r + beta * (double)cv[f] : r);
                }
                __builtin_memcpy(p, &o, sizeof(o));
            } else {
#pragma unroll
                for (int f = 0; f < 4; ++f)
                    if (row + f < m) put_c(p + f, alpha * (double)acc[f][v], beta, has_beta);
            }
        }
    }
}

// --------------------------------------------------------------------------------
// Persistent FP64 / FP32 kernel.  The grid is the device's workgroup slots (occupancy ×
// CUs); workgroup w computes the tiles sched[slot_start[w] .. slot_start[w+1]) in order, a
// longest-processing-time assignment of every tile to a slot made on the host, so the
// launch has no tail of half-empty rounds and no workgroup starts cold.  The (tile, K
// slab) steps of a workgroup form one pipeline: the global loads of the next step — the
// next slab of this tile or the first slab of the next tile — are issued before the
// MFMAs of the current step, so a tile's first slab arrives while the previous tile is
// still being multiplied.  One LDS buffer, the next step waits in registers (two
// barriers per step: the staging of gemm_k4s_kernel, four workgroups per CU at 64×64).
// Measured against gemm_k4s_kernel (64×64×16, FP64): 58 against 63 TFLOP/s at 4096³,
// 48 against 49 on 16 sectors of 2048×2048×290, 107 against 100 µs on the two-site H_eff
// L·x launch: the host schedule removes the tail and the prefetch the cold start of each
// tile, but neither is what bounds these launches, so the autotuner keeps gemm_k4s_kernel.
// --------------------------------------------------------------------------------
template <typename TC, typename TD, int BM, int BN, int BK>
__device__ __forceinline__ void gemm_pk_body(const GGroup *__restrict__ groups, int64_t ngroups,
                                             const int32_t *__restrict__ sched,
                                             const int32_t *__restrict__ slot_start, const TD *__restrict__ A,
                                             const TD *__restrict__ Bm, TD *__restrict__ C,
                                             double *__restrict__ ws, double alpha, double beta) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int AE = BM * BK / NT;
    constexpr int BE = BN * BK / NT;
    constexpr int KP = BK + 16 / (int)sizeof(TC);
    __shared__ __attribute__((aligned(16))) TC As[1][BM][KP];
    __shared__ __attribute__((aligned(16))) TC Bs[1][BN][KP];
    using acc_t = typename MfmaK4<TC>::acc_t;

    const int32_t s0 = slot_start[blockIdx.x], s1 = slot_start[blockIdx.x + 1];
    if (s0 >= s1) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // the tile of a step: its group's fields (copied by value: read through the
    // restrict-qualified table they stay scalar loads, while a pointer kept across steps
    // would turn them into vector loads whose waits also wait for the prefetch), its
    // corner, K range [kbeg, K) and slab count (a K = 0 group runs one all-zero slab, so
    // its epilogue still applies β)
    struct Tile {
        int64_t m, n, a_off, lda, b_off, ldb, c_off, ldc, ws_off;
        int ta, tb;
        bool partial;
        int64_t row0, col0, kbeg, K, z;
        int nk;
        bool full;
    };
    auto decode = [&](int32_t pos) {
        Tile t;
        const int64_t tile = sched[pos];
        const int gi = find_group(groups, ngroups, tile);
        const int64_t lt = tile - groups[gi].tile_start;
        const int64_t tiles_n = groups[gi].tiles_n;
        const int64_t tiles_mn = groups[gi].tiles_m * tiles_n;
        t.m = groups[gi].m;
        t.n = groups[gi].n;
        t.a_off = groups[gi].a_off;
        t.lda = groups[gi].lda;
        t.b_off = groups[gi].b_off;
        t.ldb = groups[gi].ldb;
        t.c_off = groups[gi].c_off;
        t.ldc = groups[gi].ldc;
        t.ws_off = groups[gi].ws_off;
        t.ta = groups[gi].ta;
        t.tb = groups[gi].tb;
        t.partial = groups[gi].splits > 1;
        const int64_t kchunk = groups[gi].kchunk;
        t.z = lt / tiles_mn;
        const int64_t lmn = lt - t.z * tiles_mn;
        t.row0 = (lmn / tiles_n) * BM;
        t.col0 = (lmn % tiles_n) * BN;
        t.kbeg = t.z * kchunk;
        t.K = min(groups[gi].k, t.kbeg + kchunk);
        t.nk = (int)max((int64_t)1, (t.K - t.kbeg + BK - 1) / BK);
        t.full = (t.row0 + BM <= t.m) && (t.col0 + BN <= t.n);
        return t;
    };
    // element e of this thread: (row i, k) of the A slab, (k, column j) of the B slab;
    // global order is coalesced (contiguous index fastest), the LDS position [row][k]
    auto a_el = [&](int ta, int e, int &i, int &k) {
        const int idx = e * NT + tid;
        if (!ta) { i = idx % BM; k = idx / BM; }
        else     { k = idx % BK; i = idx / BK; }
    };
    auto b_el = [&](int tb, int e, int &j, int &k) {
        const int idx = e * NT + tid;
        if (!tb) { k = idx % BK; j = idx / BK; }
        else     { j = idx % BN; k = idx / BN; }
    };

    // load side: the tile whose slabs are being fetched, its per-thread byte offsets from
    // the slab base (set once per tile) and the registers the slab lands in; addresses
    // are a scalar slab base plus a 32-bit offset, and interior slabs skip every bounds
    // test (as in gemm_k4s_kernel)
    Tile L;
    int32_t lpos = s0;
    int lkt = 0;
    uint32_t aoff[AE], boff[BE];
    auto set_offsets = [&]() {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            int i, k;
            a_el(L.ta, e, i, k);
            aoff[e] = (uint32_t)(sizeof(TD) * (L.ta ? (k + i * L.lda) : (i + k * L.lda)));
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            int j, k;
            b_el(L.tb, e, j, k);
            boff[e] = (uint32_t)(sizeof(TD) * (L.tb ? (j + k * L.ldb) : (k + j * L.ldb)));
        }
    };
    auto ld_at = [](const TD *base, uint32_t off) {
        return *reinterpret_cast<const TD *>(reinterpret_cast<const char *>(base) + off);
    };
    TD ra[AE], rb[BE];
    auto load_step = [&]() {
        const int64_t k0 = L.kbeg + (int64_t)lkt * BK;
        const TD *Aslab = A + L.a_off + (L.ta ? L.row0 * L.lda + k0 : L.row0 + k0 * L.lda);
        const TD *Bslab = Bm + L.b_off + (L.tb ? L.col0 + k0 * L.ldb : L.col0 * L.ldb + k0);
        if (L.full && k0 + BK <= L.K) {
#pragma unroll
            for (int e = 0; e < AE; ++e) ra[e] = ld_at(Aslab, aoff[e]);
#pragma unroll
            for (int e = 0; e < BE; ++e) rb[e] = ld_at(Bslab, boff[e]);
        } else {
            const int ilim = (int)min((int64_t)BM, L.m - L.row0), jlim = (int)min((int64_t)BN, L.n - L.col0);
            const int klim = (int)max((int64_t)0, min((int64_t)BK, L.K - k0));
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                int i, k;
                a_el(L.ta, e, i, k);
                ra[e] = (i < ilim && k < klim) ? ld_at(Aslab, aoff[e]) : TD(0);
            }
#pragma unroll
            for (int e = 0; e < BE; ++e) {
                int j, k;
                b_el(L.tb, e, j, k);
                rb[e] = (j < jlim && k < klim) ? ld_at(Bslab, boff[e]) : TD(0);
            }
        }
        ++lkt;
    };
    // the registers hold a step of tile T: stored with T's operand layouts
    auto store_step = [&](int buf, const Tile &Tt) {
        const int ta = Tt.ta, tb = Tt.tb;
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            int i, k;
            a_el(ta, e, i, k);
            As[buf][i][k] = (TC)ra[e];
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            int j, k;
            b_el(tb, e, j, k);
            Bs[buf][j][k] = (TC)rb[e];
        }
    };
    // MFMA side: the tile being multiplied
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    constexpr int KV = 16 / (int)sizeof(TC), KC = 4 * KV;
    static_assert(BK % KC == 0, "K slab must hold whole fragment rounds");
    const int kq = KV * (lane >> 4);
    // One MFMA path for every tile: the fragments of an edge tile that lie outside C
    // multiply the zeros its masked loads left in LDS, and the epilogue stores nothing
    // outside C.  (Skipping them needs a second copy of the step loop, and the merge of
    // its accumulator registers with the first costs a copy of every accumulator behind
    // the last MFMA of each step.)
    auto mma_slab = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += KC) {
            typedef TC vk_t __attribute__((ext_vector_type(KV)));
            vk_t af[FM], bfr[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                af[i] = *reinterpret_cast<const vk_t *>(&As[buf][wm * WM + i * 16 + (lane & 15)][kk + kq]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                bfr[j] = *reinterpret_cast<const vk_t *>(&Bs[buf][wn * WN + j * 16 + (lane & 15)][kk + kq]);
#pragma unroll
            for (int s = 0; s < KV; ++s)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = MfmaK4<TC>::mma(af[i][s], bfr[j][s], acc[i][j]);
        }
    };
    auto emit = [&](const Tile &M, auto *Cg, int64_t ldc, double a_, bool has_beta) {
        // the tile's corner on the scalar side, rows and columns inside it 32-bit
        auto *Ct = Cg + M.row0 + M.col0 * ldc;
        const int rlim = (int)min((int64_t)BM, M.m - M.row0), clim = (int)min((int64_t)BN, M.n - M.col0);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = wn * WN + j * 16 + (lane & 15);
                if (col < clim) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = wm * WM + i * 16 + MfmaK4<TC>::row(lane, v);
                        if (row < rlim)
                            put_c(Ct + row + (int64_t)col * ldc, a_ * (double)acc[i][j][v], beta, has_beta);
                    }
                }
                acc[i][j] = acc_t{0, 0, 0, 0};
            }
    };
    auto epilogue = [&](const Tile &M) {
        const bool partial = M.partial;
        if constexpr (std::is_same<TD, double>::value)
            emit(M, partial ? ws + M.ws_off + M.z * M.m * M.n : C + M.c_off, partial ? M.m : M.ldc,
                 partial ? 1.0 : alpha, !partial && beta != 0.0);
        else if (partial)
            emit(M, ws + M.ws_off + M.z * M.m * M.n, M.m, 1.0, false);
        else
            emit(M, C + M.c_off, M.ldc, alpha, beta != 0.0);
    };

    L = decode(s0);
    set_offsets();
    load_step();  // the first step's slab
    Tile M = L;
    int32_t mpos = s0;
    {
        // per tile: a tight loop over all but its last slab (the next slab of the same tile
        // is fetched while the current one is multiplied), then the last slab, during which
        // the first slab of the next tile is fetched
        for (bool first = true;; first = false) {
            for (int kt = 0; kt < M.nk - 1; ++kt) {
                if (!(first && kt == 0)) __syncthreads();  // every wave is done reading the slab
                store_step(0, M);
                __syncthreads();
                load_step();
                mma_slab(0);
            }
            if (!(first && M.nk == 1)) __syncthreads();
            store_step(0, M);
            __syncthreads();
            const bool more = ++mpos < s1;
            if (more) {
                L = decode(mpos);
                set_offsets();
                lkt = 0;
                load_step();
            }
            mma_slab(0);
            epilogue(M);
            if (!more) break;
            M = L;
        }
    }
}

template <typename TC, typename TD, int BM, int BN, int BK>
__global__ __launch_bounds__(NT, 4) void gemm_pks_kernel(
    const GGroup *__restrict__ groups, int64_t ngroups, const int32_t *__restrict__ sched,
    const int32_t *__restrict__ slot_start, const TD *__restrict__ A, const TD *__restrict__ Bm,
    TD *__restrict__ C, double *__restrict__ ws, double alpha, double beta) {
    static_assert(BM * BN <= 64 * 64, "four workgroups per CU: 64×64 tiles");
    gemm_pk_body<TC, TD, BM, BN, BK>(groups, ngroups, sched, slot_start, A, Bm, C, ws, alpha, beta);
}

// Split-K reduction (and, with ws = nullptr, C = β·C for α = 0).  Each group's elements
// start at a multiple of RCH in the element numbering (elem_start), so a workgroup's RCH
// consecutive elements lie in one group, found by one search per workgroup; each thread
// sums RPT elements, partials in fixed z order, loads issued before the stores.  Column-
// major C blocks with ldc = m take the element index as the offset; others divide in 32
// bits when the block has fewer than 2^31 elements.
constexpr int RPT = 4;
constexpr int64_t RCH = (int64_t)RPT * NT;
inline int64_t rch_round(int64_t e) { return (e + RCH - 1) / RCH * RCH; }

template <typename TD>
__global__ __launch_bounds__(NT) void splitk_reduce(const RGroup *__restrict__ rg, int64_t ng, int64_t total,
                                                    const double *__restrict__ ws, TD *__restrict__ C,
                                                    double alpha, double beta) {
    const bool has_beta = (beta != 0.0);
    const int64_t nchunks = total / RCH;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int64_t e0 = ch * RCH;
        int64_t lo = 0, hi = ng - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rg[mid].elem_start <= e0) lo = mid; else hi = mid - 1;
        }
        const RGroup &R = rg[lo];
        const int64_t mn = R.m * R.n, base = e0 - R.elem_start;
        double acc[RPT];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t le = base + q * NT + threadIdx.x;
            acc[q] = 0.0;
            if (ws && le < mn)
                for (int64_t zz = 0; zz < R.splits; ++zz) acc[q] += ws[R.ws_off + zz * mn + le];
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t le = base + q * NT + threadIdx.x;
            if (le >= mn) break;
            int64_t off = le;
            if (R.ldc != R.m) {
                if (mn < ((int64_t)1 << 31)) {
                    const uint32_t l32 = (uint32_t)le, m32 = (uint32_t)R.m;
                    off = (int64_t)(l32 % m32) + (int64_t)(l32 / m32) * R.ldc;
                } else {
                    off = le % R.m + (le / R.m) * R.ldc;
                }
            }
            put_c(C + R.c_off + off, alpha * acc[q], beta, has_beta);
        }
    }
}


// --------------------------------------------------------------------------------
// 3×BF16 split-product GEMM: x = x0 + x1 + x2 (bf16 each), products
// x0y0 (high), x0y1 + x1y0 + x0y2 + x1y1 + x2y0 (low) on v_mfma_f32_16x16x32_bf16,
// FP32 accumulation per BK=32 slab, high part folded into FP64 after every slab.
//
// split_pack_kernel splits every operand element once per call into the three bf16
// planes and packs each group's A rows and B columns K-contiguous, zero padded to
// whole tiles and slabs; the GEMM then stages 16-byte vectors straight into LDS with
// no conversion and no bounds test in its K loop.  One workgroup converts one
// PACK_T×PACK_T (rows × K) tile: the doubles are staged through LDS, read with lanes
// along whichever side is contiguous in memory and written as whole 8-element octets,
// PACK_T/8 lanes per row (128 contiguous bytes per plane), so both sides coalesce.
// --------------------------------------------------------------------------------
__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

typedef uint32_t u4_t __attribute__((ext_vector_type(4)));  // 16 bytes = 8 bf16

constexpr int PACK_T = 64;

// which: 0 = the A and the B tiles of every group, 1 = the A tiles only, 2 = the B tiles only.
// TD: element type of A and B (a float splits exactly: its 24-bit significand is three
// 8-bit bf16 significands)
template <typename TD>
__global__ __launch_bounds__(NT) void split_pack_kernel(const PackGroup *__restrict__ pg, int64_t ng,
                                                        const TD *__restrict__ A, const TD *__restrict__ Bm,
                                                        __bf16 *__restrict__ P, int which) {
    // staged in the element type (a float splits exactly in float arithmetic too): half
    // the LDS for FP32 data, twice the workgroups per CU
    typedef typename std::conditional<std::is_same<TD, float>::value, float, double>::type TS;
    __shared__ TS tile[PACK_T][PACK_T + 1];
    const int64_t q = blockIdx.x;
    auto first = [&](const PackGroup &g) { return which == 0 ? g.start : (which == 1 ? g.startA : g.startB); };
    int64_t lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (first(pg[mid]) <= q) lo = mid; else hi = mid - 1;
    }
    const PackGroup &G = pg[lo];
    int64_t r = q - first(G);
    bool isA = which == 1;
    if (which == 0) {
        isA = r < G.nA;
        if (!isA) r -= G.nA;
    }
    const int64_t tk = (G.kp + PACK_T - 1) / PACK_T;
    const int64_t i0 = (r / tk) * PACK_T, k0 = (r % tk) * PACK_T;
    const int64_t lim = isA ? G.m : G.n;
    const TD *src = isA ? A + G.a_off : Bm + G.b_off;
    const int64_t ld = isA ? G.lda : G.ldb;
    // element (i, k) of op(A) / of op(B)ᵀ lies at i + k·ld (rows contiguous) or k + i·ld
    const bool rows_contiguous = isA ? (G.ta == 0) : (G.tb != 0);
    const int t = threadIdx.x;
    if (rows_contiguous) {
        const int ii = t % PACK_T;
        const int64_t gi = i0 + ii;
        for (int kk = t / PACK_T; kk < PACK_T; kk += NT / PACK_T) {
            const int64_t gk = k0 + kk;
            tile[ii][kk] = (gi < lim && gk < G.k) ? (TS)src[gi + gk * ld] : (TS)0;
        }
    } else {
        const int kk = t % PACK_T;
        const int64_t gk = k0 + kk;
        for (int ii = t / PACK_T; ii < PACK_T; ii += NT / PACK_T) {
            const int64_t gi = i0 + ii;
            tile[ii][kk] = (gi < lim && gk < G.k) ? (TS)src[gk + gi * ld] : (TS)0;
        }
    }
    __syncthreads();
    __bf16 *base = P + (isA ? G.pa : G.pb);
    const int64_t plane = isA ? G.sa : G.sb;
    // rows: i0 + ii < mp (np), a multiple of 128; octets: k0 + 8·o < kp, a multiple of 32
    for (int w = t; w < PACK_T * (PACK_T / 8); w += NT) {
        const int ii = w / (PACK_T / 8), o = w % (PACK_T / 8);
        const int64_t gk = k0 + 8 * o;
        if (gk >= G.kp) continue;
        __bf16 x[3][8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const double v = (double)tile[ii][8 * o + e];
            const __bf16 x0 = to_bf16((float)v);
            const double r1 = v - (double)(float)x0;
            const __bf16 x1 = to_bf16((float)r1);
            const double r2 = r1 - (double)(float)x1;
            x[0][e] = x0;
            x[1][e] = x1;
            x[2][e] = to_bf16((float)r2);
        }
        __bf16 *dst = base + (i0 + ii) * G.kp + gk;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            u4_t v;
            __builtin_memcpy(&v, x[pl], sizeof(v));
            *reinterpret_cast<u4_t *>(dst + pl * plane) = v;
        }
    }
}

// 3×BF16 with direct global → LDS loads (64×64 tiles): the split panels are padded (rows to
// 128, K to 32, zero filled) and 16-byte aligned, so every slab loads whole with
// global_load_lds_dwordx4, no bounds test and no VGPR staging, into two LDS stages (one
// barrier per 32-deep slab).  A stage holds per operand and plane the 64 rows × 4 k
// octets of the slab, octet g of row r at slot 4r + (g ^ ((r / 4) mod 4)): a row's 64
// bytes stay contiguous for the fetch, and the 16 lanes that read one octet of 16 rows hit
// distinct banks.  The products, their order and the epilogue are gemm_bf16x3_kernel's.
template <typename TD>
__global__ __launch_bounds__(NT) void gemm_bf16x3d_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                          int64_t ntiles, const __bf16 *__restrict__ P,
                                                          TD *__restrict__ C, double *__restrict__ ws,
                                                          double alpha, double beta) {
    constexpr int BM = 64, BN = 64, BK = 32, FM = 2, FN = 2;
    constexpr int SL = BM * (BK / 8);  // 16-byte slots per operand plane (256)
    __shared__ __attribute__((aligned(16))) bf8_t Ls[2][2][3][SL];

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t kp = G.kp;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    d4_t acc64[FM][FN];
    f4_t lo[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            acc64[i][j] = d4_t{0, 0, 0, 0};
            lo[i][j] = f4_t{0, 0, 0, 0};
        }

    // slab loads: the 3 planes × 256 slots of A, then of B, are 24 runs of 64 slots; wave
    // w issues runs 6w .. 6w + 5.  Slot s of a plane holds row s / 4, octet
    // (s % 4) ^ ((s / 16) % 4).
    const __bf16 *Ap = P + G.pa + row0 * kp + kbeg;
    const __bf16 *Bp = P + G.pb + col0 * kp + kbeg;
    uint32_t off[6];  // byte offset of this lane's chunk from the slab base of its run
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int run = 6 * wave + e, opnd = run / 12, pl = (run / 4) % 3;
        const int sl = 64 * (run % 4) + lane, r = sl >> 2, g = (sl & 3) ^ ((sl >> 4) & 3);
        off[e] = (uint32_t)(2 * (pl * (opnd ? G.sb : G.sa) + r * kp + 8 * g));
    }
    auto issue = [&](int64_t kt, int st) {
        const char *ab = reinterpret_cast<const char *>(Ap + kt * BK);
        const char *bb = reinterpret_cast<const char *>(Bp + kt * BK);
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const int run = 6 * wave + e, opnd = run / 12, pl = (run / 4) % 3;
            __builtin_amdgcn_global_load_lds((const void *)((opnd ? bb : ab) + off[e]),
                                             (__attribute__((address_space(3))) void *)&Ls[st][opnd][pl][64 * (run % 4)],
                                             16, 0, 0);
        }
    };

    bool frag_ok[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) frag_ok[i][j] = row0 + wm * 32 + i * 16 < m && col0 + wn * 32 + j * 16 < n;

    const int64_t nk = (K - kbeg + BK - 1) / BK;
    const int g4 = lane >> 4;
    if (nk > 0) issue(0, 0);
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int st = (int)(kt & 1);
        // this wave's loads of slab kt have landed; the barrier covers every wave's and
        // every wave's reads of the stage that slab kt + 1 refills (a plain barrier:
        // __syncthreads' release fence would also wait for the loads issued next)
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (kt + 1 < nk) issue(kt + 1, st ^ 1);
        bf8_t a[3][FM], b[3][FN];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int r = wm * 32 + i * 16 + (lane & 15);
                a[p][i] = Ls[st][0][p][4 * r + (g4 ^ ((r >> 2) & 3))];
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int r = wn * 32 + j * 16 + (lane & 15);
                b[p][j] = Ls[st][1][p][4 * r + (g4 ^ ((r >> 2) & 3))];
            }
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (!frag_ok[i][j]) continue;
                f4_t h = f4_t{0, 0, 0, 0};
                h = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], h, 0, 0, 0);
                f4_t l = lo[i][j];
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], l, 0, 0, 0);
                lo[i][j] = l;
#pragma unroll
                for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)h[v];
            }
    }

    auto emit = [&](auto *Cg, int64_t ldc, double a_, bool has_beta) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int64_t col = col0 + wn * 32 + j * 16 + (lane & 15);
                if (col >= n) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t row = row0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
                    if (row < m)
                        put_c(Cg + row + col * ldc, a_ * (acc64[i][j][v] + (double)lo[i][j][v]), beta, has_beta);
                }
            }
    };
    const bool partial = G.splits > 1;
    if constexpr (std::is_same<TD, double>::value)
        emit(partial ? ws + G.ws_off + z * m * n : C + G.c_off, partial ? m : G.ldc, partial ? 1.0 : alpha,
             !partial && beta != 0.0);
    else if (partial)
        emit(ws + G.ws_off + z * m * n, m, 1.0, false);
    else
        emit(C + G.c_off, G.ldc, alpha, beta != 0.0);
}

// STAGES: register stages of prefetched slabs (1: the next slab loads while the current
// one is multiplied; 2: loads run two slabs ahead, for the short 32-deep bf16 slabs)
template <typename TD, int BM, int BN, int STAGES>
__global__ __launch_bounds__(NT) void gemm_bf16x3_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                         int64_t ntiles, const __bf16 *__restrict__ P,
                                                         TD *__restrict__ C, double *__restrict__ ws,
                                                         double alpha, double beta) {
    constexpr int BK = 32;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int AC = 3 * BM * (BK / 8) / NT;  // 16-byte chunks of the A slab per thread
    constexpr int BC = 3 * BN * (BK / 8) / NT;
    constexpr int KP = BK + 8;  // row pitch (bf16 elements): 80 bytes, keeps 16-byte alignment
    __shared__ __attribute__((aligned(16))) __bf16 As[3][BM][KP];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[3][BN][KP];

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t tm = lmn / G.tiles_n, tn = lmn % G.tiles_n;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);  // end of this split's K range
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int64_t kp = G.kp;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    d4_t acc64[FM][FN];
    f4_t lo[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            acc64[i][j] = d4_t{0, 0, 0, 0};
            lo[i][j] = f4_t{0, 0, 0, 0};
        }

    // chunk c of a slab: plane c / (4·rows), row (c % (4·rows)) / 4, k octet c % 4
    u4_t ra[STAGES][AC], rb[STAGES][BC];
    const __bf16 *Ap = P + G.pa + row0 * kp;
    const __bf16 *Bp = P + G.pb + col0 * kp;
    auto load_slab = [&](u4_t (&xa)[AC], u4_t (&xb)[BC], int64_t k0) {
#pragma unroll
        for (int e = 0; e < AC; ++e) {
            const int c = e * NT + tid, p = c / (4 * BM), rem = c % (4 * BM);
            xa[e] = *reinterpret_cast<const u4_t *>(Ap + p * G.sa + (rem >> 2) * kp + k0 + 8 * (rem & 3));
        }
#pragma unroll
        for (int e = 0; e < BC; ++e) {
            const int c = e * NT + tid, p = c / (4 * BN), rem = c % (4 * BN);
            xb[e] = *reinterpret_cast<const u4_t *>(Bp + p * G.sb + (rem >> 2) * kp + k0 + 8 * (rem & 3));
        }
    };
    auto store_slab = [&](const u4_t (&xa)[AC], const u4_t (&xb)[BC]) {
#pragma unroll
        for (int e = 0; e < AC; ++e) {
            const int c = e * NT + tid, p = c / (4 * BM), rem = c % (4 * BM);
            *reinterpret_cast<u4_t *>(&As[p][rem >> 2][8 * (rem & 3)]) = xa[e];
        }
#pragma unroll
        for (int e = 0; e < BC; ++e) {
            const int c = e * NT + tid, p = c / (4 * BN), rem = c % (4 * BN);
            *reinterpret_cast<u4_t *>(&Bs[p][rem >> 2][8 * (rem & 3)]) = xb[e];
        }
    };

    // the 16×16 fragments wholly outside C (zero padding of edge tiles) are skipped; the
    // test is uniform over the wave
    bool frag_ok[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
            frag_ok[i][j] = row0 + wm * WM + i * 16 < m && col0 + wn * WN + j * 16 < n;

    const int64_t nk = (K - kbeg + BK - 1) / BK;
#pragma unroll
    for (int st = 0; st < STAGES; ++st)
        if (st < nk) load_slab(ra[st], rb[st], kbeg + st * BK);
    // the stage index of slab kt is kt % STAGES; the loop is unrolled by STAGES so
    // that every register array is indexed by a constant
    for (int64_t kb = 0; kb < nk; kb += STAGES) {
#pragma unroll
    for (int st = 0; st < STAGES; ++st) {
        const int64_t kt = kb + st;
        if (kt >= nk) break;
        if (kt > 0) __syncthreads();  // previous slab fully consumed
        store_slab(ra[st], rb[st]);
        __syncthreads();
        if (kt + STAGES < nk) load_slab(ra[st], rb[st], kbeg + (kt + STAGES) * BK);
        const int kq = 8 * (lane >> 4);
        bf8_t a[3][FM], b[3][FN];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
                a[p][i] = *reinterpret_cast<const bf8_t *>(&As[p][wm * WM + i * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[p][j] = *reinterpret_cast<const bf8_t *>(&Bs[p][wn * WN + j * 16 + (lane & 15)][kq]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (!frag_ok[i][j]) continue;
                f4_t h = f4_t{0, 0, 0, 0};
                h = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], h, 0, 0, 0);
                f4_t l = lo[i][j];
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], l, 0, 0, 0);
                l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], l, 0, 0, 0);
                lo[i][j] = l;
#pragma unroll
                for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)h[v];
            }
    }
    }

    auto emit = [&](auto *Cg, int64_t ldc, double a_, bool has_beta) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int64_t col = col0 + wn * WN + j * 16 + (lane & 15);
                if (col >= n) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t row = row0 + wm * WM + i * 16 + 4 * (lane >> 4) + v;
                    if (row < m)
                        put_c(Cg + row + col * ldc, a_ * (acc64[i][j][v] + (double)lo[i][j][v]), beta, has_beta);
                }
            }
    };
    const bool partial = G.splits > 1;
    if constexpr (std::is_same<TD, double>::value)
        emit(partial ? ws + G.ws_off + z * m * n : C + G.c_off, partial ? m : G.ldc, partial ? 1.0 : alpha,
             !partial && beta != 0.0);
    else if (partial)
        emit(ws + G.ws_off + z * m * n, m, 1.0, false);
    else
        emit(C + G.c_off, G.ldc, alpha, beta != 0.0);
}


// --------------------------------------------------------------------------------
// Skinny groups (min(m, n) ≤ 32): one thread owns one index of the long output
// dimension and accumulates its whole short row/column in FP64 registers; the small
// operand is staged through LDS in 32-wide K chunks.  These products move
// O(long·(k + short)) bytes for O(long·k·short) flops, i.e. they are HBM-bound, so
// FMA in FP64 is used for every precision form.
// --------------------------------------------------------------------------------
template <int SK, typename TD>
__global__ __launch_bounds__(NT) void gemm_skinny_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                         const TD *__restrict__ A,
                                                         const TD *__restrict__ Bm, TD *__restrict__ C,
                                                         double alpha, double beta) {
    constexpr int KC = 32;
    __shared__ double Sm[KC][SK + 1];
    const int64_t tile = (int64_t)blockIdx.x;
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t m = G.m, n = G.n, K = G.k;
    const TD *Ag = A + G.a_off;
    const TD *Bg = Bm + G.b_off;
    const int64_t lda = G.lda, ldb = G.ldb, ldc = G.ldc;
    const int ta = G.ta, tb = G.tb;
    const bool longm = (G.skinny_mode == 0);
    const int64_t L = longm ? m : n;          // long dimension
    const int S = (int)(longm ? n : m);       // short dimension (≤ SK)
    const int64_t x = lt * NT + threadIdx.x;  // this thread's long index
    double acc[SK];
#pragma unroll
    for (int q = 0; q < SK; ++q) acc[q] = 0.0;

    for (int64_t k0 = 0; k0 < K; k0 += KC) {
        const int kc = (int)min((int64_t)KC, K - k0);
        // stage the short operand: longm -> op(B)[k0+l][q] ; else op(A)[q][k0+l]
        for (int e = threadIdx.x; e < KC * SK; e += NT) {
            const int l = e / SK, q = e % SK;
            double v = 0.0;
            if (l < kc && q < S) {
                const int64_t gl = k0 + l;
                if (longm)
                    v = tb ? Bg[q + gl * ldb] : Bg[gl + (int64_t)q * ldb];
                else
                    v = ta ? Ag[gl + (int64_t)q * lda] : Ag[q + gl * lda];
            }
            Sm[l][q] = v;
        }
        __syncthreads();
        if (x < L) {
            for (int l = 0; l < kc; ++l) {
                const int64_t gl = k0 + l;
                double v;
                if (longm)
                    v = ta ? Ag[gl + x * lda] : Ag[x + gl * lda];
                else
                    v = tb ? Bg[x + gl * ldb] : Bg[gl + x * ldb];
#pragma unroll
                for (int q = 0; q < SK; ++q) acc[q] = fma(v, Sm[l][q], acc[q]);
            }
        }
        __syncthreads();
    }
    if (x >= L) return;
    const bool has_beta = (beta != 0.0);
    TD *Cg = C + G.c_off;
#pragma unroll
    for (int q = 0; q < SK; ++q)
        if (q < S) put_c(longm ? (Cg + x + (int64_t)q * ldc) : (Cg + q + x * ldc), alpha * acc[q], beta, has_beta);
}

// --------------------------------------------------------------------------------
// Complex groups (ComplexF64 / ComplexF32 data, FP64 or FP32 MFMA): every group of the
// plan, tiled BM×BN with BK-deep K slabs.  The real and imaginary parts of a slab are
// staged in separate LDS arrays, each 16×16 fragment takes four real products,
//   Re += a_r·b_r − a_i·b_i,   Im += a_r·b_i + a_i·b_r,
// and op(A) / op(B) of an adjoint operand is the conjugate transpose (its imaginary parts
// negated as they are staged).  One LDS buffer, the next slab waits in registers; 64-bit
// addresses and a bounds test on every load.  Split groups write their partial products
// as (re, im) pairs, reduced by zreduce.
// --------------------------------------------------------------------------------
typedef double z2_t __attribute__((ext_vector_type(2)));
typedef float c2_t __attribute__((ext_vector_type(2)));

// α·v + β·c in FP64 (complex α, β), rounded once to C's element type
template <typename TD>
__device__ __forceinline__ void put_z(TD *p, double vr, double vi, double ar, double ai, double br, double bi,
                                      bool has_beta) {
    double rr = ar * vr - ai * vi, ri = ar * vi + ai * vr;
    if (has_beta) {
        const double cr = (double)(*p).x, ci = (double)(*p).y;
        rr += br * cr - bi * ci;
        ri += br * ci + bi * cr;
    }
    TD out;
    out.x = rr;
    out.y = ri;
    *p = out;
}

template <typename TC, typename TD, int BM, int BN, int BK>
__global__ __launch_bounds__(NT) void gemm_z_kernel(const GGroup *__restrict__ groups, int64_t ngroups,
                                                    int64_t ntiles, const TD *__restrict__ A,
                                                    const TD *__restrict__ Bm, TD *__restrict__ C,
                                                    double *__restrict__ ws, double ar, double ai, double br,
                                                    double bi, int conja, int conjb) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int AE = BM * BK / NT, BE = BN * BK / NT;
    constexpr int KP = BK + 16 / (int)sizeof(TC);
    __shared__ __attribute__((aligned(16))) TC Ars[BM][KP];
    __shared__ __attribute__((aligned(16))) TC Ais[BM][KP];
    __shared__ __attribute__((aligned(16))) TC Brs[BN][KP];
    __shared__ __attribute__((aligned(16))) TC Bis[BN][KP];
    using acc_t = typename MfmaK4<TC>::acc_t;

    const int64_t tile = xcd_remap((int64_t)blockIdx.x, ntiles);
    const GGroup &G = groups[find_group(groups, ngroups, tile)];
    const int64_t lt = tile - G.tile_start;
    const int64_t tiles_mn = G.tiles_m * G.tiles_n;
    const int64_t z = lt / tiles_mn, lmn = lt - z * tiles_mn;
    const int64_t row0 = (lmn / G.tiles_n) * BM, col0 = (lmn % G.tiles_n) * BN;
    const int64_t m = G.m, n = G.n;
    const int64_t kbeg = z * G.kchunk;
    const int64_t K = min(G.k, kbeg + G.kchunk);
    const int64_t lda = G.lda, ldb = G.ldb;
    const int ta = G.ta, tb = G.tb;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    acc_t cr[FM][FN], ci[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            cr[i][j] = acc_t{0, 0, 0, 0};
            ci[i][j] = acc_t{0, 0, 0, 0};
        }
    auto a_el = [&](int e, int &i, int &k) {
        const int idx = e * NT + tid;
        if (!ta) { i = idx % BM; k = idx / BM; }
        else     { k = idx % BK; i = idx / BK; }
    };
    auto b_el = [&](int e, int &j, int &k) {
        const int idx = e * NT + tid;
        if (!tb) { k = idx % BK; j = idx / BK; }
        else     { j = idx % BN; k = idx / BN; }
    };
    TD ra[AE], rb[BE];
    const TD zero = TD{0, 0};
    auto load = [&](int64_t k0) {
        const int64_t ilim = min((int64_t)BM, m - row0), jlim = min((int64_t)BN, n - col0);
        const int64_t klim = min((int64_t)BK, K - k0);
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            int i, k;
            a_el(e, i, k);
            const int64_t gi = row0 + i, gk = k0 + k;
            ra[e] = (i < ilim && k < klim) ? A[G.a_off + (ta ? gk + gi * lda : gi + gk * lda)] : zero;
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            int j, k;
            b_el(e, j, k);
            const int64_t gj = col0 + j, gk = k0 + k;
            rb[e] = (j < jlim && k < klim) ? Bm[G.b_off + (tb ? gj + gk * ldb : gk + gj * ldb)] : zero;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            int i, k;
            a_el(e, i, k);
            Ars[i][k] = (TC)ra[e].x;
            Ais[i][k] = conja ? -(TC)ra[e].y : (TC)ra[e].y;
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            int j, k;
            b_el(e, j, k);
            Brs[j][k] = (TC)rb[e].x;
            Bis[j][k] = conjb ? -(TC)rb[e].y : (TC)rb[e].y;
        }
    };
    constexpr int KV = 16 / (int)sizeof(TC), KC = 4 * KV;
    static_assert(BK % KC == 0, "K slab must hold whole fragment rounds");
    const int kq = KV * (lane >> 4);
    auto mma_slab = [&]() {
#pragma unroll
        for (int kk = 0; kk < BK; kk += KC) {
            typedef TC vk_t __attribute__((ext_vector_type(KV)));
            vk_t far_[FM], fai[FM], fbr[FN], fbi[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                far_[i] = *reinterpret_cast<const vk_t *>(&Ars[wm * WM + i * 16 + (lane & 15)][kk + kq]);
                fai[i] = *reinterpret_cast<const vk_t *>(&Ais[wm * WM + i * 16 + (lane & 15)][kk + kq]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                fbr[j] = *reinterpret_cast<const vk_t *>(&Brs[wn * WN + j * 16 + (lane & 15)][kk + kq]);
                fbi[j] = *reinterpret_cast<const vk_t *>(&Bis[wn * WN + j * 16 + (lane & 15)][kk + kq]);
            }
#pragma unroll
            for (int q = 0; q < KV; ++q)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        cr[i][j] = MfmaK4<TC>::mma(far_[i][q], fbr[j][q], cr[i][j]);
                        cr[i][j] = MfmaK4<TC>::mma(-fai[i][q], fbi[j][q], cr[i][j]);
                        ci[i][j] = MfmaK4<TC>::mma(far_[i][q], fbi[j][q], ci[i][j]);
                        ci[i][j] = MfmaK4<TC>::mma(fai[i][q], fbr[j][q], ci[i][j]);
                    }
        }
    };
    const int64_t nk = K > kbeg ? (K - kbeg + BK - 1) / BK : 0;
    if (nk > 0) load(kbeg);
    for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt > 0) __syncthreads();
        store();
        __syncthreads();
        if (kt + 1 < nk) load(kbeg + (kt + 1) * BK);
        mma_slab();
    }

    const bool partial = G.splits > 1;
    const bool has_beta = !partial && (br != 0.0 || bi != 0.0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t col = col0 + wn * WN + j * 16 + (lane & 15);
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = row0 + wm * WM + i * 16 + MfmaK4<TC>::row(lane, v);
                if (row >= m) continue;
                const double vr = (double)cr[i][j][v], vi = (double)ci[i][j][v];
                if (partial) {
                    double *w = ws + 2 * (G.ws_off + z * m * n + row + col * m);
                    w[0] = vr;
                    w[1] = vi;
                } else {
                    put_z(C + G.c_off + row + col * G.ldc, vr, vi, ar, ai, br, bi, has_beta);
                }
            }
        }
}

// Complex split-K reduction (and, with ws = nullptr, C = β·C for α = 0): one thread per
// output element, partial (re, im) pairs summed in fixed z order.
template <typename TD>
__global__ __launch_bounds__(NT) void zreduce(const RGroup *__restrict__ rg, int64_t ng, int64_t total,
                                              const double *__restrict__ ws, TD *__restrict__ C, double ar,
                                              double ai, double br, double bi) {
    const bool has_beta = (br != 0.0 || bi != 0.0);
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < total; e += (int64_t)gridDim.x * NT) {
        int64_t lo = 0, hi = ng - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rg[mid].elem_start <= e) lo = mid; else hi = mid - 1;
        }
        const RGroup &R = rg[lo];
        const int64_t le = e - R.elem_start;
        const int64_t mn = R.m * R.n;
        double vr = 0.0, vi = 0.0;
        if (ws)
            for (int64_t zz = 0; zz < R.splits; ++zz) {
                vr += ws[2 * (R.ws_off + zz * mn + le)];
                vi += ws[2 * (R.ws_off + zz * mn + le) + 1];
            }
        const int64_t row = le % R.m, col = le / R.m;
        put_z(C + R.c_off + row + col * R.ldc, vr, vi, ar, ai, br, bi, has_beta);
    }
}

// --------------------------------------------------------------------------------
// Dot groups (m = n = 1 with K > SKINNY_K: a contraction down to a scalar, whose only
// coupled sector is the unit): HBM-bound reductions in two passes of fixed order, FP64
// for every precision form.  Pass 1: one workgroup per K chunk of a group writes its
// partial sum; pass 2: one workgroup per group adds its partials in order and applies
// α and β.
// --------------------------------------------------------------------------------
struct DGroup {
    int64_t k, a_off, a_inc, b_off, b_inc, c_off;  // op(A)(0, i) = A[a_off + i·a_inc], op(B)(i, 0) likewise
    int64_t part_start, nparts, chunk;
};

template <typename TD>
__global__ __launch_bounds__(NT) void gemm_dot_partial(const DGroup *__restrict__ dg, int64_t ng,
                                                       const TD *__restrict__ A, const TD *__restrict__ B,
                                                       double *__restrict__ parts) {
    __shared__ double sh[NT];
    const int64_t part = (int64_t)blockIdx.x;
    int64_t lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (dg[mid].part_start <= part) lo = mid; else hi = mid - 1;
    }
    const DGroup &D = dg[lo];
    const int64_t k0 = (part - D.part_start) * D.chunk, k1 = min(D.k, k0 + D.chunk);
    const TD *a = A + D.a_off, *b = B + D.b_off;
    const int64_t ai = D.a_inc, bi = D.b_inc;
    // two independent sums per thread keep two load pairs in flight; the order of every
    // addition is fixed by (k, chunk), so results are reproducible
    double s0 = 0.0, s1 = 0.0;
    int64_t i = k0 + threadIdx.x;
    for (; i + NT < k1; i += 2 * NT) {
        s0 += (double)a[i * ai] * (double)b[i * bi];
        s1 += (double)a[(i + NT) * ai] * (double)b[(i + NT) * bi];
    }
    for (; i < k1; i += NT) s0 += (double)a[i * ai] * (double)b[i * bi];
    sh[threadIdx.x] = s0 + s1;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[part] = sh[0];
}

template <typename TD>
__global__ __launch_bounds__(NT) void gemm_dot_final(const DGroup *__restrict__ dg, const double *__restrict__ parts,
                                                     TD *__restrict__ C, double alpha, double beta) {
    __shared__ double sh[NT];
    const DGroup &D = dg[blockIdx.x];
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < D.nparts; j += NT) s += parts[D.part_start + j];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) put_c(C + D.c_off, alpha * sh[0], beta, beta != 0.0);
}

struct DotSet {
    std::vector<DGroup> groups;
    int64_t nparts = 0;
    DGroup *d_groups = nullptr;
    double *parts = nullptr;  // plan-owned partial sums
};

struct TileSet {
    std::vector<GGroup> groups;
    int64_t ntiles = 0;
    GGroup *d_groups = nullptr;
    std::vector<RGroup> rgroups;  // split-K groups to reduce
    int64_t relems = 0;
    RGroup *d_rgroups = nullptr;
    double *ws = nullptr;  // split-K partial products (plan-owned)
    int64_t split = 1;     // plan-wide split-K factor the schedule model chose
    // persistent kernel: tiles of slot w are sched[slot_start[w] .. slot_start[w+1])
    std::vector<int32_t> sched, slot_start;
    int32_t *d_sched = nullptr, *d_slot_start = nullptr;
    int64_t nslots = 0;
    // the direct-load FP64 kernel (gemm_k4d_kernel) can run these tiles: every group
    // untransposed, with slab offsets that fit in 32 bits
    bool direct_ok = false;
    // the direct-load 3×BF16 kernel (gemm_bf16x3d_kernel): panel offsets inside a slab fit
    // in 32 bits
    bool bf16_direct_ok = false;
};

struct SkinnySet {
    std::vector<GGroup> groups;
    int64_t ntiles = 0;
    int sk = 0;  // register width (8, 16 or 32)
    GGroup *d_groups = nullptr;
};

// Tile shapes the autotuner chooses from (BM × BN output tile, BK slab depth; the
// 3×BF16 kernel always uses BK = 32).
struct Cfg {
    int bm, bn, bk, nbuf;  // FP64/FP32 kernel: 2 gemm_k4 (two LDS buffers), 1 gemm_k4s, 3 gemm_k4p, 5 gemm_pks,
                           // 6 gemm_k4d (direct global → LDS loads; FP64, untransposed operands)
    int split;             // plan-wide split-K factor: 0 = the schedule model's choice
};
// nbuf 5: the persistent kernel with a host-made schedule (gemm_pks_kernel).  Configurations
// 10–12 are configuration 5 with the split-K factor fixed at 1, 2 and 3: the autotuner
// measures them against the model's choice (the model prices a split group's partial
// products and reduction only roughly).  Configurations 13–16 are the direct-load
// kernels (FP64; FP32 on FP32 data; 3×BF16) with the model's split factor and with the
// factor fixed at 1, 2 and 3; where they do not apply (transposed operands, FP32
// arithmetic on FP64 data) they run configuration 5's kernel with their split factor.
// Configurations 17–19 are the FP64 direct-load kernel on 128×128, 128×64 and 64×128
// tiles (model split factor); in FP32 and 3×BF16 they run the one-buffer kernels of
// their tiles.
constexpr Cfg CFGS[] = {{64, 64, 16, 2, 0},  {128, 128, 16, 2, 0}, {128, 64, 16, 2, 0}, {64, 128, 16, 2, 0},
                        {64, 64, 32, 2, 0},  {64, 64, 16, 1, 0},   {64, 64, 32, 1, 0},  {128, 128, 16, 1, 0},
                        {64, 64, 16, 3, 0},  {64, 64, 16, 5, 0},   {64, 64, 16, 1, 1},  {64, 64, 16, 1, 2},
                        {64, 64, 16, 1, 3},  {64, 64, 16, 6, 0},   {64, 64, 16, 6, 1},  {64, 64, 16, 6, 2},
                        {64, 64, 16, 6, 3},  {128, 128, 16, 6, 0}, {128, 64, 16, 6, 0}, {64, 128, 16, 6, 0}};
// the 3×BF16 kernel has a fixed 32-deep slab and its own LDS use: only the tile edges
// of a configuration matter to it, the prefetch depth of configuration 8, the fixed
// split factors of configurations 10–12, and configurations 13–16, which run the
// direct-load 3×BF16 kernel
constexpr bool bf16_duplicate(int c) { return (c >= 4 && c <= 7) || c == 9 || c >= 17; }
constexpr bool direct_cfg(int c) { return CFGS[c].nbuf == 6; }
// direct-load configurations with tiles other than 64×64: FP64 only
constexpr bool wide_direct(int c) { return direct_cfg(c) && (CFGS[c].bm != 64 || CFGS[c].bn != 64); }
constexpr int NCFG = sizeof(CFGS) / sizeof(CFGS[0]);

}  // namespace

struct tkmi_gemm_plan_s {
    std::vector<GGroup> base;  // MFMA groups without tiling
    // 3×BF16 split panels (allocated by the first 3×BF16 call) and the pack work list
    std::vector<PackGroup> pack;
    int64_t pack_total = 0, pack_total_a = 0, pack_total_b = 0, panel_len = 0;
    PackGroup *d_pack = nullptr;
    // operands marked constant (bit 0: A, bit 1: B) and the source pointer whose split is in
    // the panels for each (nullptr: none): a constant operand is split again only when it
    // comes from another pointer or after tkmi_gemm_plan_invalidate
    int32_t const_ops = 0;
    const void *split_src[2] = {nullptr, nullptr};
    int32_t split_dt[2] = {TKMI_F64, TKMI_F64};
    __bf16 *panels = nullptr;
    SkinnySet skinny;
    DotSet dots;
    std::vector<GGroup> skinny_c;  // the dot groups as 1×1 C blocks (α = 0 calls scale them)
    double flops_mfma = 0, flops_skinny = 0;
    double flops = 0;
    TileSet ts[NCFG];
    int choice[3] = {-1, -1, -1};  // per precision: config index, -1 = automatic
    // α = 0: C_g = β·C_g without reading A or B (mul! reduces to rmul!), one list of C
    // blocks per part (MFMA groups, skinny groups), run through splitk_reduce with no splits
    std::vector<RGroup> scale[2];
    int64_t scale_total[2] = {0, 0};
    RGroup *d_scale[2] = {nullptr, nullptr};
    // complex data: every group with m, n > 0 on the complex kernel (tiles made at the
    // first complex call), and the C blocks of all groups for α = 0
    std::vector<GGroup> zall;
    TileSet zts;
    bool zready = false;
    std::vector<RGroup> zscale;
    int64_t zscale_total = 0;
    RGroup *d_zscale = nullptr;
    // tkmi_gemm_plan_time_next: events that the next MFMA-tile launch records its own
    // start and end in (hipExtLaunchKernel), then cleared
    hipEvent_t time_ev[2] = {nullptr, nullptr};
};

namespace {

// Workgroups of configuration c's FP64 kernel that run at once on the current device
// (occupancy per CU × CUs): the slots the split-K schedule model deals tiles to.
template <int BM, int BN, int BK, int NBUF>
int kernel_occupancy() {
    int nb = 0;
    hipError_t e;
    if constexpr (NBUF == 5)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_pks_kernel<double, double, BM, BN, BK>), NT, 0);
    else if constexpr (NBUF == 3)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_k4p_kernel<double, double, BM, BN, BK>), NT, 0);
    else if constexpr (NBUF == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_k4_kernel<double, double, BM, BN, BK>), NT, 0);
    else if constexpr (NBUF == 6 && BM == 64 && BN == 64)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_k4d_kernel), NT, 0);
    else if constexpr (NBUF == 6)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_k4dw_kernel<BM, BN>), NT, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&gemm_k4s_kernel<double, double, BM, BN, BK>), NT, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

int64_t cfg_slots(int c) {
    constexpr int MAXDEV = 64;
    static int cache[MAXDEV][NCFG] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) dev = 0;
    int &v = cache[dev][c];
    if (v == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        int occ;
        switch (c) {
            case 0: occ = kernel_occupancy<64, 64, 16, 2>(); break;
            case 1: occ = kernel_occupancy<128, 128, 16, 2>(); break;
            case 2: occ = kernel_occupancy<128, 64, 16, 2>(); break;
            case 3: occ = kernel_occupancy<64, 128, 16, 2>(); break;
            case 4: occ = kernel_occupancy<64, 64, 32, 2>(); break;
            case 5: occ = kernel_occupancy<64, 64, 16, 1>(); break;
            case 6: occ = kernel_occupancy<64, 64, 32, 1>(); break;
            case 7: occ = kernel_occupancy<128, 128, 16, 1>(); break;
            case 8: occ = kernel_occupancy<64, 64, 16, 3>(); break;
            case 9: occ = kernel_occupancy<64, 64, 16, 5>(); break;
            case 13: case 14: case 15: case 16: occ = kernel_occupancy<64, 64, 16, 6>(); break;
            case 17: occ = kernel_occupancy<128, 128, 16, 6>(); break;
            case 18: occ = kernel_occupancy<128, 64, 16, 6>(); break;
            case 19: occ = kernel_occupancy<64, 128, 16, 6>(); break;
            default: occ = kernel_occupancy<64, 64, 16, 1>(); break;
        }
        v = occ * cus;
    }
    return v;
}

// Split-K factor of a group for a plan-wide factor S: long-K groups (≥ LONG_K slabs) split
// into at most S chunks of at least MIN_SLABS slabs; a chunk is whole slabs of this
// configuration and a multiple of 32 (the 3×BF16 slab).
constexpr int64_t MAX_SPLIT = 8, MIN_SLABS = 16, LONG_K = 32;

void split_group(GGroup &g, int64_t S, int bk) {
    const int64_t slabs = (g.k + bk - 1) / bk;
    int64_t sg = slabs >= LONG_K ? std::max<int64_t>(1, std::min<int64_t>(S, slabs / MIN_SLABS)) : 1;
    g.kchunk = std::max<int64_t>(32, (((slabs + sg - 1) / sg) * bk + 31) / 32 * 32);
    g.splits = g.k > 0 ? (g.k + g.kchunk - 1) / g.kchunk : 1;
}

// Modelled launch time of the plan's tiles for a plan-wide split factor S, in units of
// one BK slab on one workgroup slot: the greedy schedule (heaviest tile first, next tile
// to the first free slot — the order the tiles are numbered in) of every tile's slabs
// plus a fixed per-tile cost, plus the split groups' extra HBM traffic (each partial
// product written and read back once) at an effective FP64 MFMA : HBM ratio.
double split_model(const std::vector<GGroup> &base, int64_t S, int bm, int bn, int bk, int64_t slots) {
    constexpr double TILE_COST = 3.0;      // prologue + epilogue, in slabs (measured fit)
    constexpr double FLOP_PER_BYTE = 8.0;  // ≈ 35 TFLOP/s effective FP64 MFMA : 4.4 TB/s
    std::vector<double> cost;
    double extra = 0;
    for (GGroup g : base) {
        split_group(g, S, bk);
        const int64_t tiles = ((g.m + bm - 1) / bm) * ((g.n + bn - 1) / bn);
        for (int64_t z = 0; z < g.splits; ++z) {
            const int64_t kz = std::min(g.k, (z + 1) * g.kchunk) - z * g.kchunk;
            const double c = (double)((kz + bk - 1) / bk) + TILE_COST;
            for (int64_t t = 0; t < tiles; ++t) cost.push_back(c);
        }
        if (g.splits > 1) extra += 2.0 * (double)g.splits * (double)g.m * (double)g.n * sizeof(double);
    }
    if (cost.empty()) return 0;
    std::sort(cost.begin(), cost.end(), std::greater<double>());
    const size_t w = (size_t)std::max<int64_t>(1, std::min<int64_t>(slots, (int64_t)cost.size()));
    std::vector<double> load(w, 0.0);  // min-heap of slot finish times
    for (double c : cost) {
        std::pop_heap(load.begin(), load.end(), std::greater<double>());
        load.back() += c;
        std::push_heap(load.begin(), load.end(), std::greater<double>());
    }
    const double makespan = *std::max_element(load.begin(), load.end());
    return makespan + extra * FLOP_PER_BYTE / (2.0 * bm * bn * bk * (double)slots);
}

int build_tiles(tkmi_gemm_plan_s *p, int c) {
    TileSet &ts = p->ts[c];
    const int bm = CFGS[c].bm, bn = CFGS[c].bn;
    const int bk = CFGS[c].bk;
    ts.groups.clear();
    // plan-wide split-K factor: the one with the shortest modelled launch time (ties:
    // the smaller factor)
    const int64_t slots = cfg_slots(c);
    int64_t S = 1;
    if (CFGS[c].split > 0) {
        S = CFGS[c].split;
    } else {
        double best = split_model(p->base, 1, bm, bn, bk, slots);
        for (int64_t cand = 2; cand <= MAX_SPLIT; ++cand) {
            const double t = split_model(p->base, cand, bm, bn, bk, slots);
            if (t < best * 0.98) {
                best = t;
                S = cand;
            }
        }
    }
    ts.split = S;
    int64_t start = 0, ws_len = 0;
    ts.rgroups.clear();
    ts.relems = 0;
    std::vector<GGroup> order;
    for (auto g : p->base) {
        g.tiles_m = (g.m + bm - 1) / bm;
        g.tiles_n = (g.n + bn - 1) / bn;
        split_group(g, S, bk);
        const int64_t sg = g.splits;
        g.ws_off = 0;
        if (sg > 1) {
            g.ws_off = ws_len;
            ws_len += sg * g.m * g.n;
            RGroup r{g.m, g.n, g.c_off, g.ldc, g.ws_off, sg, ts.relems};
            ts.rgroups.push_back(r);
            ts.relems += rch_round(g.m * g.n);
        }
        order.push_back(g);
    }
    // heaviest tiles first: a tile's work is proportional to its K range
    std::stable_sort(order.begin(), order.end(), [](const GGroup &a, const GGroup &b) {
        return std::min(a.k, a.kchunk) > std::min(b.k, b.kchunk);
    });
    for (auto &g : order) {
        g.tile_start = start;
        start += g.tiles_m * g.tiles_n * g.splits;
        ts.groups.push_back(g);
    }
    ts.ntiles = start;
    if (ts.groups.empty()) return TKMI_OK;
    // the direct-load kernel's 32-bit slab offsets: A chunk rows up to bm − 2 + 15·lda, B up
    // to 14 + (bn − 1)·ldb elements
    ts.direct_ok = true;
    for (const GGroup &g : ts.groups)
        if (g.ta || g.tb || 8 * (bm - 2 + 15 * g.lda) >= ((int64_t)1 << 32) ||
            8 * (14 + (bn - 1) * g.ldb) >= ((int64_t)1 << 32))
            ts.direct_ok = false;
    // 3×BF16 panel offsets: up to two plane strides, 63 rows and 3 octets, in bytes
    ts.bf16_direct_ok = true;
    for (const GGroup &g : ts.groups)
        if (2 * (2 * std::max(g.sa, g.sb) + 63 * g.kp + 24) >= ((int64_t)1 << 32)) ts.bf16_direct_ok = false;
    int rc = TKMI_OK;
    if (CFGS[c].nbuf == 5) {
        // Workgroup w runs on XCD w % 8 (workgroups are dealt round-robin to the XCDs).
        // Chunks of XCH consecutive tiles (one group, neighbouring rows and columns) go to
        // one XCD, chunk c to XCD c % 8, so they share their A and B panels in that XCD's
        // L2; within an XCD, longest processing time first: tiles in their heaviest-first
        // numbering, each to the XCD's slot that frees first (ties: the lowest slot), at
        // the modelled cost of split_model (K slabs plus a fixed per-tile cost).
        if (ts.ntiles >= ((int64_t)1 << 31)) TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: too many tiles to schedule");
        constexpr double TILE_COST = 3.0;
        constexpr int64_t NX = 8, XCH = 16;
        const int64_t w = std::max<int64_t>(1, std::min<int64_t>(slots, ts.ntiles));
        std::vector<std::vector<int32_t>> lists((size_t)w);
        auto later = [](const std::pair<double, int64_t> &a, const std::pair<double, int64_t> &b) {
            return a.first > b.first || (a.first == b.first && a.second > b.second);
        };
        const int64_t nx = std::min<int64_t>(NX, w);
        std::vector<std::vector<std::pair<double, int64_t>>> heaps((size_t)nx);
        for (int64_t i = 0; i < w; ++i) heaps[(size_t)(i % nx)].push_back({0.0, i});
        for (auto &h : heaps) std::make_heap(h.begin(), h.end(), later);
        for (const GGroup &g : ts.groups) {
            const int64_t tmn = g.tiles_m * g.tiles_n;
            for (int64_t z = 0; z < g.splits; ++z) {
                const int64_t kz = std::max<int64_t>(0, std::min(g.k, (z + 1) * g.kchunk) - z * g.kchunk);
                const double cost = (double)std::max<int64_t>(1, (kz + bk - 1) / bk) + TILE_COST;
                for (int64_t t = 0; t < tmn; ++t) {
                    const int64_t id = g.tile_start + z * tmn + t;
                    auto &heap = heaps[(size_t)((id / XCH) % nx)];
                    std::pop_heap(heap.begin(), heap.end(), later);
                    heap.back().first += cost;
                    lists[(size_t)heap.back().second].push_back((int32_t)id);
                    std::push_heap(heap.begin(), heap.end(), later);
                }
            }
        }
        ts.sched.clear();
        ts.slot_start.assign(1, 0);
        for (auto &l : lists) {
            ts.sched.insert(ts.sched.end(), l.begin(), l.end());
            ts.slot_start.push_back((int32_t)ts.sched.size());
        }
        ts.nslots = w;
        rc = tkmi::upload((void **)&ts.d_sched, ts.sched.data(), ts.sched.size() * sizeof(int32_t));
        if (!rc)
            rc = tkmi::upload((void **)&ts.d_slot_start, ts.slot_start.data(), ts.slot_start.size() * sizeof(int32_t));
        if (rc) return rc;
    }
    rc = tkmi::upload((void **)&ts.d_groups, ts.groups.data(), ts.groups.size() * sizeof(GGroup));
    if (!rc && !ts.rgroups.empty())
        rc = tkmi::upload((void **)&ts.d_rgroups, ts.rgroups.data(), ts.rgroups.size() * sizeof(RGroup));
    if (!rc && ws_len > 0) rc = tkmi::upload((void **)&ts.ws, nullptr, ws_len * sizeof(double));
    return rc;
}

// Tiles of the complex kernel (64×64×16) over every group of the plan, with the split-K
// factor of the same schedule model; partial products take two doubles per element.
constexpr int ZBM = 64, ZBN = 64, ZBK = 16;

int build_ztiles(tkmi_gemm_plan_s *p) {
    TileSet &ts = p->zts;
    const int64_t slots = cfg_slots(5);
    int64_t S = 1;
    double best = split_model(p->zall, 1, ZBM, ZBN, ZBK, slots);
    for (int64_t cand = 2; cand <= MAX_SPLIT; ++cand) {
        const double t = split_model(p->zall, cand, ZBM, ZBN, ZBK, slots);
        if (t < best * 0.98) {
            best = t;
            S = cand;
        }
    }
    ts.split = S;
    int64_t ws_len = 0;
    std::vector<GGroup> order;
    for (auto g : p->zall) {
        g.tiles_m = (g.m + ZBM - 1) / ZBM;
        g.tiles_n = (g.n + ZBN - 1) / ZBN;
        split_group(g, S, ZBK);
        g.ws_off = 0;
        if (g.splits > 1) {
            g.ws_off = ws_len;
            ws_len += g.splits * g.m * g.n;
            ts.rgroups.push_back(RGroup{g.m, g.n, g.c_off, g.ldc, g.ws_off, g.splits, ts.relems});
            ts.relems += g.m * g.n;
        }
        order.push_back(g);
        p->zscale.push_back(RGroup{g.m, g.n, g.c_off, g.ldc, 0, 0, p->zscale_total});
        p->zscale_total += g.m * g.n;
    }
    std::stable_sort(order.begin(), order.end(), [](const GGroup &a, const GGroup &b) {
        return std::min(a.k, a.kchunk) > std::min(b.k, b.kchunk);
    });
    int64_t start = 0;
    for (auto &g : order) {
        g.tile_start = start;
        start += g.tiles_m * g.tiles_n * g.splits;
        ts.groups.push_back(g);
    }
    ts.ntiles = start;
    int rc = TKMI_OK;
    if (!ts.groups.empty()) rc = tkmi::upload((void **)&ts.d_groups, ts.groups.data(), ts.groups.size() * sizeof(GGroup));
    if (!rc && !ts.rgroups.empty())
        rc = tkmi::upload((void **)&ts.d_rgroups, ts.rgroups.data(), ts.rgroups.size() * sizeof(RGroup));
    if (!rc && ws_len > 0) rc = tkmi::upload((void **)&ts.ws, nullptr, 2 * ws_len * sizeof(double));
    if (!rc && !p->zscale.empty())
        rc = tkmi::upload((void **)&p->d_zscale, p->zscale.data(), p->zscale.size() * sizeof(RGroup));
    return rc;
}

int auto_cfg(const tkmi_gemm_plan_s *p, int prec) {
    // FP64 with untransposed operands: the direct-load kernel on 64×64 tiles with the
    // model's split factor (8–10% faster than the one-buffer kernel on the sector GEMMs
    // measured, and 2–7% faster than its 128-wide tiles on uniform 8192×8192×K products
    // for K from 128 to 2048).  Otherwise 64×64×16 with one LDS buffer: the fastest on the
    // sector GEMMs measured (K or N of a few hundred).  The autotuner moves a plan
    // elsewhere when another is > 3% faster.
    return (prec == TKMI_PREC_F64 && p->ts[13].direct_ok) ? 13 : 5;
}

int cfg_of(const tkmi_gemm_plan_s *p, int prec) { return p->choice[prec] >= 0 ? p->choice[prec] : auto_cfg(p, prec); }

// A kernel launch on s; with timing events (tkmi_gemm_plan_time_next) the launch records
// its own start and end in them.
template <typename K, typename... Args>
void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t *ev, Args... args) {
    if (ev && ev[0])
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// One launch of an FP64 / FP32 MFMA kernel over the tiles of a configuration.
template <typename TC, typename TD, int BM, int BN, int BK, int NBUF>
void launch_k4(const TileSet &ts, const TD *A, const TD *B, TD *C, double alpha, double beta, hipStream_t s,
               const hipEvent_t *ev) {
    const dim3 grid((unsigned)ts.ntiles), block(NT);
    const int64_t ng = (int64_t)ts.groups.size();
    if constexpr (NBUF == 6) {
        // the direct-load kernel: FP64 arithmetic on FP64 data with untransposed operands;
        // otherwise the one-buffer kernel of the same tiles
        if constexpr (std::is_same<TC, double>::value && std::is_same<TD, double>::value) {
            if (ts.direct_ok) {
                if constexpr (BM == 64 && BN == 64)
                    launch_timed(gemm_k4d_kernel, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C, ts.ws,
                                 alpha, beta);
                else
                    launch_timed(gemm_k4dw_kernel<BM, BN>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B,
                                 C, ts.ws, alpha, beta);
                return;
            }
        }
        if constexpr (std::is_same<TC, float>::value && std::is_same<TD, float>::value && BM == 64 && BN == 64) {
            if (ts.direct_ok) {  // (its 32-bit offset bounds imply the FP32 kernel's)
                launch_timed(gemm_k4d32_kernel, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C, ts.ws,
                             alpha, beta);
                return;
            }
        }
        launch_timed(gemm_k4s_kernel<TC, TD, BM, BN, BK>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C,
                     ts.ws, alpha, beta);
    } else if constexpr (NBUF == 5)
        launch_timed(gemm_pks_kernel<TC, TD, BM, BN, BK>, dim3((unsigned)ts.nslots), block, s, ev, ts.d_groups, ng,
                     ts.d_sched, ts.d_slot_start, A, B, C, ts.ws, alpha, beta);
    else if constexpr (NBUF == 3)
        launch_timed(gemm_k4p_kernel<TC, TD, BM, BN, BK>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C,
                     ts.ws, alpha, beta);
    else if constexpr (NBUF == 2)
        launch_timed(gemm_k4_kernel<TC, TD, BM, BN, BK>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C,
                     ts.ws, alpha, beta);
    else
        launch_timed(gemm_k4s_kernel<TC, TD, BM, BN, BK>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles, A, B, C,
                     ts.ws, alpha, beta);
}

constexpr int dtype_of(double *) { return TKMI_F64; }
constexpr int dtype_of(float *) { return TKMI_F32; }

template <typename TD, int BM, int BN, int BK, int NBUF>
void launch_k(tkmi_gemm_plan_s *p, const TileSet &ts, int prec, const TD *A, const TD *B, TD *C, double alpha,
              double beta, hipStream_t s, bool tiles, bool reduce) {
    dim3 grid((unsigned)ts.ntiles), block(NT);
    const int64_t ng = (int64_t)ts.groups.size();
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (tiles) {  // the timing events go to this launch's MFMA kernel only
        ev[0] = p->time_ev[0];
        ev[1] = p->time_ev[1];
        p->time_ev[0] = p->time_ev[1] = nullptr;
    }
    if (!tiles) {
    } else if (prec == TKMI_PREC_F64) {
        if constexpr (std::is_same<TD, double>::value)
            launch_k4<double, TD, BM, BN, BK, NBUF>(ts, A, B, C, alpha, beta, s, ev);
    } else if (prec == TKMI_PREC_F32) {
        launch_k4<float, TD, BM, BN, BK, NBUF>(ts, A, B, C, alpha, beta, s, ev);
    } else {
        // split the operands whose panels do not already hold them (constant operands from
        // the same pointer and element type as the last split keep theirs)
        const int dt = dtype_of((TD *)nullptr);
        const bool keepA = (p->const_ops & 1) && p->split_src[0] == A && p->split_dt[0] == dt;
        const bool keepB = (p->const_ops & 2) && p->split_src[1] == B && p->split_dt[1] == dt;
        const int which = !keepA && !keepB ? 0 : (!keepA ? 1 : (!keepB ? 2 : -1));
        const int64_t npack = which == 0 ? p->pack_total : (which == 1 ? p->pack_total_a : p->pack_total_b);
        if (which >= 0 && npack > 0)
            hipLaunchKernelGGL(split_pack_kernel<TD>, dim3((unsigned)npack), block, 0, s, p->d_pack,
                               (int64_t)p->pack.size(), A, B, p->panels, which);
        p->split_src[0] = (p->const_ops & 1) ? (const void *)A : nullptr;
        p->split_src[1] = (p->const_ops & 2) ? (const void *)B : nullptr;
        p->split_dt[0] = p->split_dt[1] = dt;
        if constexpr (NBUF == 3)
            launch_timed(gemm_bf16x3_kernel<TD, BM, BN, 2>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles,
                         (const __bf16 *)p->panels, C, ts.ws, alpha, beta);
        else if (NBUF == 6 && BM == 64 && BN == 64 && ts.bf16_direct_ok)
            launch_timed(gemm_bf16x3d_kernel<TD>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles,
                         (const __bf16 *)p->panels, C, ts.ws, alpha, beta);
        else
            launch_timed(gemm_bf16x3_kernel<TD, BM, BN, 1>, grid, block, s, ev, ts.d_groups, ng, ts.ntiles,
                         (const __bf16 *)p->panels, C, ts.ws, alpha, beta);
    }
    if (reduce && ts.relems > 0) {
        const int64_t blocks = std::min<int64_t>(ts.relems / RCH, (int64_t)1 << 20);
        hipLaunchKernelGGL(splitk_reduce<TD>, dim3((unsigned)blocks), dim3(NT), 0, s, ts.d_rgroups,
                           (int64_t)ts.rgroups.size(), ts.relems, ts.ws, C, alpha, beta);
    }
}

template <typename TD>
int launch_cfg(tkmi_gemm_plan_s *p, int c, int prec, const TD *A, const TD *B, TD *C, double alpha, double beta,
               hipStream_t s, bool tiles = true, bool reduce = true) {
    const TileSet &ts = p->ts[c];
    if (ts.ntiles == 0) return TKMI_OK;
    if (prec == TKMI_PREC_BF16X3 && !p->panels) {
        // first 3×BF16 call on this plan: split panels and pack list (not during graph capture)
        int rc = tkmi::upload((void **)&p->panels, nullptr, (size_t)p->panel_len * sizeof(__bf16));
        if (!rc) rc = tkmi::upload((void **)&p->d_pack, p->pack.data(), p->pack.size() * sizeof(PackGroup));
        if (rc) {
            (void)hipFree(p->panels);
            p->panels = nullptr;
            return rc;
        }
    }
    if (ts.ntiles * NT >= ((int64_t)1 << 31))
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: more output tiles than one launch can hold (2^31 work-items)");
    switch (c) {
        case 0: launch_k<TD, 64, 64, 16, 2>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 1: launch_k<TD, 128, 128, 16, 2>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 2: launch_k<TD, 128, 64, 16, 2>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 3: launch_k<TD, 64, 128, 16, 2>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 4: launch_k<TD, 64, 64, 32, 2>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 5: launch_k<TD, 64, 64, 16, 1>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 6: launch_k<TD, 64, 64, 32, 1>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 7: launch_k<TD, 128, 128, 16, 1>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 8: launch_k<TD, 64, 64, 16, 3>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 9: launch_k<TD, 64, 64, 16, 5>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 13: case 14: case 15: case 16:
            launch_k<TD, 64, 64, 16, 6>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce);
            break;
        case 17: launch_k<TD, 128, 128, 16, 6>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 18: launch_k<TD, 128, 64, 16, 6>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        case 19: launch_k<TD, 64, 128, 16, 6>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
        default: launch_k<TD, 64, 64, 16, 1>(p, ts, prec, A, B, C, alpha, beta, s, tiles, reduce); break;
    }
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

template <typename TD>
int launch_skinny(const SkinnySet &ss, const TD *A, const TD *B, TD *C, double alpha, double beta, hipStream_t s) {
    if (ss.ntiles == 0) return TKMI_OK;
    if (ss.ntiles * NT >= ((int64_t)1 << 31))
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: more skinny tiles than one launch can hold (2^31 work-items)");
    dim3 grid((unsigned)ss.ntiles), block(NT);
    const int64_t ng = (int64_t)ss.groups.size();
    if (ss.sk <= 8)
        hipLaunchKernelGGL((gemm_skinny_kernel<8, TD>), grid, block, 0, s, ss.d_groups, ng, A, B, C, alpha, beta);
    else if (ss.sk <= 16)
        hipLaunchKernelGGL((gemm_skinny_kernel<16, TD>), grid, block, 0, s, ss.d_groups, ng, A, B, C, alpha, beta);
    else
        hipLaunchKernelGGL((gemm_skinny_kernel<32, TD>), grid, block, 0, s, ss.d_groups, ng, A, B, C, alpha, beta);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

template <typename TD>
int launch_dots(const DotSet &ds, const TD *A, const TD *B, TD *C, double alpha, double beta, hipStream_t s) {
    if (ds.groups.empty()) return TKMI_OK;
    hipLaunchKernelGGL(gemm_dot_partial<TD>, dim3((unsigned)ds.nparts), dim3(NT), 0, s, ds.d_groups,
                       (int64_t)ds.groups.size(), A, B, ds.parts);
    hipLaunchKernelGGL(gemm_dot_final<TD>, dim3((unsigned)ds.groups.size()), dim3(NT), 0, s, ds.d_groups, ds.parts,
                       C, alpha, beta);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

void free_gemm(tkmi_gemm_plan_s *p) {
    if (!p) return;
    (void)hipFree(p->zts.d_groups);
    (void)hipFree(p->zts.d_rgroups);
    (void)hipFree(p->zts.ws);
    (void)hipFree(p->d_zscale);
    (void)hipFree(p->panels);
    (void)hipFree(p->d_pack);
    (void)hipFree(p->skinny.d_groups);
    (void)hipFree(p->dots.d_groups);
    (void)hipFree(p->dots.parts);
    for (auto *d : p->d_scale) (void)hipFree(d);
    for (auto &t : p->ts) {
        (void)hipFree(t.d_groups);
        (void)hipFree(t.d_rgroups);
        (void)hipFree(t.ws);
        (void)hipFree(t.d_sched);
        (void)hipFree(t.d_slot_start);
    }
    delete p;
}

int check_prec(int32_t precision) {
    if (precision < 0 || precision > 2) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: unknown precision");
    return TKMI_OK;
}

int check_dtype(int32_t precision, int32_t dtype) {
    if (int rc = check_prec(precision)) return rc;
    if (dtype != TKMI_F64 && dtype != TKMI_F32) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: unknown element type");
    if (dtype == TKMI_F32 && precision == TKMI_PREC_F64)
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: FP32 data runs in TKMI_PREC_F32 or TKMI_PREC_BF16X3");
    return TKMI_OK;
}

template <typename TD>
int exec_parts(tkmi_gemm_plan plan, int32_t precision, int32_t parts, double alpha, double beta, const TD *A,
               const TD *B, TD *C, hipStream_t s) {
    int rc = TKMI_OK;
    if (alpha == 0.0) {
        // BLAS convention: A and B are not read, C = β·C (β = 0 writes zeros)
        for (int q = 0; q < 2; ++q) {
            if (!(parts & (1 << q)) || plan->scale_total[q] == 0) continue;
            const int64_t blocks = std::min<int64_t>(plan->scale_total[q] / RCH, (int64_t)1 << 20);
            hipLaunchKernelGGL(splitk_reduce<TD>, dim3((unsigned)blocks), dim3(NT), 0, s, plan->d_scale[q],
                               (int64_t)plan->scale[q].size(), plan->scale_total[q], (const double *)nullptr, C,
                               0.0, beta);
        }
        TKMI_HIP(hipGetLastError());
        return TKMI_OK;
    }
    // bit 2: the MFMA launch without its split-K reduction; bit 3: that reduction alone
    // (timing only: either leaves the split groups of C without their sums)
    const bool tiles = (parts & 1) != 0, reduce = (parts & 1) ? !(parts & 4) : (parts & 8) != 0;
    if (tiles || reduce)
        rc = launch_cfg<TD>(plan, cfg_of(plan, precision), precision, A, B, C, alpha, beta, s, tiles, reduce);
    if (rc) return rc;
    if (parts & 2) rc = launch_skinny<TD>(plan->skinny, A, B, C, alpha, beta, s);
    if (!rc && (parts & 2)) rc = launch_dots<TD>(plan->dots, A, B, C, alpha, beta, s);
    return rc;
}

template <typename TC, typename TD>
int zexec(tkmi_gemm_plan p, double ar, double ai, double br, double bi, int conja, int conjb, const TD *A,
          const TD *B, TD *C, hipStream_t s) {
    if (!p->zready) {
        // first complex call on this plan: its tiles (not during graph capture)
        if (int rc = build_ztiles(p)) return rc;
        p->zready = true;
    }
    if (ar == 0.0 && ai == 0.0) {
        // BLAS convention: A and B are not read, C = β·C (β = 0 writes zeros)
        if (p->zscale_total > 0) {
            const int64_t blocks = std::min<int64_t>((p->zscale_total + NT - 1) / NT, 8192);
            hipLaunchKernelGGL(zreduce<TD>, dim3((unsigned)blocks), dim3(NT), 0, s, p->d_zscale,
                               (int64_t)p->zscale.size(), p->zscale_total, (const double *)nullptr, C, 0.0, 0.0,
                               br, bi);
        }
        TKMI_HIP(hipGetLastError());
        return TKMI_OK;
    }
    const TileSet &ts = p->zts;
    if (ts.ntiles == 0) return TKMI_OK;
    if (ts.ntiles * NT >= ((int64_t)1 << 31))
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: more output tiles than one launch can hold (2^31 work-items)");
    hipLaunchKernelGGL((gemm_z_kernel<TC, TD, ZBM, ZBN, ZBK>), dim3((unsigned)ts.ntiles), dim3(NT), 0, s,
                       ts.d_groups, (int64_t)ts.groups.size(), ts.ntiles, A, B, C, ts.ws, ar, ai, br, bi, conja,
                       conjb);
    if (ts.relems > 0) {
        const int64_t blocks = std::min<int64_t>((ts.relems + NT - 1) / NT, 8192);
        hipLaunchKernelGGL(zreduce<TD>, dim3((unsigned)blocks), dim3(NT), 0, s, ts.d_rgroups,
                           (int64_t)ts.rgroups.size(), ts.relems, (const double *)ts.ws, C, ar, ai, br, bi);
    }
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

template <typename TD>
int autotune(tkmi_gemm_plan plan, int32_t precision, const TD *A, const TD *B, TD *C, int32_t reps,
             hipStream_t s) {
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    TKMI_HIP(hipEventCreate(&e0));
    TKMI_HIP(hipEventCreate(&e1));
    // Interleaved rounds, best time per configuration (drift and clock ramps hit every
    // configuration alike); another configuration replaces the default only when it
    // is faster by more than 3%, so near-ties resolve the same way on every run.
    constexpr int ROUNDS = 3;
    float tbest[NCFG];
    for (int c = 0; c < NCFG; ++c) tbest[c] = 1e30f;
    int rc = TKMI_OK;
    for (int round = 0; round < ROUNDS && rc == TKMI_OK; ++round) {
        for (int c = 0; c < NCFG && rc == TKMI_OK; ++c) {
            if (plan->ts[c].ntiles == 0) continue;
            if (precision == TKMI_PREC_BF16X3 && bf16_duplicate(c)) continue;
            // the direct-load configurations run configuration 5's kernel where they do not apply
            if (direct_cfg(c) &&
                ((precision != TKMI_PREC_F64 && wide_direct(c)) ||
                 (precision == TKMI_PREC_F32 && (!std::is_same<TD, float>::value || !plan->ts[c].direct_ok)) ||
                 (precision == TKMI_PREC_F64 && (!std::is_same<TD, double>::value || !plan->ts[c].direct_ok)) ||
                 (precision == TKMI_PREC_BF16X3 && !plan->ts[c].bf16_direct_ok)))
                continue;
            if ((rc = launch_cfg<TD>(plan, c, precision, A, B, C, 1.0, 0.0, s))) break;
            if (hipEventRecord(e0, s) != hipSuccess) { rc = TKMI_ERR_HIP; break; }
            for (int r = 0; r < reps && rc == TKMI_OK; ++r) rc = launch_cfg<TD>(plan, c, precision, A, B, C, 1.0, 0.0, s);
            if (rc) break;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = TKMI_ERR_HIP; break; }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tbest[c] = std::min(tbest[c], ms);
        }
    }
    const int d = auto_cfg(plan, precision);
    int best_c = d;
    for (int c = 0; c < NCFG; ++c)
        if (tbest[c] < tbest[best_c]) best_c = c;
    if (tbest[d] <= 1.03f * tbest[best_c]) best_c = d;  // within 3% of the fastest: keep the default
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    plan->choice[precision] = best_c;
    return TKMI_OK;
}

}  // namespace

extern "C" {

int tkmi_gemm_plan_create(tkmi_gemm_plan *plan, int64_t ngroups, const int64_t *m, const int64_t *n,
                          const int64_t *k, const int64_t *a_off, const int64_t *lda, const int32_t *transa,
                          const int64_t *b_off, const int64_t *ldb, const int32_t *transb, const int64_t *c_off,
                          const int64_t *ldc) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_gemm_plan_create: null plan pointer");
    *plan = nullptr;
    if (ngroups < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_gemm_plan_create: negative group count");
    auto *p = new tkmi_gemm_plan_s();
    // When the plan launches the MFMA kernel anyway, skinny-shaped groups that fit in a
    // few 64×64 tiles ride along in it instead of costing a launch of their own; the
    // skinny kernel keeps the plans made of skinny groups only (the MPO steps).
    auto skinny_shape = [](int64_t m, int64_t n, int64_t k) { return std::min(m, n) <= SKINNY && k <= SKINNY_K; };
    bool any_mfma = false;
    auto dot_shape = [](int64_t m, int64_t n, int64_t k) { return m == 1 && n == 1 && k > SKINNY_K; };
    for (int64_t g = 0; g < ngroups; ++g)
        if (m[g] > 0 && n[g] > 0 && k[g] > 0 && !skinny_shape(m[g], n[g], k[g]) && !dot_shape(m[g], n[g], k[g]))
            any_mfma = true;
    for (int64_t g = 0; g < ngroups; ++g) {
        GGroup G;
        std::memset(&G, 0, sizeof(G));
        G.m = m[g];
        G.n = n[g];
        G.k = k[g];
        G.a_off = a_off[g];
        G.lda = lda[g];
        G.ta = transa[g] ? 1 : 0;
        G.b_off = b_off[g];
        G.ldb = ldb[g];
        G.tb = transb[g] ? 1 : 0;
        G.c_off = c_off[g];
        G.ldc = ldc[g];
        if (G.m < 0 || G.n < 0 || G.k < 0 || G.ldc < std::max<int64_t>(1, G.m)) {
            free_gemm(p);
            TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_gemm_plan_create: inconsistent group dimensions");
        }
        const bool few_tiles = ((G.m + 63) / 64) * ((G.n + 63) / 64) <= 4;
        const bool to_skinny = skinny_shape(G.m, G.n, G.k) && !(any_mfma && few_tiles);
        if (G.k > 0 && G.m > 0 && G.n > 0 && !to_skinny && !dot_shape(G.m, G.n, G.k) &&
            (G.lda * (G.ta ? 128 : 32) >= ((int64_t)1 << 29) || G.ldb * (G.tb ? 32 : 128) >= ((int64_t)1 << 29))) {
            // the MFMA kernels address a slab with 32-bit byte offsets from its base
            free_gemm(p);
            TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "tkmi_gemm_plan_create: leading dimension too large for 32-bit tile offsets");
        }
        if (G.k > 0) {
            const int64_t need_a = G.ta ? G.k : G.m, need_b = G.tb ? G.n : G.k;
            if (G.lda < std::max<int64_t>(1, need_a) || G.ldb < std::max<int64_t>(1, need_b)) {
                free_gemm(p);
                TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_gemm_plan_create: leading dimension too small");
            }
        }
        const double f = 2.0 * (double)G.m * (double)G.n * (double)G.k;
        p->flops += f;
        if (G.m == 0 || G.n == 0) continue;
        p->zall.push_back(G);
        if (dot_shape(G.m, G.n, G.k)) {
            // a dot product: op(A) is 1×k, op(B) is k×1
            constexpr int64_t DOT_CHUNK = 8192, DOT_MAXPARTS = 1024;
            DGroup D;
            std::memset(&D, 0, sizeof(D));
            D.k = G.k;
            D.a_off = G.a_off;
            D.a_inc = G.ta ? 1 : G.lda;
            D.b_off = G.b_off;
            D.b_inc = G.tb ? G.ldb : 1;
            D.c_off = G.c_off;
            D.nparts = std::min<int64_t>(DOT_MAXPARTS, (G.k + DOT_CHUNK - 1) / DOT_CHUNK);
            D.chunk = (G.k + D.nparts - 1) / D.nparts;
            D.part_start = p->dots.nparts;
            p->dots.nparts += D.nparts;
            p->dots.groups.push_back(D);
            p->flops_skinny += f;
            p->skinny_c.push_back(G);  // its C element, for α = 0 calls
            continue;
        }
        if (to_skinny) {
            p->flops_skinny += f;
            G.skinny_mode = (G.n <= G.m) ? 0 : 1;
            const int64_t L = G.skinny_mode == 0 ? G.m : G.n;
            G.tiles_m = (L + NT - 1) / NT;
            G.tiles_n = 1;
            G.tile_start = p->skinny.ntiles;
            p->skinny.ntiles += G.tiles_m;
            p->skinny.sk = (int)std::max<int64_t>(p->skinny.sk, std::min(G.m, G.n));
            p->skinny.groups.push_back(G);
        } else {
            p->flops_mfma += f;
            p->base.push_back(G);
        }
    }
    int rc = TKMI_OK;
    if (!p->skinny.groups.empty())
        rc = tkmi::upload((void **)&p->skinny.d_groups, p->skinny.groups.data(),
                          p->skinny.groups.size() * sizeof(GGroup));
    if (!rc && !p->dots.groups.empty()) {
        rc = tkmi::upload((void **)&p->dots.d_groups, p->dots.groups.data(), p->dots.groups.size() * sizeof(DGroup));
        if (!rc) rc = tkmi::upload((void **)&p->dots.parts, nullptr, (size_t)p->dots.nparts * sizeof(double));
    }
    // largest groups first: their tiles start earliest
    std::stable_sort(p->base.begin(), p->base.end(), [](const GGroup &a, const GGroup &b) {
        return a.m * a.n * std::max<int64_t>(a.k, 1) > b.m * b.n * std::max<int64_t>(b.k, 1);
    });
    // 3×BF16 panels: rows (A) and columns (B) padded to 128, K to 32
    for (auto &G : p->base) {
        G.kp = (G.k + 31) / 32 * 32;
        const int64_t mp = (G.m + 127) / 128 * 128, np = (G.n + 127) / 128 * 128;
        G.sa = mp * G.kp;
        G.sb = np * G.kp;
        G.pa = p->panel_len;
        p->panel_len += 3 * G.sa;
        G.pb = p->panel_len;
        p->panel_len += 3 * G.sb;
        PackGroup P;
        std::memset(&P, 0, sizeof(P));
        P.m = G.m; P.n = G.n; P.k = G.k;
        P.a_off = G.a_off; P.lda = G.lda; P.ta = G.ta;
        P.b_off = G.b_off; P.ldb = G.ldb; P.tb = G.tb;
        P.mp = mp; P.np = np; P.kp = G.kp;
        P.pa = G.pa; P.pb = G.pb; P.sa = G.sa; P.sb = G.sb;
        const int64_t tk = (G.kp + PACK_T - 1) / PACK_T;  // K tiles of the pack pass
        P.start = p->pack_total;
        P.nA = (mp / PACK_T) * tk;
        P.nB = (np / PACK_T) * tk;
        P.startA = p->pack_total_a;
        P.startB = p->pack_total_b;
        p->pack_total += ((mp + np) / PACK_T) * tk;
        p->pack_total_a += P.nA;
        p->pack_total_b += P.nB;
        p->pack.push_back(P);
    }
    // C blocks of each part for α = 0 calls
    for (int q = 0; q < 2; ++q) {
        for (const std::vector<GGroup> *list : q == 0 ? std::vector<const std::vector<GGroup> *>{&p->base}
                                                      : std::vector<const std::vector<GGroup> *>{&p->skinny.groups,
                                                                                                 &p->skinny_c}) {
            for (const GGroup &G : *list) {
                if (G.m * G.n == 0) continue;
                p->scale[q].push_back(RGroup{G.m, G.n, G.c_off, G.ldc, 0, 0, p->scale_total[q]});
                p->scale_total[q] += rch_round(G.m * G.n);
            }
        }
        if (!rc && !p->scale[q].empty())
            rc = tkmi::upload((void **)&p->d_scale[q], p->scale[q].data(), p->scale[q].size() * sizeof(RGroup));
    }
    for (int c = 0; c < NCFG && !rc; ++c) rc = build_tiles(p, c);
    if (rc) {
        free_gemm(p);
        return rc;
    }
    *plan = p;
    return TKMI_OK;
}

int tkmi_gemm_plan_set_tile(tkmi_gemm_plan plan, int32_t precision, int32_t tile) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_prec(precision)) return rc;
    int c;
    if (tile == 0) c = -1;
    else if (tile == 64) c = 0;
    else if (tile == 128) c = 1;
    else if (tile >= 1000 && tile < 1000 + NCFG) c = tile - 1000;
    else TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: tile must be 0, 64, 128 or 1000 + configuration index");
    plan->choice[precision] = c;
    return TKMI_OK;
}

int tkmi_gemm_config_info(int32_t config, int32_t *bm, int32_t *bn, int32_t *bk, int32_t *nbuf) {
    if (config < 0 || config >= NCFG) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: unknown configuration");
    if (bm) *bm = CFGS[config].bm;
    if (bn) *bn = CFGS[config].bn;
    if (bk) *bk = CFGS[config].bk;
    if (nbuf) *nbuf = CFGS[config].nbuf;
    return TKMI_OK;
}

int tkmi_gemm_exec_parts(tkmi_gemm_plan plan, int32_t precision, int32_t parts, double alpha, double beta,
                         const double *A, const double *B, double *C, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_prec(precision)) return rc;
    return exec_parts<double>(plan, precision, parts, alpha, beta, A, B, C, tkmi::as_stream(stream));
}

int tkmi_gemm_exec(tkmi_gemm_plan plan, int32_t precision, double alpha, double beta, const double *A,
                   const double *B, double *C, void *stream) {
    return tkmi_gemm_exec_parts(plan, precision, 3, alpha, beta, A, B, C, stream);
}

int tkmi_gemm_exec_dtype(tkmi_gemm_plan plan, int32_t precision, int32_t dtype, int32_t parts, double alpha,
                         double beta, const void *A, const void *B, void *C, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_dtype(precision, dtype)) return rc;
    hipStream_t s = tkmi::as_stream(stream);
    if (dtype == TKMI_F32)
        return exec_parts<float>(plan, precision, parts, alpha, beta, (const float *)A, (const float *)B, (float *)C, s);
    return exec_parts<double>(plan, precision, parts, alpha, beta, (const double *)A, (const double *)B, (double *)C, s);
}

int tkmi_gemm_exec_z(tkmi_gemm_plan plan, int32_t precision, int32_t dtype, int32_t conj_a, int32_t conj_b,
                     double alpha_re, double alpha_im, double beta_re, double beta_im, const void *A, const void *B,
                     void *C, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_prec(precision)) return rc;
    if (precision == TKMI_PREC_BF16X3)
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "gemm: complex data runs in TKMI_PREC_F64 or TKMI_PREC_F32");
    if (dtype != TKMI_C128 && dtype != TKMI_C64) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: dtype must be TKMI_C128 or TKMI_C64");
    hipStream_t s = tkmi::as_stream(stream);
    const int ca = conj_a != 0, cb = conj_b != 0;
    if (dtype == TKMI_C128) {
        auto a = (const z2_t *)A, b = (const z2_t *)B;
        auto c = (z2_t *)C;
        return precision == TKMI_PREC_F64
                   ? zexec<double, z2_t>(plan, alpha_re, alpha_im, beta_re, beta_im, ca, cb, a, b, c, s)
                   : zexec<float, z2_t>(plan, alpha_re, alpha_im, beta_re, beta_im, ca, cb, a, b, c, s);
    }
    auto a = (const c2_t *)A, b = (const c2_t *)B;
    auto c = (c2_t *)C;
    return precision == TKMI_PREC_F64
               ? zexec<double, c2_t>(plan, alpha_re, alpha_im, beta_re, beta_im, ca, cb, a, b, c, s)
               : zexec<float, c2_t>(plan, alpha_re, alpha_im, beta_re, beta_im, ca, cb, a, b, c, s);
}

int tkmi_gemm_plan_autotune(tkmi_gemm_plan plan, int32_t precision, const double *A, const double *B, double *C,
                            int32_t reps, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_prec(precision)) return rc;
    return autotune<double>(plan, precision, A, B, C, reps, tkmi::as_stream(stream));
}

int tkmi_gemm_plan_autotune_dtype(tkmi_gemm_plan plan, int32_t precision, int32_t dtype, const void *A,
                                  const void *B, void *C, int32_t reps, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_dtype(precision, dtype)) return rc;
    hipStream_t s = tkmi::as_stream(stream);
    if (dtype == TKMI_F32)
        return autotune<float>(plan, precision, (const float *)A, (const float *)B, (float *)C, reps, s);
    return autotune<double>(plan, precision, (const double *)A, (const double *)B, (double *)C, reps, s);
}

int tkmi_gemm_plan_set_constant(tkmi_gemm_plan plan, int32_t operands) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (operands < 0 || operands > 3) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: operands is a mask of 1 (A) and 2 (B)");
    plan->const_ops = operands;
    if (!(operands & 1)) plan->split_src[0] = nullptr;
    if (!(operands & 2)) plan->split_src[1] = nullptr;
    return TKMI_OK;
}

int tkmi_gemm_plan_invalidate(tkmi_gemm_plan plan) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    plan->split_src[0] = plan->split_src[1] = nullptr;
    return TKMI_OK;
}

int tkmi_gemm_plan_query(tkmi_gemm_plan plan, int32_t precision, int32_t *config, int64_t *ntiles) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (int rc = check_prec(precision)) return rc;
    const int c = cfg_of(plan, precision);
    if (config) *config = c;
    if (ntiles) *ntiles = plan->ts[c].ntiles;
    return TKMI_OK;
}

int tkmi_gemm_plan_time_next(tkmi_gemm_plan plan, void *start_event, void *stop_event) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (!start_event != !stop_event) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: time_next needs both events or neither");
    plan->time_ev[0] = (hipEvent_t)start_event;
    plan->time_ev[1] = (hipEvent_t)stop_event;
    return TKMI_OK;
}

int tkmi_calib_mfma_f64(int64_t blocks, int32_t iters, double *out, void *stream) {
    if (blocks < 1 || blocks > ((int64_t)1 << 20) || iters < 1 || !out)
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "calib: blocks in [1, 2^20], iters >= 1 and an output buffer");
    hipLaunchKernelGGL(mfma_f64_calib_kernel, dim3((unsigned)blocks), dim3(NT), 0, (hipStream_t)stream, out, iters,
                       1.0);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

int tkmi_gemm_plan_split(tkmi_gemm_plan plan, double *flops_mfma, double *flops_skinny, int64_t *skinny_groups) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (flops_mfma) *flops_mfma = plan->flops_mfma;
    if (flops_skinny) *flops_skinny = plan->flops_skinny;
    if (skinny_groups) *skinny_groups = (int64_t)plan->skinny.groups.size();
    return TKMI_OK;
}

int tkmi_gemm_plan_splitk(tkmi_gemm_plan plan, int32_t config, int64_t *splitk, int64_t *slots) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (config < 0 || config >= NCFG) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: no such tile configuration");
    if (splitk) *splitk = plan->ts[config].split;
    if (slots) *slots = cfg_slots(config);
    return TKMI_OK;
}

int tkmi_gemm_plan_info(tkmi_gemm_plan plan, int64_t *ntiles, double *flops) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "gemm: null plan");
    if (ntiles) *ntiles = plan->ts[0].ntiles;
    if (flops) *flops = plan->flops;
    return TKMI_OK;
}

int tkmi_gemm_plan_destroy(tkmi_gemm_plan plan) {
    free_gemm(plan);
    return TKMI_OK;
}

}  // extern "C"
