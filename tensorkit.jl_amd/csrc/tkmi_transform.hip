// Grouped permute / braid of block-sparse tensors (K1 abelian, K2 generic recoupling).
//
// Host-side plan creation canonicalises each work item:
//   * dims of size 1 are dropped, dims that stay contiguous on both sides are merged;
//   * dW = the dim with the smallest stride on the written side,
//     dR = the dim with the smallest stride on the read side;
//   * dW == dR  -> "copy" mode: the plane of dW and the dim with the next smallest
//                  written stride (stored as dR) is numbered element after element
//                  along dW (contiguous on both sides) and a tile is the next `area`
//                  elements of that numbering, across row ends, so every tile but an
//                  item's last is full; lanes consecutive on both sides;
//     dW != dR  -> "transpose" mode: a (tW × tR) tile (powers of two) is read with
//                  lanes along dR, staged in LDS, and written with lanes along dW;
//                  when tW·tR leaves the tile's area unused (small dims), the tile
//                  takes tB slices along a third dim dB as well;
//     dW == dR with a short row (≤ 24 elements) and different next dims on the two
//                  sides -> "vector transpose": the short dim dV is carried whole and
//                  the tile is a tW × tR transpose of those vectors (tR along the read
//                  side's next dim, tW along the write side's), staged in LDS; both
//                  sides then move runs of |dV|·tR or |dV|·tW elements instead of rows
//                  of |dV|.
// Work items are split into classes by the number of trees they touch (L = 1, 2, 4,
// 8); each class is one launch of transform_kernel<T, L>.  Inside a launch the tiles
// of all items are numbered consecutively (tile_start prefix sums) and each
// workgroup finds its item by binary search: a device-side block schedule.
//
// Every thread owns E elements of each of the L trees of a tile (E = 8 for L = 1,
// else 16/L), all loaded before the first LDS store; the tile area is 256·E
// elements per tree.  A launch that writes more than KEEP_BYTES streams: its loads and
// stores are non-temporal (every element is read once; non-temporal stores are 11%
// faster on a 4.8 GB permute).  A smaller launch loads and stores normally, so that it
// finds its source in the 256 MB Infinity Cache where the kernel before left it and
// leaves its result there for the kernel after (measured on the U(1) D = 2052 two-site
// H_eff chain: cached stores 2% faster per application, cached loads 3%, the fused MPO
// steps 31 and 35 → 27 µs).
//
// Generic (non-abelian) blocks carry R destination and C source trees of identical
// shape plus a dense R × C recoupling matrix U; each written element is
//   β·dst_j + Σ_i (α·U[j,i])·src_i,
// i.e. the reference's pack → buffer_src·Uᵀ → unpack (indexmanipulations.jl:692-721)
// fused in registers.  Destination trees are dealt out in chunks of at most 8 (one
// work item each); source trees are consumed in chunks of L inside the kernel with
// the sums kept in registers, so a fusion block of any size runs in one pass.
#include "tkmi_internal.h"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr int MAXD = 12;
constexpr int NTHREADS = 256;
constexpr int NCLASS = 4;  // L = 1, 2, 4, 8
constexpr int NSIDE = 2 * NCLASS;  // per class: the other tiles, then the vector-transpose tiles
constexpr int MAXW = 8;    // destination trees per work item
constexpr int64_t SMALL_PLAN = (int64_t)1 << 25;  // written elements below which classes merge
constexpr int64_t KEEP_BYTES = (int64_t)128 << 20;  // written bytes up to which stores stay cached

struct TBlock {
    int32_t ndim;
    int32_t mode;           // 0 copy, 1 transpose, 2 transpose of short vectors along dV
    int32_t nread, nwrite;  // nwrite ≤ L; nread may exceed L (chunked in the kernel)
    int32_t dW, dR;
    int32_t tW, tR;  // tile extents (transpose mode: powers of two)
    int32_t lgW, lgR;
    int32_t nother;
    int32_t u_ld;  // leading dimension of the (column-major) coefficient matrix
    int32_t dV;    // mode 2: the short dim, innermost on both sides, carried whole in a tile
    int32_t off32; // every element of a tile lies within ±2^31 elements of the tile's corner in
                   // every tree: offsets from the corner are formed in 32-bit arithmetic
    double inv_w;  // copy mode: 1 / size[dW], for the row of a tile's first element
    int32_t odims[MAXD];
    int64_t size[MAXD];
    int64_t ntW, ntR, nbatch;
    int64_t tile_start;
    int64_t read_tree, write_tree;
    int64_t u_start;
};

static_assert(sizeof(TBlock) % sizeof(int64_t) == 0, "the compact tile_start array follows the blocks");

struct TTree {
    int64_t off;
    int64_t str[MAXD];
};

// per-class tile geometry, shared by the kernel and the host tiler: E elements per
// thread per tree (8 for single-tree items, else 16/L), tile area 256·E per tree
// (measured on 82.8 GB SU(2) permutes: larger areas lose more to occupancy than
// they gain in loads in flight)
constexpr int class_elems(int L) { return L == 1 ? 8 : 16 / L; }
constexpr int class_area(int L) { return NTHREADS * class_elems(L); }
constexpr int class_twmax(int L) { return L <= 2 ? 256 : (L == 4 ? 128 : 64); }
constexpr int class_L(int c) { return 1 << c; }

// β·c + v with β·c and the sum rounded separately, as the reference's tensoradd!
// (scale(C, β) + scale(A, α): Julia does not fuse a multiply and an add).  The empty asm
// statement takes the rounded product as a register operand, so the compiler cannot fuse
// the multiply into the add; it emits no instruction.  (A `#pragma clang fp contract(off)`
// here switched contraction off in the whole calling kernel, the recoupling sums too.)
__device__ __forceinline__ double scaled_sum(double beta, double c, double v) {
    double bc = beta * c;
    asm volatile("" : "+v"(bc));
    return bc + v;
}

// Complex elements: (re, im) pairs, accumulated in FP64 pairs.  The recoupling
// coefficients are real; α, β and the conjugation of the source may be complex / set.
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

template <typename T>
struct Acc {
    typedef double type;
    static constexpr bool cplx = false;
};
template <>
struct Acc<f2_t> {
    typedef d2_t type;
    static constexpr bool cplx = true;
};
template <>
struct Acc<d2_t> {
    typedef d2_t type;
    static constexpr bool cplx = true;
};

__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ d2_t widen(d2_t v) { return v; }
__device__ __forceinline__ d2_t widen(f2_t v) { return __builtin_convertvector(v, d2_t); }
template <typename T>
__device__ __forceinline__ T narrow(typename Acc<T>::type v) {
    if constexpr (std::is_same<T, f2_t>::value)
        return __builtin_convertvector(v, f2_t);
    else
        return (T)v;
}
__device__ __forceinline__ double fma_acc(double u, double x, double o) { return fma(u, x, o); }
__device__ __forceinline__ d2_t fma_acc(double u, d2_t x, d2_t o) { return d2_t{fma(u, x.x, o.x), fma(u, x.y, o.y)}; }
template <typename T>
__device__ __forceinline__ T conj_of(T v) {
    if constexpr (Acc<T>::cplx) v.y = -v.y;
    return v;
}

// (a.x + i a.y)·(b.x + i b.y) with the four products rounded before they are combined,
// as Julia's complex multiply (no fused multiply-add)
__device__ __forceinline__ d2_t cmul(d2_t a, d2_t b) {
    double p1 = a.x * b.x, p2 = a.y * b.y, p3 = a.x * b.y, p4 = a.y * b.x;
    asm volatile("" : "+v"(p1), "+v"(p2), "+v"(p3), "+v"(p4));
    return d2_t{p1 - p2, p3 + p4};
}

// The scalars of one transform call: α = ar + i·ai, β = br + i·bi (imaginary parts zero for
// real data) and whether the source is read conjugated (complex data only).
struct Scal {
    double ar, ai, br, bi;
    int conj;
};

// One tile of one work item.  COPY: the tile is nr runs of nw elements along dW
// (contiguous on both sides), numbered run after run; otherwise a tW × tR tile
// read with lanes along dR and written with lanes along dW through LDS.
template <typename T, int L, bool COPY, bool SLICED, bool VEC = false>
__device__ __forceinline__ void transform_tile(const TBlock &B, const TTree *__restrict__ trees,
                                               const double *__restrict__ Umat, const T *__restrict__ src,
                                               T *__restrict__ dst, const Scal &sc, bool keep,
                                               int64_t t, T *__restrict__ lds, int dB_in, int lgB) {
    using A = typename Acc<T>::type;
    const double alpha = sc.ar, beta = sc.br;
    constexpr int E = class_elems(L);
    const int nread = B.nread, nwrite = B.nwrite;
    const int dW = B.dW, dR = B.dR;
    const int tW = B.tW, tR = B.tR;

    const int64_t iW = t % B.ntW;
    t /= B.ntW;
    const int64_t iR = t % B.ntR;
    t /= B.ntR;
    // slices along dB only in the SLICED instantiation (the others keep their registers);
    // a vector-transpose tile carries its short dim dV whole, addressed like a slice dim
    const int dB = SLICED ? dB_in : (VEC ? B.dV : -1);
    const int tB = SLICED ? 1 << lgB : 1;
    int64_t iB = 0;
    if (SLICED && dB >= 0) {
        const int64_t ntB = (B.size[dB] + tB - 1) >> lgB;
        iB = t % ntB;
        t /= ntB;
    }
    const int64_t bt = t;
    const int64_t b0 = iB * tB;
    const int nb = (SLICED && dB >= 0) ? (int)min((int64_t)tB, B.size[dB] - b0) : 1;
    int64_t w0, r0;
    int nw, nr;
    // copy mode: tile iR holds elements [iR·area, iR·area + cnt) of the item's plane
    // numbered along dW, starting at row r0, column w0 (r0 from the reciprocal, then
    // corrected); element e of the tile is e places further along that numbering
    const int64_t W = B.size[dW];
    int cnt = 0;
    if (COPY) {
        constexpr int AREA = class_area(L);
        const int64_t f0 = iR * AREA;
        int64_t rs = (int64_t)((double)f0 * B.inv_w);
        if (rs * W > f0) --rs;
        else if ((rs + 1) * W <= f0) ++rs;
        r0 = rs;
        w0 = f0 - rs * W;
        cnt = (int)min((int64_t)AREA, W * B.size[dR] - f0);
        nw = nr = 0;
    } else {
        w0 = iW * tW;
        r0 = iR * tR;
        nw = (int)min((int64_t)tW, B.size[dW] - w0);
        nr = (int)min((int64_t)tR, B.size[dR] - r0);
    }
    const int tid = threadIdx.x;

    // offset of the tile corner in one tree (the batch index bt is decomposed
    // again per tree: cheaper than holding the multi-index in scalar registers)
    auto corner = [&](const TTree &tr) {
        int64_t o = tr.off + w0 * tr.str[dW] + r0 * tr.str[dR] + (dB >= 0 ? b0 * tr.str[dB] : 0);
        int64_t b = bt;
        for (int k = 0; k < B.nother; ++k) {
            const int d = B.odims[k];
            const int64_t sz = B.size[d];
            int64_t q;
            if ((uint64_t)b < 0x80000000ull)
                q = (int64_t)((uint32_t)b / (uint32_t)sz);  // sz ≤ b here or q = 0
            else
                q = b / sz;
            o += (b - q * sz) * tr.str[d];
            b = q;
        }
        return o;
    };

    // Element k of this thread sits at (w, r) inside the tile.  Transpose tiles
    // have tW, tR ≤ NTHREADS, so on the read side r is fixed per thread and w steps
    // by NTHREADS/tR with k (on the write side w is fixed and r steps): addresses
    // advance by a uniform stride and only the loaded values stay live.
    // Copy mode returns w = e and r = q: relative to the tile corner the element sits
    // at e·sW + q·(sR − W·sW) (the copy stride pair below), which keeps both small.
    // Rows shorter than the tile: q from a float reciprocal (w0 + e < 2·area, exact
    // after the ±1 correction); longer rows hold a tile in at most two rows.
    const bool long_rows = W >= class_area(L);
    const int ws = long_rows ? 0 : (int)w0;
    const int wshort = long_rows ? 1 : (int)W;
    const float inv = 1.0f / (float)(COPY ? wshort : nw);
    // vector transpose: element e of the tile is (v, e / V) with v = e mod V; the quotient
    // from a float reciprocal (e < 2048·V ≪ 2^24), corrected by ±1
    const int V = VEC ? (int)B.size[B.dV] : 1;
    const float vinv = VEC ? (float)B.inv_w : 1.0f;
    auto vsplit = [&](int e, int &v) {
        int q = (int)((float)e * vinv);
        const int rem = e - q * V;
        if (rem < 0) --q;
        else if (rem >= V) ++q;
        v = e - q * V;
        return q;
    };
    auto pos_read = [&](int k, int &w, int &r, int &b) {
        b = 0;
        if (VEC) {
            // lanes along (dV, dR): runs of V·tR elements on the read side
            const int q = vsplit(tid + k * NTHREADS, b);
            r = q & (tR - 1);
            w = q >> B.lgR;
            return w < nw && r < nr;
        }
        if (COPY) {
            const int e = tid + k * NTHREADS;
            if (long_rows) {
                r = (int64_t)e >= W - w0;
            } else {
                const int c = ws + e;
                int q = (int)((float)c * inv);
                const int rem = c - q * wshort;
                if (rem < 0) --q;
                else if (rem >= wshort) ++q;
                r = q;
            }
            w = e;
            return e < cnt;
        }
        // lanes along dR; with slices the index along the other side runs over tB
        // slices of tW
        r = tid & (tR - 1);
        const int wv = (tid >> B.lgR) + k * (NTHREADS >> B.lgR);
        if (!SLICED) {
            w = wv;
            return w < nw && r < nr;
        }
        w = wv & (tW - 1);
        b = wv >> B.lgW;
        return w < nw && r < nr && b < nb;
    };
    auto pos_write = [&](int k, int &w, int &r, int &b) {
        if (COPY) return pos_read(k, w, r, b);
        b = 0;
        if (VEC) {
            // lanes along (dV, dW): runs of V·tW elements on the write side
            const int q = vsplit(tid + k * NTHREADS, b);
            w = q & (tW - 1);
            r = q >> B.lgW;
            return w < nw && r < nr;
        }
        w = tid & (tW - 1);
        const int rv = (tid >> B.lgW) + k * (NTHREADS >> B.lgW);
        if (!SLICED) {
            r = rv;
            return w < nw && r < nr;
        }
        r = rv & (tR - 1);
        b = rv >> B.lgR;
        return w < nw && r < nr && b < nb;
    };

    const bool has_alpha = (alpha != 0.0 || sc.ai != 0.0);  // α = 0: the source is not read (BLAS convention)
    // One source tree (every abelian entry, and generic blocks of one tree): the
    // reference scales by α·coeff (tensoradd! with α*coeff, indexmanipulations.jl:658-669,
    // 684-690), so α is folded into the coefficient.  Several source trees: the reference
    // forms Σ_i U[j,i]·src_i (the recoupling GEMM, :710-711) and scales that by α when it
    // inserts the column (:715-720), so α is applied to the sum at the store.
    // Sums are FP64 for every element type: FP32 data with the Float64 scalars of the
    // reference's tensoradd! is promoted to Float64 and rounded once when stored.  Complex
    // data: a complex α·coeff multiplies the source as a complex product.
    const bool single = (nread == 1);
    A o[L][E];
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
        for (int k = 0; k < E; ++k) o[j][k] = (A)0.0;

    // positions of this thread's elements in the tile (the same in every tree)
    int pw[E], pr[E], pb[E];
    bool pok[E];
#pragma unroll
    for (int k = 0; k < E; ++k) pok[k] = pos_read(k, pw[k], pr[k], pb[k]);

    for (int ic = 0; ic < nread && has_alpha; ic += L) {
        const int nc = min(L, nread - ic);
        T x[L][E];
        // Each present tree's E loads without a branch per element: elements outside the
        // tile read the tile's first element (always inside it) and keep that value, which
        // only reaches sums that are never stored (stores and LDS staging test the position
        // again).  Absent trees (i ≥ nc) hold zeros.  The sums below run over all L trees
        // with the absent ones' coefficients zero, so nothing right after a tree's loads
        // needs their values, and the loads of every tree issue before the first sum (a
        // per-tree branch around loads and sums together had it load, sum and wait tree
        // after tree).
#pragma unroll
        for (int i = 0; i < L; ++i) {
            if (i < nc) {
                const TTree &tr = trees[B.read_tree + ic + i];
                const T *sp = src + corner(tr);
                const int64_t sW = tr.str[dW], sR = COPY ? tr.str[dR] - W * tr.str[dW] : tr.str[dR];
                const int64_t sB = dB >= 0 ? tr.str[dB] : 0;
                // non-temporal loads only in launches that stream (see the file header)
                auto ld = [&](const T *p) { return keep ? *p : __builtin_nontemporal_load(p); };
                if (B.off32) {
                    // formed for every position (unsigned: wraps harmlessly outside the tile),
                    // then 0 selected outside it: a select, not a branch per element
                    const uint32_t sw = (uint32_t)sW, sr = (uint32_t)sR, sb = (uint32_t)sB;
#pragma unroll
                    for (int k = 0; k < E; ++k) {
                        const uint32_t ou = (uint32_t)pw[k] * sw + (uint32_t)pr[k] * sr + (uint32_t)pb[k] * sb;
                        const int32_t o = pok[k] ? (int32_t)ou : 0;
                        x[i][k] = ld(&sp[o]);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < E; ++k) {
                        const int64_t o =
                            pok[k] ? (int64_t)pw[k] * sW + (int64_t)pr[k] * sR + (int64_t)pb[k] * sB : 0;
                        x[i][k] = ld(&sp[o]);
                    }
                }
                if constexpr (Acc<T>::cplx)
                    if (sc.conj)
#pragma unroll
                        for (int k = 0; k < E; ++k) x[i][k] = conj_of(x[i][k]);
            } else {
#pragma unroll
                for (int k = 0; k < E; ++k) x[i][k] = (T)0;
            }
        }
        if (!COPY) {
            // stage through LDS: read layout (lanes along dR) -> write layout (lanes along dW).
            // Vector tiles: [w][r][v] with a row pitch of V·tR + 1 per w (odd: the write
            // side's lanes, consecutive in v then w, fall in distinct banks)
            const int pitch = tR + 1;
            const int vpitch = V * tR + 1;
            auto lidx = [&](int i, int w, int r, int b) {
                return VEC ? (i * tW + w) * vpitch + r * V + b : ((i * tB + b) * tW + w) * pitch + r;
            };
            if (ic > 0) __syncthreads();
#pragma unroll
            for (int i = 0; i < L; ++i)
                if (i < nc)
#pragma unroll
                    for (int k = 0; k < E; ++k) {
                        int w, r, b;
                        if (pos_read(k, w, r, b)) lds[lidx(i, w, r, b)] = x[i][k];
                    }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < L; ++i)
                if (i < nc)
#pragma unroll
                    for (int k = 0; k < E; ++k) {
                        int w, r, b;
                        x[i][k] = pos_write(k, w, r, b) ? lds[lidx(i, w, r, b)] : (T)0;
                    }
        }
#pragma unroll
        for (int j = 0; j < L; ++j) {
            if (j < nwrite) {
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    {
                        // absent trees: coefficient 0 and zero sources (+0 added to a sum that
                        // starts at +0 changes nothing, so results stay bit for bit)
                        const double uc = i < nc ? Umat[B.u_start + j + (int64_t)(ic + i) * B.u_ld] : 0.0;
                        if (Acc<T>::cplx && single && sc.ai != 0.0) {
                            if constexpr (Acc<T>::cplx) {
                                const d2_t u = {alpha * uc, sc.ai * uc};
#pragma unroll
                                for (int k = 0; k < E; ++k) o[j][k] += cmul(u, widen(x[i][k]));
                            }
                        } else {
                            // an explicit fused multiply-add per term (per component for
                            // complex data): the same arithmetic for every element type
                            const double u = single ? alpha * uc : uc;
#pragma unroll
                            for (int k = 0; k < E; ++k) o[j][k] = fma_acc(u, widen(x[i][k]), o[j][k]);
                        }
                    }
                }
            }
        }
    }

    // several source trees: α scales the recoupling sum (α = 1 leaves it as it is)
    if (!single && (alpha != 1.0 || sc.ai != 0.0)) {
#pragma unroll
        for (int j = 0; j < L; ++j)
#pragma unroll
            for (int k = 0; k < E; ++k) {
                if constexpr (Acc<T>::cplx)
                    o[j][k] = sc.ai != 0.0 ? cmul(d2_t{alpha, sc.ai}, o[j][k]) : alpha * o[j][k];
                else
                    o[j][k] = alpha * o[j][k];
            }
    }
    const bool has_beta = (beta != 0.0 || sc.bi != 0.0);
    auto with_beta = [&](A v, T c) -> A {
        if constexpr (Acc<T>::cplx) {
            const d2_t cw = widen(c);
            if (sc.bi == 0.0) return d2_t{scaled_sum(beta, cw.x, v.x), scaled_sum(beta, cw.y, v.y)};
            return cmul(d2_t{beta, sc.bi}, cw) + v;
        } else {
            return scaled_sum(beta, (double)c, v);
        }
    };
    // write positions (the read positions for copy tiles), the same in every tree
    int qw[E], qr[E], qb[E];
    bool qok[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (COPY) {
            qw[k] = pw[k];
            qr[k] = pr[k];
            qb[k] = pb[k];
            qok[k] = pok[k];
        } else {
            qok[k] = pos_write(k, qw[k], qr[k], qb[k]);
        }
    }
    auto store_all = [&](auto store) {
#pragma unroll
        for (int j = 0; j < L; ++j) {
            if (j < nwrite) {
                const TTree &tw = trees[B.write_tree + j];
                T *dp = dst + corner(tw);
                const int64_t sW = tw.str[dW], sR = COPY ? tw.str[dR] - W * tw.str[dW] : tw.str[dR];
                const int64_t sB = dB >= 0 ? tw.str[dB] : 0;
                const uint32_t sw = (uint32_t)sW, sr = (uint32_t)sR, sb = (uint32_t)sB;
#pragma unroll
                for (int k = 0; k < E; ++k) {
                    if (qok[k]) {
                        T *a = B.off32 ? dp + (int32_t)((uint32_t)qw[k] * sw + (uint32_t)qr[k] * sr + (uint32_t)qb[k] * sb)
                                       : dp + ((int64_t)qw[k] * sW + (int64_t)qr[k] * sR + (int64_t)qb[k] * sB);
                        store(narrow<T>(has_beta ? with_beta(o[j][k], *a) : o[j][k]), a);
                    }
                }
            }
        }
    };
    if (keep)
        store_all([](T v, T *a) { *a = v; });
    else
        store_all([](T v, T *a) { __builtin_nontemporal_store(v, a); });
}

// STAGED: the launch has transpose tiles, which stage through LDS.  A launch of copy
// tiles only (STAGED = false) declares no LDS, so LDS does not cap how many of its
// workgroups share a CU (for L = 2 and 4 the staging array is 36 KB).
// SLICED: the launch has transpose tiles with slices along a third dim, described per
// block in `slices` (dB | lgB << 8, or -1), which only this instantiation reads.
template <typename T, int L, bool STAGED, bool SLICED>
__global__ __launch_bounds__(NTHREADS) void transform_kernel(
    const TBlock *__restrict__ blocks, int64_t nblocks, const TTree *__restrict__ trees,
    const double *__restrict__ Umat, const T *__restrict__ src, T *__restrict__ dst, Scal sc, int keep,
    int64_t tile0, const int32_t *__restrict__ slices) {
    const int64_t gt = tile0 + (int64_t)blockIdx.x;
    // binary search: last item with tile_start <= gt
    const int64_t *__restrict__ tstart = reinterpret_cast<const int64_t *>(blocks + nblocks);
    int64_t lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= gt)
            lo = mid;
        else
            hi = mid - 1;
    }
    const TBlock &B = blocks[lo];
    if constexpr (STAGED) {
        constexpr int LDSN = L * (class_area(L) + class_twmax(L));  // tW·(tR+1) ≤ area + tW per tree
        __shared__ T lds[LDSN];
        if (B.mode == 0) {
            transform_tile<T, L, true, false>(B, trees, Umat, src, dst, sc, keep, gt - B.tile_start, lds,
                                              -1, 0);
        } else if (SLICED) {
            const int32_t sl = slices[lo];
            transform_tile<T, L, false, true>(B, trees, Umat, src, dst, sc, keep, gt - B.tile_start, lds,
                                              sl < 0 ? -1 : (sl & 0xff), sl < 0 ? 0 : (sl >> 8));
        } else {
            transform_tile<T, L, false, false>(B, trees, Umat, src, dst, sc, keep, gt - B.tile_start, lds,
                                               -1, 0);
        }
    } else {
        transform_tile<T, L, true, false>(B, trees, Umat, src, dst, sc, keep, gt - B.tile_start, nullptr,
                                          -1, 0);
    }
}

// Vector-transpose tiles (mode 2) only: a launch of its own, with the same LDS staging bound.
template <typename T, int L>
__global__ __launch_bounds__(NTHREADS) void transform_vec_kernel(
    const TBlock *__restrict__ blocks, int64_t nblocks, const TTree *__restrict__ trees,
    const double *__restrict__ Umat, const T *__restrict__ src, T *__restrict__ dst, Scal sc, int keep,
    int64_t tile0) {
    const int64_t gt = tile0 + (int64_t)blockIdx.x;
    const int64_t *__restrict__ tstart = reinterpret_cast<const int64_t *>(blocks + nblocks);
    int64_t lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= gt)
            lo = mid;
        else
            hi = mid - 1;
    }
    constexpr int LDSN = L * (class_area(L) + class_twmax(L));
    __shared__ T lds[LDSN];
    transform_tile<T, L, false, false, true>(blocks[lo], trees, Umat, src, dst, sc, keep, gt - blocks[lo].tile_start,
                                             lds, -1, 0);
}

// U[e] = coef[e]·g[idx[e]] (idx < 0: 0): the coefficients of a transform that depend on
// a small device tensor (a contraction with a small operand folded into the transform).
// A launch of its own ahead of the transform: forming them inside every tile puts two
// dependent loads in front of each tile's arithmetic, which measured slower.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void gather_coeffs(const double *__restrict__ coef,
                                                          const int64_t *__restrict__ idx, const T *__restrict__ g,
                                                          int64_t n, double *__restrict__ U) {
    for (int64_t e = (int64_t)blockIdx.x * NTHREADS + threadIdx.x; e < n; e += (int64_t)gridDim.x * NTHREADS) {
        const int64_t i = idx[e];
        U[e] = i < 0 ? 0.0 : coef[e] * (double)g[i];
    }
}

// -------------------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------------------
struct HostItem {               // one work item before tiling
    std::vector<int64_t> size;  // canonical dims (dst order)
    int nread, nwrite;
    int64_t read_tree, write_tree, u_start;
};

int64_t pow2ceil(int64_t x) {
    int64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

int log2i(int64_t x) {
    int l = 0;
    while (((int64_t)1 << l) < x) ++l;
    return l;
}

struct PlanSide {
    std::vector<TBlock> blocks;
    int64_t ntiles = 0;
    TBlock *d_blocks = nullptr;
    bool staged = false;  // some block uses a transpose tile (LDS staging)
    bool reads = false;   // some block reads source trees (blocks with none only scale by β)
    bool sliced = false;  // some transpose tile takes slices along a third dim
    std::vector<int32_t> slices;  // per block: dB | lgB << 8, or -1 (uploaded when sliced)
    int32_t *d_slices = nullptr;
    int64_t written = 0;  // elements written by all classes of this side together
};

}  // namespace

struct tkmi_transform_plan_s {
    int32_t N = 0;
    // optional gather of the forward coefficients: U[e] = gcoef[e]·g[gidx[e]] (0 for gidx < 0),
    // from a device vector given at execution time
    int64_t gather_len = 0;
    double *d_gcoef = nullptr;
    int64_t *d_gidx = nullptr;
    int64_t dst_len = 0, src_len = 0, elements = 0, nitems = 0;
    // every element is read and written at the same offset by the same thread (an
    // elementwise y = β·y + α·x: add!/scale!), so the source may be the destination
    bool inplace = false;
    std::vector<TTree> trees;
    std::vector<double> U, Ut;
    PlanSide fwd[NSIDE], adj[NSIDE];
    TTree *d_trees = nullptr;
    double *d_U = nullptr, *d_Ut = nullptr;
};

namespace {

// Transpose-mode tile extents: powers of two with tW·tR ≤ area, tW ≤ twmax ≤ NTHREADS
// and tR ≤ NTHREADS,
// runs of at least 16 elements where the dims allow it (whole 128-byte lines for
// doubles), chosen to minimise tiles × (fixed cost + tile slots).
void choose_tile(int64_t sW, int64_t sR, int area, int twmax, int &tW, int &tR) {
    const int64_t pW = pow2ceil(sW), pR = pow2ceil(sR);
    double best = 1e300;
    tW = tR = 0;
    for (int64_t r = 1; r <= NTHREADS && r <= pR; r <<= 1) {
        if (r < std::min<int64_t>(16, pR)) continue;
        for (int64_t w = 1; w <= twmax && w <= pW && w * r <= area; w <<= 1) {
            if (w < std::min<int64_t>(16, pW)) continue;
            const double tiles = (double)((sW + w - 1) / w) * (double)((sR + r - 1) / r);
            const double cost = tiles * (256.0 + 0.25 * (double)(w * r));
            const bool better = cost < best ||
                                (cost == best && std::min(w, r) > std::min<int64_t>(tW, tR)) ||
                                (cost == best && std::min(w, r) == std::min<int64_t>(tW, tR) && r > tR);
            if (better) {
                best = cost;
                tW = (int)w;
                tR = (int)r;
            }
        }
    }
}

// Vector-transpose tile extents (mode 2): powers of two, V·tW·tR ≤ area and tW ≤ twmax
// (the LDS staging bound), by the cost model of choose_tile.
void choose_vec_tile(int64_t V, int64_t sW, int64_t sR, int area, int twmax, int &tW, int &tR) {
    const int64_t pW = pow2ceil(sW), pR = pow2ceil(sR);
    double best = 1e300;
    tW = tR = 1;
    for (int64_t r = 1; r <= NTHREADS && r <= pR; r <<= 1)
        for (int64_t w = 1; w <= twmax && w <= pW && V * w * r <= area; w <<= 1) {
            const double tiles = (double)((sW + w - 1) / w) * (double)((sR + r - 1) / r);
            const double cost = tiles * (256.0 + 0.25 * (double)(V * w * r));
            if (cost < best || (cost == best && std::min(w, r) > std::min<int64_t>(tW, tR))) {
                best = cost;
                tW = (int)w;
                tR = (int)r;
            }
        }
}

// rows up to this many elements take the vector transpose when the two sides' next dims differ
constexpr int64_t VEC_MAX = 24;

// Tiles of a block: ntW · ntR · (slices along dB) · nbatch.
int64_t block_tiles(const TBlock &b, int dB, int lgB) {
    const int64_t ntB = dB >= 0 ? (b.size[dB] + ((int64_t)1 << lgB) - 1) >> lgB : 1;
    return b.ntW * b.ntR * ntB * b.nbatch;
}

// Fill tiling fields of a block given which trees are read/written.
void tile_block(TBlock &b, const std::vector<TTree> &trees, int L, bool allow_slices, int &dB, int &lgB) {
    const TTree &tw = trees[b.write_tree];
    const TTree &tr = trees[b.read_tree];
    const int area = (L == 1) ? class_area(1) : (L == 2 ? class_area(2) : (L == 4 ? class_area(4) : class_area(8)));
    const int twmax = (L == 1) ? class_twmax(1) : (L == 2 ? class_twmax(2) : (L == 4 ? class_twmax(4) : class_twmax(8)));
    int dW = 0, dR = 0;
    int64_t bestW = INT64_MAX, bestR = INT64_MAX;
    for (int d = 0; d < b.ndim; ++d) {
        if (b.size[d] <= 1) continue;
        int64_t sw = std::llabs(tw.str[d]), sr = std::llabs(tr.str[d]);
        if (sw < bestW) { bestW = sw; dW = d; }
        if (sr < bestR) { bestR = sr; dR = d; }
    }
    // the next dims after the shared innermost one, on the read and on the write side
    int aR = -1, aW = -1;
    if (dW == dR) {
        int64_t br = INT64_MAX, bw = INT64_MAX;
        for (int d = 0; d < b.ndim; ++d) {
            if (d == dW || b.size[d] <= 1) continue;
            const int64_t sr = std::llabs(tr.str[d]), sw = std::llabs(tw.str[d]);
            if (sr < br) { br = sr; aR = d; }
            if (sw < bw) { bw = sw; aW = d; }
        }
    }
    if (dW == dR && b.size[dW] <= VEC_MAX && aR >= 0 && aW >= 0 && aR != aW) {
        b.mode = 2;
        b.dV = dW;
        b.inv_w = 1.0 / (double)b.size[dW];
        dW = aW;
        dR = aR;
        choose_vec_tile(b.size[b.dV], b.size[dW], b.size[dR], area, twmax, b.tW, b.tR);
        b.lgW = log2i(b.tW);
        b.lgR = log2i(b.tR);
    } else if (dW == dR) {
        b.mode = 0;
        // second tiled dim: next smallest write stride among the others (or a size-1 dim)
        int second = -1;
        int64_t best2 = INT64_MAX;
        for (int d = 0; d < b.ndim; ++d) {
            if (d == dW) continue;
            int64_t s = std::llabs(tw.str[d]);
            if (b.size[d] > 1 && s < best2) { best2 = s; second = d; }
        }
        if (second < 0) second = (dW == 0) ? 1 : 0;
        dR = second;
        // tiles of `area` consecutive elements of the plane numbered along dW (the kernel
        // places them from the flat index): one tile column, ntR tiles per batch index
        b.tW = (int)std::min<int64_t>(b.size[dW], INT32_MAX);
        b.tR = 1;
        b.lgW = b.lgR = 0;
        b.inv_w = 1.0 / (double)b.size[dW];
    } else {
        b.mode = 1;
        choose_tile(b.size[dW], b.size[dR], area, twmax, b.tW, b.tR);
        b.lgW = log2i(b.tW);
        b.lgR = log2i(b.tR);
    }
    // with allow_slices, a transpose tile that leaves at least half its area unused
    // (small dims) takes slices along one more dim: the other dim with the smallest read
    // stride, tB of them (a power of two) with tW·tR·tB within the area and tW·tB within
    // twmax (the LDS staging bound, so the launch's LDS size does not grow)
    dB = -1;
    lgB = 0;
    if (allow_slices && b.mode == 1 && (int64_t)b.tW * b.tR * 2 <= area) {
        int64_t best = INT64_MAX;
        for (int d = 0; d < b.ndim; ++d) {
            if (d == dW || d == dR || b.size[d] <= 1) continue;
            const int64_t sr = std::llabs(tr.str[d]);
            if (sr < best) { best = sr; dB = d; }
        }
        if (dB >= 0) {
            int tb = 1;
            while ((int64_t)b.tW * b.tR * tb * 2 <= area && b.tW * tb * 2 <= twmax && tb < b.size[dB]) tb *= 2;
            if (tb > 1) {
                lgB = log2i(tb);
            } else {
                dB = -1;
            }
        }
    }
    b.dW = dW;
    b.dR = dR;
    if (b.mode == 0) {
        b.ntW = 1;
        b.ntR = (b.size[dW] * b.size[dR] + area - 1) / area;
    } else {
        b.ntW = (b.size[dW] + b.tW - 1) / b.tW;
        b.ntR = (b.size[dR] + b.tR - 1) / b.tR;
    }
    b.nother = 0;
    b.nbatch = 1;
    for (int d = 0; d < b.ndim; ++d) {
        if (d == dW || d == dR || d == dB || (b.mode == 2 && d == b.dV)) continue;
        b.odims[b.nother++] = d;
        b.nbatch *= b.size[d];
    }
}

// Whether every element of every tile of block b lies within 2^31 − 1 elements of its tile's
// corner in each of the trees [rt, rt + nr) and [wt, wt + nw): the largest in-tile position
// along each tiled dim times |stride| (copy tiles: up to `area` along dW and the rows a
// tile spans, with the row step of the kernel's copy numbering), summed.
bool tile_offsets_fit_32(const TBlock &b, const std::vector<TTree> &trees, int64_t rt, int nr, int64_t wt, int nw,
                         int area, int dB, int lgB) {
    const int64_t W = b.size[b.dW];
    int64_t mw, mr, mb = 0;
    int db = -1;
    if (b.mode == 0) {
        mw = area;
        mr = W >= area ? 1 : (area + W - 1) / W + 1;
    } else {
        mw = b.tW;
        mr = b.tR;
        if (b.mode == 2) {
            db = b.dV;
            mb = b.size[b.dV];
        } else if (dB >= 0) {
            db = dB;
            mb = (int64_t)1 << lgB;
        }
    }
    auto fits = [&](const TTree &t) {
        const int64_t sw = std::llabs(t.str[b.dW]);
        const int64_t sr = b.mode == 0 ? std::llabs(t.str[b.dR] - W * t.str[b.dW]) : std::llabs(t.str[b.dR]);
        const int64_t sb = db >= 0 ? std::llabs(t.str[db]) : 0;
        const double span = (double)mw * (double)sw + (double)mr * (double)sr + (double)mb * (double)sb;
        return span < 2147483647.0;
    };
    for (int i = 0; i < nr; ++i)
        if (!fits(trees[rt + i])) return false;
    for (int j = 0; j < nw; ++j)
        if (!fits(trees[wt + j])) return false;
    return true;
}

// Canonicalise one work item: `sizes` are the dst-order sizes; `strides[t][k]` the
// strides (in dst order) of every tree t that takes part.  Returns the kept dims.
std::vector<std::pair<int64_t, std::vector<int>>> canon_dims(
    const std::vector<int64_t> &sizes, const std::vector<const int64_t *> &strides) {
    // keep dims with size > 1
    std::vector<int> keep;
    for (size_t k = 0; k < sizes.size(); ++k)
        if (sizes[k] > 1) keep.push_back((int)k);
    // merge runs: dim k follows the group if for every tree str[k] == str[first]*size(group)
    std::vector<std::pair<int64_t, std::vector<int>>> groups;  // (size, member dims)
    for (int k : keep) {
        if (!groups.empty()) {
            auto &g = groups.back();
            bool ok = true;
            int first = g.second.front();
            for (auto *s : strides) {
                if (s[k] != s[first] * g.first) { ok = false; break; }
            }
            if (ok) {
                g.first *= sizes[k];
                g.second.push_back(k);
                continue;
            }
        }
        groups.push_back({sizes[k], {k}});
    }
    return groups;
}

int class_of(int nread, int nwrite) {
    const int m = std::max(std::min(nread, MAXW), nwrite);
    return m <= 1 ? 0 : (m <= 2 ? 1 : (m <= 4 ? 2 : 3));
}

int finalize_plan(tkmi_transform_plan_s *p, std::vector<HostItem> &items) {
    // forward and adjoint tile lists, one per class; destination trees are split
    // into chunks of at most MAXW per work item
    std::vector<TTree> &trees = p->trees;
    auto build = [&](bool adjoint, PlanSide *sides) {
        for (int c = 0; c < NSIDE; ++c) {
            sides[c].blocks.clear();
            sides[c].ntiles = 0;
            sides[c].staged = false;
            sides[c].reads = false;
            sides[c].sliced = false;
            sides[c].slices.clear();
            sides[c].written = 0;
        }
        // A small plan runs as one launch of its largest class when that class holds at
        // most 4 trees: the ramp-up and tail of a second launch cost more than the
        // smaller per-thread work of one-tree items in the bigger class.
        int cmax = 0;
        int64_t moved = 0;
        for (auto &it : items) {
            const int R = adjoint ? it.nread : it.nwrite;
            const int Cn = adjoint ? it.nwrite : it.nread;
            int64_t vol = 1;
            for (int64_t d : it.size) vol *= d;
            moved += vol * R;
            for (int jc = 0; jc < R; jc += MAXW) cmax = std::max(cmax, class_of(Cn, std::min(MAXW, R - jc)));
        }
        const bool merge = moved <= SMALL_PLAN && class_L(cmax) <= 4;
        for (int c = 0; c < NSIDE; ++c) sides[c].written = moved;
        // Slicing is decided per class: its launch then uses the kernel instantiation with
        // slices, which holds more registers (the L = 1 kernel drops from 7 to 6 waves per
        // SIMD), so a class takes it only when slices cut its tile count by at least 60%
        // (measured: 2.65× on an SU(2) rank-4 permute with 8–24 element dims, whose tiles
        // drop to 0.24 of their number; a class at 0.6 with 16–48 element dims and the
        // 82.8 GB permute with 32–96 element dims ran slower in the sliced kernel).
        bool slice_class[NCLASS];
        {
            int64_t plain[NCLASS] = {0}, sliced[NCLASS] = {0};
            for (auto &it : items) {
                const int R = adjoint ? it.nread : it.nwrite;
                const int Cn = adjoint ? it.nwrite : it.nread;
                const int64_t rt = adjoint ? it.write_tree : it.read_tree;
                const int64_t wt = adjoint ? it.read_tree : it.write_tree;
                for (int jc = 0; jc < R; jc += MAXW) {
                    TBlock b;
                    std::memset(&b, 0, sizeof(b));
                    b.ndim = (int)it.size.size();
                    for (int d = 0; d < b.ndim; ++d) b.size[d] = it.size[d];
                    b.nread = Cn;
                    b.nwrite = std::min(MAXW, R - jc);
                    b.read_tree = rt;
                    b.write_tree = wt + jc;
                    const int c = merge ? cmax : class_of(b.nread, b.nwrite);
                    int sdb, slg;
                    tile_block(b, trees, class_L(c), false, sdb, slg);
                    plain[c] += block_tiles(b, sdb, slg);
                    tile_block(b, trees, class_L(c), true, sdb, slg);
                    sliced[c] += block_tiles(b, sdb, slg);
                }
            }
            for (int c = 0; c < NCLASS; ++c) slice_class[c] = sliced[c] * 10 <= plain[c] * 4;
        }
        for (auto &it : items) {
            const int R = adjoint ? it.nread : it.nwrite;   // written trees on this side
            const int Cn = adjoint ? it.nwrite : it.nread;  // read trees on this side
            const int64_t rt = adjoint ? it.write_tree : it.read_tree;
            const int64_t wt = adjoint ? it.read_tree : it.write_tree;
            for (int jc = 0; jc < R; jc += MAXW) {
                TBlock b;
                std::memset(&b, 0, sizeof(b));
                b.ndim = (int)it.size.size();
                for (int d = 0; d < b.ndim; ++d) b.size[d] = it.size[d];
                b.nread = Cn;
                b.nwrite = std::min(MAXW, R - jc);
                b.read_tree = rt;
                b.write_tree = wt + jc;
                b.u_start = it.u_start + jc;
                b.u_ld = R;
                const int c = merge ? cmax : class_of(b.nread, b.nwrite);
                int sdb, slg;
                tile_block(b, trees, class_L(c), slice_class[c], sdb, slg);
                b.off32 = tile_offsets_fit_32(b, trees, rt, Cn, b.write_tree, b.nwrite, class_area(class_L(c)),
                                              sdb, slg) ? 1 : 0;
                // vector-transpose tiles run in a launch of their own kernel, so that the
                // other kernels keep their register counts
                PlanSide &sd = sides[c + (b.mode == 2 ? NCLASS : 0)];
                b.tile_start = sd.ntiles;
                sd.ntiles += block_tiles(b, sdb, slg);
                sd.slices.push_back(sdb >= 0 ? (int32_t)(sdb | (slg << 8)) : -1);
                sd.staged |= (b.mode != 0);
                sd.reads |= (b.nread > 0);
                sd.sliced |= (sdb >= 0);
                sd.blocks.push_back(b);
            }
        }
    };
    build(false, p->fwd);
    build(true, p->adj);
    p->nitems = (int64_t)items.size();

    int rc;
    if ((rc = tkmi::upload((void **)&p->d_trees, trees.data(), trees.size() * sizeof(TTree)))) return rc;
    if ((rc = tkmi::upload((void **)&p->d_U, p->U.data(), p->U.size() * sizeof(double)))) return rc;
    if ((rc = tkmi::upload((void **)&p->d_Ut, p->Ut.data(), p->Ut.size() * sizeof(double)))) return rc;
    for (int c = 0; c < NSIDE; ++c) {
        for (PlanSide *side : {&p->fwd[c], &p->adj[c]}) {
            if (side->blocks.empty()) continue;
            // the blocks, then their tile_start values as a compact array: the kernels'
            // binary search reads 8 bytes per probe from a few cache lines (the scalar
            // cache keeps them) instead of one probe per ~250-byte block record
            const size_t nb = side->blocks.size();
            std::vector<unsigned char> buf(nb * (sizeof(TBlock) + sizeof(int64_t)));
            std::memcpy(buf.data(), side->blocks.data(), nb * sizeof(TBlock));
            for (size_t b = 0; b < nb; ++b)
                std::memcpy(buf.data() + nb * sizeof(TBlock) + b * sizeof(int64_t), &side->blocks[b].tile_start,
                            sizeof(int64_t));
            if ((rc = tkmi::upload((void **)&side->d_blocks, buf.data(), buf.size()))) return rc;
            if (side->sliced && (rc = tkmi::upload((void **)&side->d_slices, side->slices.data(),
                                                   side->slices.size() * sizeof(int32_t))))
                return rc;
        }
    }
    return TKMI_OK;
}

void free_plan(tkmi_transform_plan_s *p) {
    if (!p) return;
    (void)hipFree(p->d_gcoef);
    (void)hipFree(p->d_gidx);
    (void)hipFree(p->d_trees);
    (void)hipFree(p->d_U);
    (void)hipFree(p->d_Ut);
    for (int c = 0; c < NSIDE; ++c) {
        (void)hipFree(p->fwd[c].d_blocks);
        (void)hipFree(p->adj[c].d_blocks);
        (void)hipFree(p->fwd[c].d_slices);
        (void)hipFree(p->adj[c].d_slices);
    }
    delete p;
}

int check_perm(int32_t N, const int32_t *perm) {
    if (N < 0 || N > MAXD) TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "transform: number of indices exceeds 12");
    std::vector<int> seen(N, 0);
    for (int k = 0; k < N; ++k) {
        if (perm[k] < 0 || perm[k] >= N || seen[perm[k]])
            TKMI_FAIL(TKMI_ERR_INDEX, "transform: perm is not a permutation of 0..N-1");
        seen[perm[k]] = 1;
    }
    return TKMI_OK;
}

// Add one work item given dst-order sizes and per-tree (dst-order) strides.
// read trees come first in `strs` (nread of them), then write trees.
int add_item(tkmi_transform_plan_s *p, std::vector<HostItem> &items, const std::vector<int64_t> &sizes,
             const std::vector<std::vector<int64_t>> &strs, const std::vector<int64_t> &offs, int nread,
             int nwrite, int64_t u_start) {
    std::vector<const int64_t *> sp;
    for (auto &s : strs) sp.push_back(s.data());
    auto groups = canon_dims(sizes, sp);
    HostItem it;
    for (auto &g : groups) it.size.push_back(g.first);
    while (it.size.size() < 2) it.size.push_back(1);
    it.nread = nread;
    it.nwrite = nwrite;
    it.u_start = u_start;
    it.read_tree = (int64_t)p->trees.size();
    it.write_tree = it.read_tree + nread;
    for (size_t t = 0; t < strs.size(); ++t) {
        TTree tr;
        std::memset(&tr, 0, sizeof(tr));
        tr.off = offs[t];
        for (size_t g = 0; g < groups.size(); ++g) tr.str[g] = strs[t][groups[g].second.front()];
        p->trees.push_back(tr);
    }
    int64_t n = 1;
    for (auto s : sizes) n *= s;
    p->elements += n * nwrite;
    items.push_back(std::move(it));
    return TKMI_OK;
}

template <typename T, int L>
void launch_class(const tkmi_transform_plan_s *p, const PlanSide &side, const double *dU, const T *src, T *dst,
                  const Scal &sc, hipStream_t stream) {
    const int keep = side.written * (int64_t)sizeof(T) <= KEEP_BYTES;
    // the dispatch packet holds the grid size in work-items as 32 bits: keep
    // blocks × NTHREADS below 2^31 per launch and continue at tile0 = t0
    const int64_t chunk = ((int64_t)1 << 31) / NTHREADS;
    for (int64_t t0 = 0; t0 < side.ntiles; t0 += chunk) {
        const unsigned grid = (unsigned)std::min(chunk, side.ntiles - t0);
        if (side.staged && side.sliced)
            hipLaunchKernelGGL((transform_kernel<T, L, true, true>), dim3(grid), dim3(NTHREADS), 0, stream,
                               side.d_blocks, (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, sc,
                               keep, t0, (const int32_t *)side.d_slices);
        else if (side.staged)
            hipLaunchKernelGGL((transform_kernel<T, L, true, false>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                               (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, sc, keep, t0,
                               (const int32_t *)nullptr);
        else
            hipLaunchKernelGGL((transform_kernel<T, L, false, false>), dim3(grid), dim3(NTHREADS), 0, stream,
                               side.d_blocks, (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, sc,
                               keep, t0, (const int32_t *)nullptr);
    }
}

template <typename T, int L>
void launch_vec(const tkmi_transform_plan_s *p, const PlanSide &side, const double *dU, const T *src, T *dst,
                const Scal &sc, hipStream_t stream) {
    const int keep = side.written * (int64_t)sizeof(T) <= KEEP_BYTES;
    const int64_t chunk = ((int64_t)1 << 31) / NTHREADS;
    for (int64_t t0 = 0; t0 < side.ntiles; t0 += chunk) {
        const unsigned grid = (unsigned)std::min(chunk, side.ntiles - t0);
        hipLaunchKernelGGL((transform_vec_kernel<T, L>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                           (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, sc, keep, t0);
    }
}

template <typename T>
int launch_side(const tkmi_transform_plan_s *p, const PlanSide *sides, const double *dU, const T *src, T *dst,
                const Scal &sc, hipStream_t stream) {
    if (sides[0].ntiles) launch_class<T, 1>(p, sides[0], dU, src, dst, sc, stream);
    if (sides[1].ntiles) launch_class<T, 2>(p, sides[1], dU, src, dst, sc, stream);
    if (sides[2].ntiles) launch_class<T, 4>(p, sides[2], dU, src, dst, sc, stream);
    if (sides[3].ntiles) launch_class<T, 8>(p, sides[3], dU, src, dst, sc, stream);
    if (sides[NCLASS + 0].ntiles) launch_vec<T, 1>(p, sides[NCLASS + 0], dU, src, dst, sc, stream);
    if (sides[NCLASS + 1].ntiles) launch_vec<T, 2>(p, sides[NCLASS + 1], dU, src, dst, sc, stream);
    if (sides[NCLASS + 2].ntiles) launch_vec<T, 4>(p, sides[NCLASS + 2], dU, src, dst, sc, stream);
    if (sides[NCLASS + 3].ntiles) launch_vec<T, 8>(p, sides[NCLASS + 3], dU, src, dst, sc, stream);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

int64_t total_tiles(const PlanSide *sides) {
    int64_t n = 0;
    for (int c = 0; c < NSIDE; ++c) n += sides[c].ntiles;
    return n;
}

size_t elem_size(int32_t dtype) {
    switch (dtype) {
        case TKMI_F64: return 8;
        case TKMI_F32: return 4;
        case TKMI_C128: return 16;
        case TKMI_C64: return 8;
        default: return 0;
    }
}

int exec_common(tkmi_transform_plan p, int32_t dtype, const void *src, void *dst, const Scal &sc, void *stream,
                bool adjoint, const void *gsrc = nullptr) {
    if (!p) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    const double alpha = sc.ar;
    const bool has_alpha = sc.ar != 0.0 || sc.ai != 0.0;
    const PlanSide *sides = adjoint ? p->adj : p->fwd;
    const double *dU = adjoint ? p->d_Ut : p->d_U;
    // a source may be empty (null) when no block reads it: e.g. a contraction with an
    // empty operand into a non-empty C only scales C by β
    bool reads = false;
    for (int c = 0; c < NSIDE; ++c) reads |= sides[c].ntiles > 0 && sides[c].reads;
    (void)alpha;
    if ((total_tiles(sides) > 0 && !dst) || (reads && has_alpha && !src))
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null data pointer");
    const size_t es0 = elem_size(dtype);
    if (!es0) TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "transform: dtype must be TKMI_F64, TKMI_F32, TKMI_C128 or TKMI_C64");
    if ((dtype == TKMI_F64 || dtype == TKMI_F32) && (sc.ai != 0.0 || sc.bi != 0.0))
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: complex scalars need complex data");
    if (p->gather_len > 0 && dtype != TKMI_F64 && dtype != TKMI_F32)
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "transform: gathered coefficients need real data");
    // the source and destination vectors must not overlap (tiles read and write them in
    // no particular order), except an elementwise plan run in place, or α = 0 (no reads)
    if (reads && has_alpha && src && dst) {
        const uintptr_t es = es0;
        const uintptr_t s0 = (uintptr_t)src, d0 = (uintptr_t)dst;
        const uintptr_t s1 = s0 + es * (uintptr_t)(adjoint ? p->dst_len : p->src_len);
        const uintptr_t d1 = d0 + es * (uintptr_t)(adjoint ? p->src_len : p->dst_len);
        if (s0 < d1 && d0 < s1 && !(p->inplace && s0 == d0))
            TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: source and destination overlap");
    }
    // a gather plan's coefficients are formed from gsrc at every call
    if (p->gather_len > 0 && (adjoint || !gsrc))
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: the plan's coefficients come from a gather source: "
                                     "use tkmi_transform_exec_gather (no adjoint)");
    hipStream_t s = tkmi::as_stream(stream);
    if (p->gather_len > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>((p->gather_len + NTHREADS - 1) / NTHREADS, 1024);
        if (dtype == TKMI_F64)
            hipLaunchKernelGGL(gather_coeffs<double>, dim3(grid), dim3(NTHREADS), 0, s, p->d_gcoef, p->d_gidx,
                               (const double *)gsrc, p->gather_len, p->d_U);
        else
            hipLaunchKernelGGL(gather_coeffs<float>, dim3(grid), dim3(NTHREADS), 0, s, p->d_gcoef, p->d_gidx,
                               (const float *)gsrc, p->gather_len, p->d_U);
    }
    switch (dtype) {
        case TKMI_F64: return launch_side<double>(p, sides, dU, (const double *)src, (double *)dst, sc, s);
        case TKMI_F32: return launch_side<float>(p, sides, dU, (const float *)src, (float *)dst, sc, s);
        case TKMI_C128: return launch_side<d2_t>(p, sides, dU, (const d2_t *)src, (d2_t *)dst, sc, s);
        default: return launch_side<f2_t>(p, sides, dU, (const f2_t *)src, (f2_t *)dst, sc, s);
    }
}

bool in_range(const int64_t *sz, const int64_t *st, int64_t off, int N, int64_t len) {
    // every element of the strided view lies in [0, len)
    int64_t lo = off, hi = off;
    for (int k = 0; k < N; ++k) {
        if (sz[k] <= 0) return false;
        int64_t span = (sz[k] - 1) * st[k];
        if (span < 0) lo += span; else hi += span;
    }
    return lo >= 0 && hi < len;
}

}  // namespace

extern "C" {

int tkmi_abelian_plan_create(tkmi_transform_plan *plan, int32_t numind, const int32_t *perm, int64_t nentries,
                             const double *coeffs, const int64_t *dst_structs, const int64_t *src_structs,
                             int64_t dst_len, int64_t src_len) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_abelian_plan_create: null plan pointer");
    *plan = nullptr;
    int rc = check_perm(numind, perm);
    if (rc) return rc;
    if (nentries < 0 || (nentries > 0 && (!coeffs || !dst_structs || !src_structs)))
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_abelian_plan_create: bad table");
    const int N = numind;
    auto *p = new tkmi_transform_plan_s();
    p->N = N;
    p->dst_len = dst_len;
    p->src_len = src_len;
    p->inplace = dst_len == src_len;
    for (int k = 0; k < N; ++k) p->inplace &= perm[k] == k;
    std::vector<HostItem> items;
    for (int64_t e = 0; e < nentries; ++e) {
        const int64_t *ds = dst_structs + e * (2 * N + 1);
        const int64_t *ss = src_structs + e * (2 * N + 1);
        p->inplace &= std::equal(ds, ds + 2 * N + 1, ss);
        std::vector<int64_t> sizes(N);
        std::vector<std::vector<int64_t>> strs(2, std::vector<int64_t>(N));
        for (int k = 0; k < N; ++k) {
            if (ds[k] != ss[perm[k]]) {
                free_plan(p);
                TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_abelian_plan_create: dst sizes are not the permuted src sizes");
            }
            sizes[k] = ds[k];
            strs[0][k] = ss[N + perm[k]];  // read (src) strides in dst order
            strs[1][k] = ds[N + k];        // write (dst) strides
        }
        if (!in_range(ds, ds + N, ds[2 * N], N, dst_len) || !in_range(ss, ss + N, ss[2 * N], N, src_len)) {
            free_plan(p);
            TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_abelian_plan_create: sub-block outside the data vector");
        }
        p->U.push_back(coeffs[e]);
        p->Ut.push_back(coeffs[e]);
        add_item(p, items, sizes, strs, {ss[2 * N], ds[2 * N]}, 1, 1, (int64_t)p->U.size() - 1);
    }
    rc = finalize_plan(p, items);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *plan = p;
    return TKMI_OK;
}

int tkmi_generic_plan_create(tkmi_transform_plan *plan, int32_t numind, const int32_t *perm, int64_t nblocks,
                             const int64_t *rows, const int64_t *cols, const double *U, const int64_t *sz_dst,
                             const int64_t *sz_src, const int64_t *so_dst, const int64_t *so_src,
                             const double *w_dst, const double *w_src, int64_t dst_len, int64_t src_len) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_generic_plan_create: null plan pointer");
    *plan = nullptr;
    int rc = check_perm(numind, perm);
    if (rc) return rc;
    const int N = numind;
    if (nblocks < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_generic_plan_create: negative block count");
    // cols = 0 is allowed: a destination with no source contributions is only scaled by β
    // (a partial trace whose traced sectors never match for that destination sub-block)
    for (int64_t b = 0; b < nblocks; ++b)
        if (rows[b] < 1 || cols[b] < 0 || rows[b] > INT32_MAX || cols[b] > INT32_MAX)
            TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_generic_plan_create: bad recoupling matrix size");
    auto *p = new tkmi_transform_plan_s();
    p->N = N;
    p->dst_len = dst_len;
    p->src_len = src_len;
    std::vector<HostItem> items;
    int64_t uoff = 0, tdst = 0, tsrc = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t R = rows[b], C = cols[b];
        const int64_t *szd = sz_dst + b * N, *szs = sz_src + b * N;
        std::vector<int64_t> sizes(N);
        for (int k = 0; k < N; ++k) {
            if (szd[k] != szs[perm[k]]) {
                free_plan(p);
                TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_generic_plan_create: dst sizes are not the permuted src sizes");
            }
            sizes[k] = szd[k];
        }
        std::vector<std::vector<int64_t>> strs;
        std::vector<int64_t> offs;
        for (int64_t i = 0; i < C; ++i) {
            const int64_t *so = so_src + (tsrc + i) * (N + 1);
            if (!in_range(szs, so, so[N], N, src_len)) {
                free_plan(p);
                TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_generic_plan_create: source sub-block outside the data vector");
            }
            std::vector<int64_t> s(N);
            for (int k = 0; k < N; ++k) s[k] = so[perm[k]];
            strs.push_back(s);
            offs.push_back(so[N]);
        }
        for (int64_t j = 0; j < R; ++j) {
            const int64_t *so = so_dst + (tdst + j) * (N + 1);
            if (!in_range(szd, so, so[N], N, dst_len)) {
                free_plan(p);
                TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_generic_plan_create: destination sub-block outside the data vector");
            }
            strs.push_back(std::vector<int64_t>(so, so + N));
            offs.push_back(so[N]);
        }
        // U is R×C column-major; adjoint matrix is its transpose (C×R column-major)
        const int64_t ustart = (int64_t)p->U.size();
        for (int64_t i = 0; i < R * C; ++i) p->U.push_back(U[uoff + i]);
        p->Ut.resize(p->U.size());
        // adjoint w.r.t. the dim(c)-weighted inner product: Ut = W_src⁻¹ Uᵀ W_dst
        for (int64_t i = 0; i < C; ++i)
            for (int64_t j = 0; j < R; ++j) {
                const double wd = w_dst ? w_dst[tdst + j] : 1.0, ws = w_src ? w_src[tsrc + i] : 1.0;
                p->Ut[ustart + i + j * C] = U[uoff + j + i * R] * wd / ws;
            }
        add_item(p, items, sizes, strs, offs, (int)C, (int)R, ustart);
        uoff += R * C;
        tsrc += C;
        tdst += R;
    }
    rc = finalize_plan(p, items);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *plan = p;
    return TKMI_OK;
}

int tkmi_transform_exec(tkmi_transform_plan plan, int32_t dtype, const void *src, void *dst, double alpha,
                        double beta, void *stream) {
    return exec_common(plan, dtype, src, dst, Scal{alpha, 0.0, beta, 0.0, 0}, stream, false);
}

int tkmi_transform_exec_ex(tkmi_transform_plan plan, int32_t dtype, int32_t conj_src, const void *src, void *dst,
                           double alpha_re, double alpha_im, double beta_re, double beta_im, void *stream) {
    const int conj = (conj_src != 0) && (dtype == TKMI_C128 || dtype == TKMI_C64);
    return exec_common(plan, dtype, src, dst, Scal{alpha_re, alpha_im, beta_re, beta_im, conj}, stream, false);
}

int tkmi_transform_adjoint_exec(tkmi_transform_plan plan, int32_t dtype, const void *ddst, void *dsrc,
                                double alpha, double beta, void *stream) {
    return exec_common(plan, dtype, ddst, dsrc, Scal{alpha, 0.0, beta, 0.0, 0}, stream, true);
}

int tkmi_transform_plan_set_gather(tkmi_transform_plan plan, int64_t n, const double *coef, const int64_t *idx) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    if (n != (int64_t)plan->U.size())
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_transform_plan_set_gather: one entry per recoupling coefficient");
    if (plan->d_gcoef) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_transform_plan_set_gather: gather already set");
    int rc;
    if ((rc = tkmi::upload((void **)&plan->d_gcoef, coef, (size_t)n * sizeof(double)))) return rc;
    if ((rc = tkmi::upload((void **)&plan->d_gidx, idx, (size_t)n * sizeof(int64_t)))) return rc;
    plan->gather_len = n;
    return TKMI_OK;
}

int tkmi_transform_exec_gather(tkmi_transform_plan plan, int32_t dtype, const void *src, void *dst,
                               const void *gsrc, double alpha, double beta, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    if (!plan->d_gcoef) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_transform_exec_gather: plan has no gather");
    if (plan->gather_len > 0 && !gsrc) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_transform_exec_gather: null gather source");
    return exec_common(plan, dtype, src, dst, Scal{alpha, 0.0, beta, 0.0, 0}, stream, false, gsrc);
}

int tkmi_transform_plan_info(tkmi_transform_plan plan, int64_t *ntiles, int64_t *nentries, int64_t *elements) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    if (ntiles) *ntiles = total_tiles(plan->fwd);
    if (nentries) *nentries = plan->nitems;
    if (elements) *elements = plan->elements;
    return TKMI_OK;
}

int tkmi_transform_plan_destroy(tkmi_transform_plan plan) {
    free_plan(plan);
    return TKMI_OK;
}

}  // extern "C"
