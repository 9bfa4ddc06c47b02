// Grouped permute / braid of block-sparse tensors (K1 abelian, K2 generic recoupling).
//
// One launch covers every sub-block (abelian) or every fusion block (generic) of a
// transform.  Host-side plan creation canonicalises each work item:
//   * dims of size 1 are dropped, dims that stay contiguous on both sides are merged;
//   * dW = the dim with the smallest stride on the written side,
//     dR = the dim with the smallest stride on the read side;
//   * dW == dR  -> "copy" mode: lanes run along dW for both the read and the write;
//     dW != dR  -> "transpose" mode: a (tW × tR) tile is read with lanes along dR,
//                  staged in LDS, and written with lanes along dW.
// Tiles of all work items are numbered consecutively (tile_start prefix sums); each
// workgroup owns one tile and finds its work item by binary search, so the whole
// transform is a single launch with a device-side block schedule.
//
// Generic (non-abelian) blocks carry L_read source trees and L_write destination
// trees of identical shape plus a dense L_write × L_read recoupling matrix U; each
// written element is  β·dst_j + Σ_i (α·U[j,i])·src_i, i.e. the reference's
// pack → buffer_src·Uᵀ → unpack (indexmanipulations.jl:692-721) fused in registers.
#include "tkmi_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr int MAXD = 12;
constexpr int NTHREADS = 256;

struct TBlock {
    int32_t ndim;
    int32_t mode;  // 0 copy, 1 transpose
    int32_t nread, nwrite;
    int32_t dW, dR;
    int32_t tW, tR;
    int32_t nother;
    int32_t odims[MAXD];
    int64_t size[MAXD];
    int64_t ntW, ntR, nbatch;
    int64_t tile_start;
    int64_t read_tree, write_tree;
    int64_t u_start;
};

struct TTree {
    int64_t off;
    int64_t str[MAXD];
};

template <int L>
struct TileCfg;
template <>
struct TileCfg<1> {
    static constexpr int CAP = 64;
};
template <>
struct TileCfg<2> {
    static constexpr int CAP = 32;
};
template <>
struct TileCfg<4> {
    static constexpr int CAP = 32;
};
template <>
struct TileCfg<8> {
    static constexpr int CAP = 16;
};

template <typename T, int L>
__global__ __launch_bounds__(NTHREADS) void transform_kernel(
    const TBlock *__restrict__ blocks, int64_t nblocks, const TTree *__restrict__ trees,
    const double *__restrict__ Umat, const T *__restrict__ src, T *__restrict__ dst, double alpha,
    double beta, int64_t tile0) {
    constexpr int CAP = TileCfg<L>::CAP;
    __shared__ T lds[L * CAP * (CAP + 1)];

    const int64_t gt = tile0 + (int64_t)blockIdx.x;
    // binary search: last block with tile_start <= gt
    int64_t lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (blocks[mid].tile_start <= gt)
            lo = mid;
        else
            hi = mid - 1;
    }
    const TBlock &B = blocks[lo];
    const int nread = B.nread, nwrite = B.nwrite;
    const int dW = B.dW, dR = B.dR;
    const int tW = B.tW, tR = B.tR;

    int64_t t = gt - B.tile_start;
    const int64_t iW = t % B.ntW;
    t /= B.ntW;
    const int64_t iR = t % B.ntR;
    int64_t bt = t / B.ntR;

    int64_t rbase[L], wbase[L];
    int64_t rsW[L], rsR[L], wsW[L], wsR[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        if (i < nread) {
            const TTree &tr = trees[B.read_tree + i];
            rbase[i] = tr.off;
            rsW[i] = tr.str[dW];
            rsR[i] = tr.str[dR];
        }
        if (i < nwrite) {
            const TTree &tw = trees[B.write_tree + i];
            wbase[i] = tw.off;
            wsW[i] = tw.str[dW];
            wsR[i] = tw.str[dR];
        }
    }
    for (int k = 0; k < B.nother; ++k) {
        const int d = B.odims[k];
        const int64_t sz = B.size[d];
        const int64_t idx = bt % sz;
        bt /= sz;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            if (i < nread) rbase[i] += idx * trees[B.read_tree + i].str[d];
            if (i < nwrite) wbase[i] += idx * trees[B.write_tree + i].str[d];
        }
    }

    T u[L][L];
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
        for (int i = 0; i < L; ++i)
            u[j][i] = (i < nread && j < nwrite) ? (T)(alpha * Umat[B.u_start + j + (int64_t)i * nwrite])
                                                : (T)0;
    const T tbeta = (T)beta;
    const bool has_beta = (beta != 0.0);
    const bool has_alpha = (alpha != 0.0);  // α = 0: the source is not read (BLAS convention)

    const int64_t w0 = iW * tW, r0 = iR * tR;
    const int nw = (int)min((int64_t)tW, B.size[dW] - w0);
    const int nr = (int)min((int64_t)tR, B.size[dR] - r0);
    const int tid = threadIdx.x;

    if (B.mode == 0) {
        // copy mode: the tile is `nr` runs of `nw` elements along dW (contiguous on both
        // sides); its elements are numbered run after run and dealt to the threads in
        // order, so consecutive lanes touch consecutive addresses even for short runs.
        const int ne = nw * nr;
        const float inv = 1.0f / (float)nw;
        for (int e = tid; e < ne; e += NTHREADS) {
            int r = (int)((float)e * inv);
            int w = e - r * nw;
            if (w < 0) { --r; w += nw; }
            else if (w >= nw) { ++r; w -= nw; }
            const int64_t ww = w0 + w, rr = r0 + r;
            T v[L];
#pragma unroll
            for (int i = 0; i < L; ++i)
                v[i] = (i < nread && has_alpha) ? src[rbase[i] + ww * rsW[i] + rr * rsR[i]] : (T)0;
#pragma unroll
            for (int j = 0; j < L; ++j) {
                if (j < nwrite) {
                    T o = 0;
#pragma unroll
                    for (int i = 0; i < L; ++i)
                        if (i < nread) o += u[j][i] * v[i];
                    const int64_t a = wbase[j] + ww * wsW[j] + rr * wsR[j];
                    dst[a] = has_beta ? o + tbeta * dst[a] : o;
                }
            }
        }
    } else {
        // transpose mode: read with lanes along dR into LDS, write with lanes along dW
        {
            const int tx = tid % tR, ty = tid / tR, wstep = NTHREADS / tR;
            if (tx < nr) {
                const int64_t rr = r0 + tx;
                for (int w = ty; w < nw; w += wstep) {
                    const int64_t ww = w0 + w;
#pragma unroll
                    for (int i = 0; i < L; ++i)
                        if (i < nread)
                            lds[(i * CAP + w) * (CAP + 1) + tx] =
                                has_alpha ? src[rbase[i] + ww * rsW[i] + rr * rsR[i]] : (T)0;
                }
            }
        }
        __syncthreads();
        {
            const int tx = tid % tW, ty = tid / tW, rstep = NTHREADS / tW;
            if (tx < nw) {
                const int64_t ww = w0 + tx;
                for (int r = ty; r < nr; r += rstep) {
                    const int64_t rr = r0 + r;
#pragma unroll
                    for (int j = 0; j < L; ++j) {
                        if (j < nwrite) {
                            T o = 0;
#pragma unroll
                            for (int i = 0; i < L; ++i)
                                if (i < nread) o += u[j][i] * lds[(i * CAP + tx) * (CAP + 1) + r];
                            const int64_t a = wbase[j] + ww * wsW[j] + rr * wsR[j];
                            dst[a] = has_beta ? o + tbeta * dst[a] : o;
                        }
                    }
                }
            }
        }
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

int pow2ceil(int64_t x) {
    int p = 1;
    while (p < x && p < (1 << 20)) p <<= 1;
    return p;
}

struct PlanSide {
    std::vector<TBlock> blocks;
    int64_t ntiles = 0;
    TBlock *d_blocks = nullptr;
};

}  // namespace

struct tkmi_transform_plan_s {
    int32_t N = 0;
    int L = 1;  // template width (max trees per item, rounded to 1/2/4/8)
    int64_t dst_len = 0, src_len = 0, elements = 0, nitems = 0;
    std::vector<TTree> trees;
    std::vector<double> U, Ut;
    PlanSide fwd, adj;
    TTree *d_trees = nullptr;
    double *d_U = nullptr, *d_Ut = nullptr;
};

namespace {

// Fill tiling fields of a block given which trees are read/written.
void tile_block(TBlock &b, const std::vector<TTree> &trees, int L) {
    const TTree &tw = trees[b.write_tree];
    const TTree &tr = trees[b.read_tree];
    int dW = 0, dR = 0;
    int64_t bestW = INT64_MAX, bestR = INT64_MAX;
    for (int d = 0; d < b.ndim; ++d) {
        if (b.size[d] <= 1) continue;
        int64_t sw = std::llabs(tw.str[d]), sr = std::llabs(tr.str[d]);
        if (sw < bestW) { bestW = sw; dW = d; }
        if (sr < bestR) { bestR = sr; dR = d; }
    }
    if (dW == dR) {
        b.mode = 0;
        // second tiled dim: next smallest write stride among the others (or a padded dim)
        int second = -1;
        int64_t best2 = INT64_MAX;
        for (int d = 0; d < b.ndim; ++d) {
            if (d == dW) continue;
            int64_t s = std::llabs(tw.str[d]);
            if (b.size[d] > 1 && s < best2) { best2 = s; second = d; }
        }
        if (second < 0) second = (dW == 0) ? 1 : 0;
        dR = second;
        // tiles of whole runs: up to COPY_TILE elements; runs longer than that are split
        constexpr int64_t COPY_TILE = 4096;
        if (b.size[dW] >= COPY_TILE / 2) {
            b.tW = (int)COPY_TILE;
            b.tR = 1;
        } else {
            b.tW = (int)b.size[dW];
            b.tR = (int)std::max<int64_t>(1, std::min<int64_t>(b.size[dR], COPY_TILE / b.size[dW]));
        }
    } else {
        b.mode = 1;
        const int cap = (L == 1) ? 64 : (L <= 4 ? 32 : 16);
        b.tW = std::min(cap, pow2ceil(b.size[dW]));
        b.tR = std::min(cap, pow2ceil(b.size[dR]));
    }
    b.dW = dW;
    b.dR = dR;
    b.ntW = (b.size[dW] + b.tW - 1) / b.tW;
    b.ntR = (b.size[dR] + b.tR - 1) / b.tR;
    b.nother = 0;
    b.nbatch = 1;
    for (int d = 0; d < b.ndim; ++d) {
        if (d == dW || d == dR) continue;
        b.odims[b.nother++] = d;
        b.nbatch *= b.size[d];
    }
}

// Canonicalise one work item: `sizes` are the dst-order sizes; `wstr[t][k]` the
// strides (in dst order) of every tree t that takes part.  Returns the kept dims.
std::vector<std::pair<int64_t, std::vector<int>>> canon_dims(
    const std::vector<int64_t> &sizes, const std::vector<const int64_t *> &strides) {
    // keep dims with size > 1
    std::vector<int> keep;
    for (size_t k = 0; k < sizes.size(); ++k)
        if (sizes[k] > 1) keep.push_back((int)k);
    // merge runs: dim b follows a if for every tree str[b] == str[a]*size[a]
    std::vector<std::pair<int64_t, std::vector<int>>> groups;  // (size, member dims)
    for (int k : keep) {
        if (!groups.empty()) {
            auto &g = groups.back();
            int a = g.second.back();
            bool ok = true;
            int first = g.second.front();
            for (auto *s : strides) {
                // merged group's stride is str[first]; the size so far is g.first
                if (s[k] != s[first] * g.first) { ok = false; break; }
            }
            (void)a;
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

int finalize_plan(tkmi_transform_plan_s *p, std::vector<HostItem> &items) {
    // build forward and adjoint tile lists
    std::vector<TTree> &trees = p->trees;
    auto build = [&](bool adjoint, PlanSide &side) {
        side.blocks.clear();
        int64_t tiles = 0;
        for (auto &it : items) {
            TBlock b;
            std::memset(&b, 0, sizeof(b));
            b.ndim = (int)it.size.size();
            for (int d = 0; d < b.ndim; ++d) b.size[d] = it.size[d];
            if (!adjoint) {
                b.nread = it.nread;
                b.nwrite = it.nwrite;
                b.read_tree = it.read_tree;
                b.write_tree = it.write_tree;
            } else {
                b.nread = it.nwrite;
                b.nwrite = it.nread;
                b.read_tree = it.write_tree;
                b.write_tree = it.read_tree;
            }
            b.u_start = it.u_start;
            tile_block(b, trees, p->L);
            b.tile_start = tiles;
            tiles += b.ntW * b.ntR * b.nbatch;
            side.blocks.push_back(b);
        }
        side.ntiles = tiles;
    };
    build(false, p->fwd);
    build(true, p->adj);
    p->nitems = (int64_t)items.size();

    int rc;
    if ((rc = tkmi::upload((void **)&p->d_trees, trees.data(), trees.size() * sizeof(TTree)))) return rc;
    if ((rc = tkmi::upload((void **)&p->d_U, p->U.data(), p->U.size() * sizeof(double)))) return rc;
    if ((rc = tkmi::upload((void **)&p->d_Ut, p->Ut.data(), p->Ut.size() * sizeof(double)))) return rc;
    if ((rc = tkmi::upload((void **)&p->fwd.d_blocks, p->fwd.blocks.data(),
                           p->fwd.blocks.size() * sizeof(TBlock))))
        return rc;
    if ((rc = tkmi::upload((void **)&p->adj.d_blocks, p->adj.blocks.data(),
                           p->adj.blocks.size() * sizeof(TBlock))))
        return rc;
    return TKMI_OK;
}

void free_plan(tkmi_transform_plan_s *p) {
    if (!p) return;
    (void)hipFree(p->d_trees);
    (void)hipFree(p->d_U);
    (void)hipFree(p->d_Ut);
    (void)hipFree(p->fwd.d_blocks);
    (void)hipFree(p->adj.d_blocks);
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
// read trees come first in `trs` (nread of them), then write trees.
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

template <typename T>
int launch_side(const tkmi_transform_plan_s *p, const PlanSide &side, const double *dU, const T *src, T *dst,
                double alpha, double beta, hipStream_t stream) {
    if (side.ntiles == 0) return TKMI_OK;
    const int64_t chunk = (int64_t)1 << 30;
    for (int64_t t0 = 0; t0 < side.ntiles; t0 += chunk) {
        unsigned grid = (unsigned)std::min(chunk, side.ntiles - t0);
        switch (p->L) {
            case 1:
                hipLaunchKernelGGL((transform_kernel<T, 1>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                                   (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, alpha, beta, t0);
                break;
            case 2:
                hipLaunchKernelGGL((transform_kernel<T, 2>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                                   (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, alpha, beta, t0);
                break;
            case 4:
                hipLaunchKernelGGL((transform_kernel<T, 4>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                                   (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, alpha, beta, t0);
                break;
            default:
                hipLaunchKernelGGL((transform_kernel<T, 8>), dim3(grid), dim3(NTHREADS), 0, stream, side.d_blocks,
                                   (int64_t)side.blocks.size(), p->d_trees, dU, src, dst, alpha, beta, t0);
                break;
        }
    }
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

int exec_common(tkmi_transform_plan p, int32_t dtype, const void *src, void *dst, double alpha, double beta,
                void *stream, bool adjoint) {
    if (!p) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    const PlanSide &side = adjoint ? p->adj : p->fwd;
    const double *dU = adjoint ? p->d_Ut : p->d_U;
    if (side.ntiles > 0 && (!src || !dst)) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null data pointer");
    hipStream_t s = tkmi::as_stream(stream);
    if (dtype == TKMI_F64)
        return launch_side<double>(p, side, dU, (const double *)src, (double *)dst, alpha, beta, s);
    if (dtype == TKMI_F32)
        return launch_side<float>(p, side, dU, (const float *)src, (float *)dst, alpha, beta, s);
    TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "transform: dtype must be TKMI_F64 or TKMI_F32");
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
    p->L = 1;
    p->dst_len = dst_len;
    p->src_len = src_len;
    std::vector<HostItem> items;
    for (int64_t e = 0; e < nentries; ++e) {
        const int64_t *ds = dst_structs + e * (2 * N + 1);
        const int64_t *ss = src_structs + e * (2 * N + 1);
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
    int64_t maxL = 1;
    for (int64_t b = 0; b < nblocks; ++b) maxL = std::max(maxL, std::max(rows[b], cols[b]));
    if (maxL > 8) TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "tkmi_generic_plan_create: more than 8 trees in a fusion block");
    auto *p = new tkmi_transform_plan_s();
    p->N = N;
    p->L = maxL <= 1 ? 1 : (maxL <= 2 ? 2 : (maxL <= 4 ? 4 : 8));
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
    return exec_common(plan, dtype, src, dst, alpha, beta, stream, false);
}

int tkmi_transform_adjoint_exec(tkmi_transform_plan plan, int32_t dtype, const void *ddst, void *dsrc,
                                double alpha, double beta, void *stream) {
    return exec_common(plan, dtype, ddst, dsrc, alpha, beta, stream, true);
}

int tkmi_transform_plan_info(tkmi_transform_plan plan, int64_t *ntiles, int64_t *nentries, int64_t *elements) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "transform: null plan");
    if (ntiles) *ntiles = plan->fwd.ntiles;
    if (nentries) *nentries = plan->nitems;
    if (elements) *elements = plan->elements;
    return TKMI_OK;
}

int tkmi_transform_plan_destroy(tkmi_transform_plan plan) {
    free_plan(plan);
    return TKMI_OK;
}

}  // extern "C"
