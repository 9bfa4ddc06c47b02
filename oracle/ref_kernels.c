/*
 * Oracle (test infrastructure): the per-sub-block permute loop of TensorKit's
 * add_transform_kernel! in plain C.
 *
 *   abelian  src/tensors/indexmanipulations.jl:658-669
 *            for each entry: dst_sub = β·dst_sub + (α·coeff)·permutedims(src_sub, p)
 *            (TO.tensoradd! on StridedViews)
 *   generic  src/tensors/indexmanipulations.jl:670-726
 *            for each fusion block: buffer_dst = buffer_src·Uᵀ, then
 *            dst_j = β·dst_j + α·permutedims(buffer_dst[:, j], p)
 *
 * Built with -ffp-contract=off: β·dst and (α·coeff)·src are rounded separately and
 * then added, as Julia evaluates scale(C, β) + scale(A, α) (no fused multiply-add).
 *
 * Entries are distributed over OpenMP threads (the reference's tforeach /
 * @tasks over sub-blocks when TRANSFORMER_THREADS > 1).  Strided loops use an
 * odometer over the destination index; the destination's fastest dim is the
 * innermost loop.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 16

/* dst[view] = b·dst[view] + a·src[view'] over the outer iterations [o0, o1) of one
 * strided view (outer = every dimension but the first, first dimension innermost). */
static void strided_add_range(int N, const int64_t *sz, const int64_t *dstr, int64_t doff,
                              const int64_t *sstr_dst_order, int64_t soff, const double *src, double *dst, double a,
                              double b, int64_t o0, int64_t o1) {
    int64_t idx[MAXN];
    int64_t rem = o0;
    for (int k = 1; k < N; ++k) {
        idx[k] = rem % sz[k];
        rem /= sz[k];
    }
    const int64_t n0 = N > 0 ? sz[0] : 1;
    const int64_t ds0 = N > 0 ? dstr[0] : 0, ss0 = N > 0 ? sstr_dst_order[0] : 0;
    for (int64_t outer = o0; outer < o1; ++outer) {
        int64_t dp = doff, sp = soff;
        for (int k = 1; k < N; ++k) {
            dp += idx[k] * dstr[k];
            sp += idx[k] * sstr_dst_order[k];
        }
        if (b == 0.0) {
            for (int64_t i = 0; i < n0; ++i) dst[dp + i * ds0] = a * src[sp + i * ss0];
        } else {
            for (int64_t i = 0; i < n0; ++i) dst[dp + i * ds0] = b * dst[dp + i * ds0] + a * src[sp + i * ss0];
        }
        for (int k = 1; k < N; ++k) {
            if (++idx[k] < sz[k]) break;
            idx[k] = 0;
        }
    }
}

static void strided_add(int N, const int64_t *sz, const int64_t *dstr, int64_t doff, const int64_t *sstr_dst_order,
                        int64_t soff, const double *src, double *dst, double a, double b) {
    int64_t total = 1;
    for (int k = 0; k < N; ++k) total *= sz[k];
    if (total == 0) return;
    const int64_t n0 = N > 0 ? sz[0] : 1;
    strided_add_range(N, sz, dstr, doff, sstr_dst_order, soff, src, dst, a, b, 0, total / n0);
}

/* structs: nentries × (2N+1) = sizes[N], strides[N], offset.  The work is split into
 * items of about CHUNK elements (whole runs of the first dimension), so that one large
 * sub-block spreads over every thread. */
#define CHUNK ((int64_t)1 << 14)
void oracle_abelian_permute(int N, const int32_t *perm, int64_t nentries, const double *coeffs,
                            const int64_t *dst_structs, const int64_t *src_structs, const double *src, double *dst,
                            double alpha, double beta, int nthreads) {
    int64_t nitems = 0;
    for (int64_t e = 0; e < nentries; ++e) {
        const int64_t *ds = dst_structs + e * (2 * N + 1);
        int64_t total = 1;
        for (int k = 0; k < N; ++k) total *= ds[k];
        if (total == 0) continue;
        const int64_t n0 = N > 0 ? ds[0] : 1, outer = total / n0;
        const int64_t step = n0 >= CHUNK ? 1 : CHUNK / n0;
        nitems += (outer + step - 1) / step;
    }
    int64_t *items = (int64_t *)malloc(sizeof(int64_t) * 3 * (nitems > 0 ? nitems : 1));
    int64_t q = 0;
    for (int64_t e = 0; e < nentries; ++e) {
        const int64_t *ds = dst_structs + e * (2 * N + 1);
        int64_t total = 1;
        for (int k = 0; k < N; ++k) total *= ds[k];
        if (total == 0) continue;
        const int64_t n0 = N > 0 ? ds[0] : 1, outer = total / n0;
        const int64_t step = n0 >= CHUNK ? 1 : CHUNK / n0;
        for (int64_t o = 0; o < outer; o += step) {
            items[3 * q] = e;
            items[3 * q + 1] = o;
            items[3 * q + 2] = o + step < outer ? o + step : outer;
            ++q;
        }
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t it = 0; it < nitems; ++it) {
        const int64_t e = items[3 * it];
        const int64_t *ds = dst_structs + e * (2 * N + 1);
        const int64_t *ss = src_structs + e * (2 * N + 1);
        int64_t sstr[MAXN];
        for (int k = 0; k < N; ++k) sstr[k] = ss[N + perm[k]];
        strided_add_range(N, ds, ds + N, ds[2 * N], sstr, ss[2 * N], src, dst, alpha * coeffs[e], beta,
                          items[3 * it + 1], items[3 * it + 2]);
    }
    free(items);
}

/* generic: per block b, rows[b]×cols[b] U (column-major), sizes N per block,
 * per tree (strides[N], offset). */
void oracle_generic_permute(int N, const int32_t *perm, int64_t nblocks, const int64_t *rows, const int64_t *cols,
                            const double *U, const int64_t *sz_dst, const int64_t *sz_src, const int64_t *so_dst,
                            const int64_t *so_src, const double *src, double *dst, double alpha, double beta,
                            int nthreads) {
    int64_t *uoff = (int64_t *)malloc(sizeof(int64_t) * (nblocks + 1));
    int64_t *toff_d = (int64_t *)malloc(sizeof(int64_t) * (nblocks + 1));
    int64_t *toff_s = (int64_t *)malloc(sizeof(int64_t) * (nblocks + 1));
    uoff[0] = toff_d[0] = toff_s[0] = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        uoff[b + 1] = uoff[b] + rows[b] * cols[b];
        toff_d[b + 1] = toff_d[b] + rows[b];
        toff_s[b + 1] = toff_s[b] + cols[b];
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t R = rows[b], C = cols[b];
        const int64_t *szs = sz_src + b * N, *szd = sz_dst + b * N;
        if (R == 1 && C == 1) {
            /* a block of one tree: tensoradd! with α·coeff, no recoupling GEMM
             * (indexmanipulations.jl:684-690) */
            const int64_t *sod = so_dst + toff_d[b] * (N + 1), *sos = so_src + toff_s[b] * (N + 1);
            int64_t sstr[MAXN];
            for (int k = 0; k < N; ++k) sstr[k] = sos[perm[k]];
            strided_add(N, szd, sod, sod[N], sstr, sos[N], src, dst, alpha * U[uoff[b]], beta);
            continue;
        }
        int64_t bs = 1;
        for (int k = 0; k < N; ++k) bs *= szs[k];
        /* 1. extract: buffer_src[:, i] = src_i in canonical (column-major) order */
        double *bsrc = (double *)malloc(sizeof(double) * bs * (C > 0 ? C : 1));
        double *bdst = (double *)malloc(sizeof(double) * bs * (R > 0 ? R : 1));
        int64_t cstr[MAXN];
        int64_t s = 1;
        for (int k = 0; k < N; ++k) {
            cstr[k] = s;
            s *= szs[k];
        }
        for (int64_t i = 0; i < C; ++i) {
            const int64_t *so = so_src + (toff_s[b] + i) * (N + 1);
            strided_add(N, szs, cstr, i * bs, so, so[N], src, bsrc, 1.0, 0.0);
        }
        /* 2. recouple: buffer_dst = buffer_src · Uᵀ */
        const double *Ub = U + uoff[b];
        for (int64_t j = 0; j < R; ++j)
            for (int64_t x = 0; x < bs; ++x) {
                double acc = 0.0;
                for (int64_t i = 0; i < C; ++i) acc += bsrc[i * bs + x] * Ub[j + i * R];
                bdst[j * bs + x] = acc;
            }
        /* 3. insert with the permutation p */
        for (int64_t j = 0; j < R; ++j) {
            const int64_t *so = so_dst + (toff_d[b] + j) * (N + 1);
            int64_t sstr[MAXN];
            for (int k = 0; k < N; ++k) sstr[k] = cstr[perm[k]];
            strided_add(N, szd, so, so[N], sstr, j * bs, bdst, dst, alpha, beta);
        }
        free(bsrc);
        free(bdst);
    }
    free(uoff);
    free(toff_d);
    free(toff_s);
}
