// Block SVD and truncation selection for svd_trunc, all FP64 on the device.
//
// tkmi_svd_exec   per coupled block A_c (d1×d2, column-major): A_c = U_c·diag(S_c)·Vt_c
//                 (thin: r_c = min(d1, d2)) with a rocSOLVER driver on the caller's stream:
//                 QR iteration (dgesvd), divide and conquer (dgesdd) or one-sided Jacobi
//                 (dgesvdj), MatrixAlgebraKit's ROCSOLVER_* algorithm choices
//                 (MAK.svd_compact! per block, src/factorizations/matrixalgebrakit.jl:34-51).
//                 Blocks are independent: they run on up to 8 concurrent lanes (stream +
//                 rocBLAS handle each) forked from and joined back to the caller's stream.
// tkmi_eigh_exec  per square block A_c = V_c·diag(λ_c)·V_cᵀ (symmetric, λ ascending) with
//                 dsyevd / dsyevj / dsyev on the same lanes (MAK.eigh_full! per block).
// tkmi_truncate   keep-mask over all singular values of all blocks for
//                 truncrank(howmany) ∩ trunctol(atol) with quantum-dimension weights, and
//                 the truncation error ε = sqrt(Σ_discarded w·σ²)
//                 (findtruncated for TruncationByOrder / TruncationByValue / intersection,
//                 truncation_error!: src/factorizations/truncation.jl:171-322).
//                 Ranking is O(R²) over R ≤ a few thousand values: every value counts the
//                 values ranked before it (descending σ, ties by index), so the kept set is
//                 exactly the reference's sequential scan of the sorted values.
#include "tkmi_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int ST = 256;

__global__ __launch_bounds__(ST) void rank_select(const double *__restrict__ S, const double *__restrict__ w,
                                                  int64_t R, double howmany, double atol,
                                                  unsigned char *__restrict__ keep, double *__restrict__ disc) {
    // the values and weights pass through LDS in tiles of ST, read by every thread
    __shared__ double ts[ST], tw[ST];
    const int64_t i = (int64_t)blockIdx.x * ST + threadIdx.x;
    const double si = i < R ? fabs(S[i]) : 0.0;
    double before = 0.0;  // weight of values ranked before i, added in index order
    for (int64_t j0 = 0; j0 < R; j0 += ST) {
        const int64_t j = j0 + threadIdx.x;
        ts[threadIdx.x] = j < R ? fabs(S[j]) : 0.0;
        tw[threadIdx.x] = j < R ? w[j] : 0.0;
        __syncthreads();
        const int n = (int)min((int64_t)ST, R - j0);
        for (int q = 0; q < n; ++q) {
            const double sj = ts[q];
            if (sj > si || (sj == si && j0 + q < i)) before += tw[q];
        }
        __syncthreads();
    }
    if (i >= R) return;
    const bool k = (before + w[i] <= howmany) && (si >= atol);
    keep[i] = k ? 1 : 0;
    disc[i] = k ? 0.0 : w[i] * S[i] * S[i];
}

__global__ __launch_bounds__(ST) void sum_sqrt(const double *__restrict__ x, int64_t n, double *__restrict__ out) {
    __shared__ double sh[ST];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += ST) s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = ST / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(sh[0]);
}

}  // namespace

// One lane = one stream + rocBLAS handle + scratch: the blocks of a plan are dealt
// out over up to MAX_LANES lanes (largest first, to the least loaded lane) and the
// lanes run concurrently, each driven by its own host thread so that a driver that
// waits on the host for its stream holds up only its own lane.
constexpr int MAX_LANES = 8;

struct SvdLane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    rocblas_handle handle = nullptr;
    double *work = nullptr;  // packed copy of one block (the rocSOLVER drivers overwrite their input)
    double *E = nullptr;
    double *residual = nullptr;
    rocblas_int *info = nullptr, *sweeps = nullptr;
    std::vector<size_t> blocks;
    int64_t maxmn = 0, maxr = 0;
};

struct tkmi_svd_plan_s {
    int32_t kind = 0;  // 0: SVD (U, S, Vt), 1: symmetric eigendecomposition (V in U's slot, λ in S's)
    std::vector<int64_t> m, n, a_off, lda, r, u_off, s_off, v_off;
    int64_t u_len = 0, s_len = 0, v_len = 0;
    int32_t alg = TKMI_SVD_QRITERATION;
    int device = 0;
    hipEvent_t start = nullptr;
    std::vector<SvdLane> lanes;
    tkmi::JacobiWork *jac = nullptr;  // block Jacobi workspace, made by the first call with that algorithm
};

namespace {
void free_svd(tkmi_svd_plan_s *p) {
    if (!p) return;
    // work still queued on a lane uses its workspace: drain every lane first
    for (auto &l : p->lanes)
        if (l.stream) (void)hipStreamSynchronize(l.stream);
    for (auto &l : p->lanes) {
        if (l.handle) rocblas_destroy_handle(l.handle);
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.stream) (void)hipStreamDestroy(l.stream);
        (void)hipFree(l.work);
        (void)hipFree(l.E);
        (void)hipFree(l.residual);
        (void)hipFree(l.info);
        (void)hipFree(l.sweeps);
    }
    if (p->start) (void)hipEventDestroy(p->start);
    tkmi::jacobi_destroy(p->jac);
    delete p;
}

// Factorise the blocks of one lane on its stream; returns a TKMI status (errors are
// reported by the calling thread).
int run_lane(const tkmi_svd_plan_s *p, SvdLane &l, const double *A, double *U, double *S, double *Vt,
             std::string &err) {
    if (hipSetDevice(p->device) != hipSuccess) { err = "svd: hipSetDevice failed"; return TKMI_ERR_HIP; }
    if (rocblas_set_stream(l.handle, l.stream) != rocblas_status_success) {
        err = "svd: rocblas_set_stream failed";
        return TKMI_ERR_HIP;
    }
    for (size_t b : l.blocks) {
        const int64_t m = p->m[b], n = p->n[b], r = p->r[b];
        const rocblas_int im = (rocblas_int)m, in = (rocblas_int)n, ir = (rocblas_int)r;
        double *Sb = S + p->s_off[b], *Ub = U + p->u_off[b], *Vb = Vt ? Vt + p->v_off[b] : nullptr;
        // the drivers overwrite their input: SVD works on a packed copy, the
        // eigensolvers on the block copied into the eigenvector output itself
        if (hipMemcpy2DAsync(p->kind == 1 ? Ub : l.work, m * sizeof(double), A + p->a_off[b],
                             p->lda[b] * sizeof(double), m * sizeof(double), n, hipMemcpyDeviceToDevice,
                             l.stream) != hipSuccess) {
            err = "svd: block copy failed";
            return TKMI_ERR_HIP;
        }
        rocblas_status st;
        if (p->kind == 1) {
            // eigenvalues ascending, orthonormal eigenvectors in place (lower triangle read)
            if (p->alg == TKMI_SVD_DIVIDEANDCONQUER)
                st = rocsolver_dsyevd(l.handle, rocblas_evect_original, rocblas_fill_lower, in, Ub, in, Sb, l.E,
                                      l.info);
            else if (p->alg == TKMI_SVD_JACOBI)
                st = rocsolver_dsyevj(l.handle, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_lower,
                                      in, Ub, in, 0.0, l.residual, 100, l.sweeps, Sb, l.info);
            else
                st = rocsolver_dsyev(l.handle, rocblas_evect_original, rocblas_fill_lower, in, Ub, in, Sb, l.E,
                                     l.info);
            if (st != rocblas_status_success) {
                err = "eigh: rocSOLVER driver failed";
                return TKMI_ERR_HIP;
            }
            continue;
        }
        if (p->alg == TKMI_SVD_DIVIDEANDCONQUER)
            st = rocsolver_dgesdd(l.handle, rocblas_svect_singular, rocblas_svect_singular, im, in, l.work, im, Sb,
                                  Ub, im, Vb, ir, l.info);
        else if (p->alg == TKMI_SVD_JACOBI)
            // abstol ≤ 0: converge to machine precision
            st = rocsolver_dgesvdj(l.handle, rocblas_svect_singular, rocblas_svect_singular, im, in, l.work, im, 0.0,
                                   l.residual, 100, l.sweeps, Sb, Ub, im, Vb, ir, l.info);
        else
            st = rocsolver_dgesvd(l.handle, rocblas_svect_singular, rocblas_svect_singular, im, in, l.work, im, Sb,
                                  Ub, im, Vb, ir, l.E, rocblas_outofplace, l.info);
        if (st != rocblas_status_success) {
            err = "svd: rocSOLVER driver failed";
            return TKMI_ERR_HIP;
        }
    }
    if (hipEventRecord(l.done, l.stream) != hipSuccess) { err = "svd: hipEventRecord failed"; return TKMI_ERR_HIP; }
    return TKMI_OK;
}
}  // namespace

extern "C" {

static int factor_plan_create(tkmi_svd_plan *plan, int32_t kind, int64_t nblocks, const int64_t *m,
                              const int64_t *n, const int64_t *a_off, const int64_t *lda) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_svd_plan_create: null plan pointer");
    *plan = nullptr;
    auto *p = new tkmi_svd_plan_s();
    p->kind = kind;
    std::vector<size_t> order;
    for (int64_t b = 0; b < nblocks; ++b) {
        if (m[b] < 0 || n[b] < 0 || lda[b] < std::max<int64_t>(1, m[b]) || m[b] > INT32_MAX || n[b] > INT32_MAX) {
            free_svd(p);
            TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_svd_plan_create: bad block dimensions");
        }
        if (kind == 1 && m[b] != n[b]) {
            free_svd(p);
            TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_eigh_plan_create: blocks must be square");
        }
        const int64_t r = std::min(m[b], n[b]);
        p->m.push_back(m[b]);
        p->n.push_back(n[b]);
        p->a_off.push_back(a_off[b]);
        p->lda.push_back(lda[b]);
        p->r.push_back(r);
        p->u_off.push_back(p->u_len);
        p->s_off.push_back(p->s_len);
        p->v_off.push_back(p->v_len);
        p->u_len += m[b] * r;
        p->s_len += r;
        p->v_len += kind == 1 ? 0 : r * n[b];
        if (r > 0) order.push_back((size_t)b);
    }
    // cost model m·n·r; largest block first onto the least loaded lane
    auto cost = [&](size_t b) { return (double)p->m[b] * (double)p->n[b] * (double)p->r[b]; };
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
    const int nl = (int)std::min<size_t>(MAX_LANES, order.size());
    p->lanes.resize(nl);
    std::vector<double> load(nl, 0.0);
    for (size_t b : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += cost(b);
        SvdLane &l = p->lanes[k];
        l.blocks.push_back(b);
        l.maxmn = std::max(l.maxmn, p->m[b] * p->n[b]);
        l.maxr = std::max(l.maxr, p->r[b]);
    }
    if (hipGetDevice(&p->device) != hipSuccess || hipEventCreateWithFlags(&p->start, hipEventDisableTiming) != hipSuccess) {
        free_svd(p);
        TKMI_FAIL(TKMI_ERR_HIP, "tkmi_svd_plan_create: event creation failed");
    }
    for (auto &l : p->lanes) {
        if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess ||
            rocblas_create_handle(&l.handle) != rocblas_status_success) {
            free_svd(p);
            TKMI_FAIL(TKMI_ERR_HIP, "tkmi_svd_plan_create: stream / rocBLAS handle creation failed");
        }
        int rc;
        if ((rc = tkmi::upload((void **)&l.work, nullptr, std::max<int64_t>(1, kind == 1 ? 1 : l.maxmn) * sizeof(double))) ||
            (rc = tkmi::upload((void **)&l.E, nullptr, std::max<int64_t>(1, l.maxr) * sizeof(double))) ||
            (rc = tkmi::upload((void **)&l.residual, nullptr, sizeof(double))) ||
            (rc = tkmi::upload((void **)&l.sweeps, nullptr, sizeof(rocblas_int))) ||
            (rc = tkmi::upload((void **)&l.info, nullptr, sizeof(rocblas_int)))) {
            free_svd(p);
            return rc;
        }
    }
    *plan = p;
    return TKMI_OK;
}

int tkmi_svd_plan_create(tkmi_svd_plan *plan, int64_t nblocks, const int64_t *m, const int64_t *n,
                         const int64_t *a_off, const int64_t *lda) {
    return factor_plan_create(plan, 0, nblocks, m, n, a_off, lda);
}

int tkmi_eigh_plan_create(tkmi_svd_plan *plan, int64_t nblocks, const int64_t *n, const int64_t *a_off,
                          const int64_t *lda) {
    return factor_plan_create(plan, 1, nblocks, n, n, a_off, lda);
}

int tkmi_svd_plan_sizes(tkmi_svd_plan plan, int64_t *u_len, int64_t *s_len, int64_t *vt_len) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: null plan");
    if (u_len) *u_len = plan->u_len;
    if (s_len) *s_len = plan->s_len;
    if (vt_len) *vt_len = plan->v_len;
    return TKMI_OK;
}

int tkmi_eigh_exec(tkmi_svd_plan plan, const double *A, double *W, double *V, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "eigh: null plan");
    if (plan->kind != 1) TKMI_FAIL(TKMI_ERR_ARGUMENT, "eigh: plan was not made by tkmi_eigh_plan_create");
    return tkmi_svd_exec(plan, A, V, W, nullptr, stream);
}

int tkmi_svd_exec(tkmi_svd_plan plan, const double *A, double *U, double *S, double *Vt, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: null plan");
    if (plan->lanes.empty()) return TKMI_OK;
    hipStream_t s = tkmi::as_stream(stream);
    if (plan->alg == TKMI_SVD_BLOCKJACOBI) {
        // every block in the same launches on the caller's stream
        if (!plan->jac) {
            const size_t nb = plan->m.size();
            int rc = tkmi::jacobi_create(&plan->jac, (int64_t)nb, plan->m.data(), plan->n.data(), plan->a_off.data(),
                                         plan->lda.data(), plan->u_off.data(), plan->s_off.data(), plan->v_off.data());
            if (rc) return rc;
        }
        return tkmi::jacobi_svd(plan->jac, A, U, S, Vt, s, nullptr);
    }
    // fork: every lane starts after the work already queued on the caller's stream
    TKMI_HIP(hipEventRecord(plan->start, s));
    for (auto &l : plan->lanes) TKMI_HIP(hipStreamWaitEvent(l.stream, plan->start, 0));
    const size_t nl = plan->lanes.size();
    std::vector<int> rc(nl, TKMI_OK);
    std::vector<std::string> err(nl);
    if (nl == 1) {
        rc[0] = run_lane(plan, plan->lanes[0], A, U, S, Vt, err[0]);
    } else {
        std::vector<std::thread> th;
        for (size_t k = 0; k < nl; ++k)
            th.emplace_back([&, k] { rc[k] = run_lane(plan, plan->lanes[k], A, U, S, Vt, err[k]); });
        for (auto &t : th) t.join();
    }
    for (size_t k = 0; k < nl; ++k)
        if (rc[k] != TKMI_OK) TKMI_FAIL(rc[k], err[k]);
    // join: the caller's stream continues once every lane is done
    for (auto &l : plan->lanes) TKMI_HIP(hipStreamWaitEvent(s, l.done, 0));
    return TKMI_OK;
}

int tkmi_svd_plan_set_algorithm(tkmi_svd_plan plan, int32_t alg) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: null plan");
    if (alg != TKMI_SVD_QRITERATION && alg != TKMI_SVD_DIVIDEANDCONQUER && alg != TKMI_SVD_JACOBI &&
        alg != TKMI_SVD_BLOCKJACOBI)
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: unknown algorithm");
    if (alg == TKMI_SVD_BLOCKJACOBI && plan->kind == 1)
        TKMI_FAIL(TKMI_ERR_ARGUMENT, "eigh: block Jacobi is an SVD algorithm");
    plan->alg = alg;
    return TKMI_OK;
}

int tkmi_svd_plan_destroy(tkmi_svd_plan plan) {
    free_svd(plan);
    return TKMI_OK;
}

int tkmi_truncate_select(int64_t nvals, const double *S, const double *weights, double howmany, double atol,
                         unsigned char *keep, double *disc_work, double *err, void *stream) {
    if (nvals < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "truncate: negative length");
    hipStream_t s = tkmi::as_stream(stream);
    if (nvals > 0) {
        hipLaunchKernelGGL(rank_select, dim3((unsigned)((nvals + ST - 1) / ST)), dim3(ST), 0, s, S, weights, nvals,
                           howmany, atol, keep, disc_work);
    }
    hipLaunchKernelGGL(sum_sqrt, dim3(1), dim3(ST), 0, s, disc_work, nvals, err);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

}  // extern "C"
