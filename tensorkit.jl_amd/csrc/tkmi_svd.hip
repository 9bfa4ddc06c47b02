// Block SVD and truncation selection for svd_trunc, all FP64 on the device.
//
// tkmi_svd_exec   per coupled block A_c (d1×d2, column-major): A_c = U_c·diag(S_c)·Vt_c
//                 (thin: r_c = min(d1, d2)) with rocSOLVER dgesvd on the caller's stream
//                 (MAK.svd_compact! per block, src/factorizations/matrixalgebrakit.jl:34-51).
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
#include <vector>

namespace {

constexpr int ST = 256;

__global__ __launch_bounds__(ST) void rank_select(const double *__restrict__ S, const double *__restrict__ w,
                                                  int64_t R, double howmany, double atol,
                                                  unsigned char *__restrict__ keep, double *__restrict__ disc) {
    const int64_t i = (int64_t)blockIdx.x * ST + threadIdx.x;
    if (i >= R) return;
    const double si = fabs(S[i]);
    double before = 0.0;  // weight of values ranked before i
    for (int64_t j = 0; j < R; ++j) {
        const double sj = fabs(S[j]);
        if (sj > si || (sj == si && j < i)) before += w[j];
    }
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

struct tkmi_svd_plan_s {
    std::vector<int64_t> m, n, a_off, lda, r, u_off, s_off, v_off;
    int64_t u_len = 0, s_len = 0, v_len = 0, maxmn = 0, maxr = 0;
    rocblas_handle handle = nullptr;
    double *work = nullptr;  // copy of one block (dgesvd overwrites its input)
    double *E = nullptr;
    rocblas_int *info = nullptr;
};

namespace {
void free_svd(tkmi_svd_plan_s *p) {
    if (!p) return;
    if (p->handle) rocblas_destroy_handle(p->handle);
    (void)hipFree(p->work);
    (void)hipFree(p->E);
    (void)hipFree(p->info);
    delete p;
}
}  // namespace

extern "C" {

int tkmi_svd_plan_create(tkmi_svd_plan *plan, int64_t nblocks, const int64_t *m, const int64_t *n,
                         const int64_t *a_off, const int64_t *lda) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_svd_plan_create: null plan pointer");
    *plan = nullptr;
    auto *p = new tkmi_svd_plan_s();
    for (int64_t b = 0; b < nblocks; ++b) {
        if (m[b] < 0 || n[b] < 0 || lda[b] < std::max<int64_t>(1, m[b]) || m[b] > INT32_MAX || n[b] > INT32_MAX) {
            free_svd(p);
            TKMI_FAIL(TKMI_ERR_SPACE_MISMATCH, "tkmi_svd_plan_create: bad block dimensions");
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
        p->v_len += r * n[b];
        p->maxmn = std::max(p->maxmn, m[b] * n[b]);
        p->maxr = std::max(p->maxr, r);
    }
    if (rocblas_create_handle(&p->handle) != rocblas_status_success) {
        free_svd(p);
        TKMI_FAIL(TKMI_ERR_HIP, "tkmi_svd_plan_create: rocblas_create_handle failed");
    }
    int rc;
    if ((rc = tkmi::upload((void **)&p->work, nullptr, std::max<int64_t>(1, p->maxmn) * sizeof(double))) ||
        (rc = tkmi::upload((void **)&p->E, nullptr, std::max<int64_t>(1, p->maxr) * sizeof(double))) ||
        (rc = tkmi::upload((void **)&p->info, nullptr, sizeof(rocblas_int)))) {
        free_svd(p);
        return rc;
    }
    *plan = p;
    return TKMI_OK;
}

int tkmi_svd_plan_sizes(tkmi_svd_plan plan, int64_t *u_len, int64_t *s_len, int64_t *vt_len) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: null plan");
    if (u_len) *u_len = plan->u_len;
    if (s_len) *s_len = plan->s_len;
    if (vt_len) *vt_len = plan->v_len;
    return TKMI_OK;
}

int tkmi_svd_exec(tkmi_svd_plan plan, const double *A, double *U, double *S, double *Vt, void *stream) {
    if (!plan) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: null plan");
    hipStream_t s = tkmi::as_stream(stream);
    if (rocblas_set_stream(plan->handle, s) != rocblas_status_success)
        TKMI_FAIL(TKMI_ERR_HIP, "svd: rocblas_set_stream failed");
    for (size_t b = 0; b < plan->m.size(); ++b) {
        const int64_t m = plan->m[b], n = plan->n[b], r = plan->r[b];
        if (r == 0) continue;
        // dgesvd overwrites A: work on a packed copy of the block
        TKMI_HIP(hipMemcpy2DAsync(plan->work, m * sizeof(double), A + plan->a_off[b], plan->lda[b] * sizeof(double),
                                  m * sizeof(double), n, hipMemcpyDeviceToDevice, s));
        rocblas_status st = rocsolver_dgesvd(plan->handle, rocblas_svect_singular, rocblas_svect_singular,
                                             (rocblas_int)m, (rocblas_int)n, plan->work, (rocblas_int)m,
                                             S + plan->s_off[b], U + plan->u_off[b], (rocblas_int)m,
                                             Vt + plan->v_off[b], (rocblas_int)r, plan->E, rocblas_outofplace,
                                             plan->info);
        if (st != rocblas_status_success) TKMI_FAIL(TKMI_ERR_HIP, "svd: rocsolver_dgesvd failed");
    }
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
