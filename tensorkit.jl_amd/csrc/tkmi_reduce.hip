// Deterministic dot product of two device vectors (two passes, fixed order).
// Used for TensorKit's `inner` (Σ_c dim(c)·⟨A_c, B_c⟩) in the reverse-mode rules.
#include "tkmi_internal.h"

namespace {

constexpr int RT = 256;
constexpr int MAXPARTS = 1024;

template <typename T>
__global__ __launch_bounds__(RT) void dot_partial(const T *__restrict__ a, const T *__restrict__ b, int64_t n,
                                                  double *__restrict__ parts) {
    __shared__ double sh[RT];
    // four independent strided partial sums per thread keep four load pairs in flight;
    // the order of every addition is fixed by (n, grid), so results are reproducible
    const int64_t stride = (int64_t)gridDim.x * RT;
    int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        s0 += (double)a[i] * (double)b[i];
        s1 += (double)a[i + stride] * (double)b[i + stride];
        s2 += (double)a[i + 2 * stride] * (double)b[i + 2 * stride];
        s3 += (double)a[i + 3 * stride] * (double)b[i + 3 * stride];
    }
    for (; i < n; i += stride) s0 += (double)a[i] * (double)b[i];
    sh[threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[1 + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RT) void dot_final(double *__restrict__ parts, int nparts) {
    __shared__ double sh[RT];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += RT) s += parts[1 + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[0] = sh[0];
}

}  // namespace

extern "C" int tkmi_dot(int32_t dtype, const void *a, const void *b, int64_t n, double *work, void *stream) {
    if (!work) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: null workspace");
    if (n < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: negative length");
    if (n > 0 && (!a || !b)) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: null operand");
    hipStream_t s = tkmi::as_stream(stream);
    int nparts = (int)std::min<int64_t>(MAXPARTS, std::max<int64_t>(1, (n + RT * 8 - 1) / (RT * 8)));
    if (dtype == TKMI_F64)
        hipLaunchKernelGGL(dot_partial<double>, dim3(nparts), dim3(RT), 0, s, (const double *)a, (const double *)b, n,
                           work);
    else if (dtype == TKMI_F32)
        hipLaunchKernelGGL(dot_partial<float>, dim3(nparts), dim3(RT), 0, s, (const float *)a, (const float *)b, n,
                           work);
    else
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "tkmi_dot: dtype must be TKMI_F64 or TKMI_F32");
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(RT), 0, s, work, nparts);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}
