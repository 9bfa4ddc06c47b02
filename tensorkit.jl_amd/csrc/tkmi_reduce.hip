// Deterministic dot product of two device vectors (two passes, fixed order).
// Used for TensorKit's `inner` (Σ_c dim(c)·⟨A_c, B_c⟩) in the reverse-mode rules.
#include "tkmi_internal.h"

#include <algorithm>
#include <mutex>
#include <vector>

namespace {

constexpr int RT = 256;
constexpr int MAXPARTS = 1024;

// Partial sums of tkmi_dot_noalloc: 1 + MAXPARTS doubles (8 200 bytes) per device, held
// until the process ends.  Host code never passes `scratch` itself (on the host it is a
// stand-in, not a device address): the device address is resolved once per device with
// hipGetSymbolAddress and only that address reaches kernels.
constexpr int SCRATCH_LEN = 1 + MAXPARTS;
__device__ __align__(256) double scratch[SCRATCH_LEN];

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device index, resolved with that device current

// Device address of this device's `scratch` (resolved on first use per device).
int scratch_for_current_device(double **out) {
    int dev = -1;
    TKMI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        TKMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) TKMI_FAIL(TKMI_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = (double *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return TKMI_OK;
}

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

// out[0] = Σ parts[1 .. nparts] (out may be parts itself)
__global__ __launch_bounds__(RT) void dot_final(const double *parts, int nparts, double *out) {
    __shared__ double sh[RT];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += RT) s += parts[1 + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

int dot_launch(int32_t dtype, const void *a, const void *b, int64_t n, double *parts, double *out,
               hipStream_t s) {
    const int nparts = (int)std::min<int64_t>(MAXPARTS, std::max<int64_t>(1, (n + RT * 8 - 1) / (RT * 8)));
    if (dtype == TKMI_F64)
        hipLaunchKernelGGL(dot_partial<double>, dim3(nparts), dim3(RT), 0, s, (const double *)a, (const double *)b, n,
                           parts);
    else if (dtype == TKMI_F32)
        hipLaunchKernelGGL(dot_partial<float>, dim3(nparts), dim3(RT), 0, s, (const float *)a, (const float *)b, n,
                           parts);
    else
        TKMI_FAIL(TKMI_ERR_UNSUPPORTED, "tkmi_dot: dtype must be TKMI_F64 or TKMI_F32");
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(RT), 0, s, parts, nparts, out);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

}  // namespace

namespace tkmi {
// Resolve this device's scratch address at set-up (tkmi_init), with the device current.
int reduce_setup() {
    double *p = nullptr;
    return scratch_for_current_device(&p);
}
}  // namespace tkmi

extern "C" int tkmi_dot(int32_t dtype, const void *a, const void *b, int64_t n, double *work, void *stream) {
    if (!work) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: null workspace");
    if (n < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: negative length");
    if (n > 0 && (!a || !b)) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot: null operand");
    return dot_launch(dtype, a, b, n, work, work, tkmi::as_stream(stream));
}

extern "C" int tkmi_dot_noalloc(int32_t dtype, const void *a, const void *b, int64_t n, double *result,
                                void *stream) {
    if (!result) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot_noalloc: null result pointer");
    if (n < 0) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot_noalloc: negative length");
    if (n > 0 && (!a || !b)) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_dot_noalloc: null operand");
    // the partials of any n fit: their number is capped at MAXPARTS (each partial sums a
    // strided share of the vector), so the scratch never grows with the input
    static_assert(1 + MAXPARTS <= SCRATCH_LEN, "partials exceed the static scratch");
    double *parts = nullptr;
    int rc = scratch_for_current_device(&parts);
    if (rc) return rc;
    return dot_launch(dtype, a, b, n, parts, result, tkmi::as_stream(stream));
}
