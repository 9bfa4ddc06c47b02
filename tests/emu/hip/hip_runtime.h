// CPU stand-in for the part of the HIP runtime and of the device built-ins that libtkmi's
// kernels use, so that tkmi_transform.hip, tkmi_gemm.hip and tkmi_jacobi.hip compile unchanged as host C++
// (clang++ -x c++) for a bounds-checked run under AddressSanitizer.  Test infrastructure:
// nothing here is part of the library.
//
// Execution model: a launch runs its workgroups one after another; a workgroup's threads
// are real host threads (a pool), so __syncthreads() is a barrier over the workgroup and
// every MFMA builtin is a rendezvous of the 64 lanes of one wave.  __shared__ arrays become
// function-local statics: one copy, used by one workgroup at a time.  Device memory is host
// memory (hipMalloc = malloc), so an out-of-bounds access in a kernel is an out-of-bounds
// access of a heap (or static LDS) array, which AddressSanitizer reports.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_index {
    unsigned x, y, z;
};
extern thread_local emu_index threadIdx;
extern emu_index blockIdx, blockDim, gridDim;

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;

inline hipError_t hipMalloc(void **p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T **p, size_t n) {
    return hipMalloc((void **)p, n);
}
inline hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemset(void *d, int v, size_t n) {
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
    *v = 256;
    return hipSuccess;
}
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t) {
    *n = 4;
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) {
    *e = nullptr;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 1.0f;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }

using std::fma;

template <typename T>
inline T min(T a, T b) {
    return b < a ? b : a;
}
template <typename T>
inline T max(T a, T b) {
    return a < b ? b : a;
}

namespace emu {
// Runs fn on block.x threads for each of grid.x workgroups (in order).
void launch(dim3 grid, dim3 block, const std::function<void()> &fn);
void syncthreads();
void wave_sync();
// per-wave exchange area for the MFMA rendezvous (double-sized slots, 64 lanes × 8)
double *wave_area(int which);
}  // namespace emu

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    ::emu::launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })
#define __syncthreads() ::emu::syncthreads()

namespace emu {
// D = C + A·B over one wave: lane l supplies A[i][k], B[k][j] (i or j = l % 16,
// k = KE·(l / 16) + e for its KE values) and C/D at the documented output rows.
template <int KE, bool ROWS_INTERLEAVED, typename AV, typename CV>
inline CV mfma(const AV &a, const AV &b, CV c) {
    const int lane = threadIdx.x & 63;
    double *A = wave_area(0), *B = wave_area(1);
    for (int e = 0; e < KE; ++e) {
        A[lane * KE + e] = (double)a[e];
        B[lane * KE + e] = (double)b[e];
    }
    wave_sync();
    const int j = lane % 16;
    for (int v = 0; v < 4; ++v) {
        const int i = ROWS_INTERLEAVED ? (lane / 16) + 4 * v : 4 * (lane / 16) + v;
        double s = (double)c[v];
        for (int g = 0; g < 4; ++g)           // lane group g holds k = KE·g + e
            for (int e = 0; e < KE; ++e) s += A[(16 * g + i) * KE + e] * B[(16 * g + j) * KE + e];
        c[v] = s;
    }
    wave_sync();  // every lane has read the exchange area before it is written again
    return c;
}
template <typename V>
struct one {
    V v;
    double operator[](int) const { return (double)v; }
};
}  // namespace emu

// global → LDS loads (global_load_lds_dwordx4): lane l's `size` bytes land at the wave's
// LDS pointer + offset + l·size, at once (so the waits before reading them are no-ops);
// the address-space qualifier of the LDS pointer has no host meaning
namespace emu {
inline void global_load_lds(const void *src, void *lds, unsigned size, unsigned offset) {
    std::memcpy(static_cast<char *>(lds) + offset + (threadIdx.x & 63) * size, src, size);
}
}  // namespace emu
#define address_space(n)
#define __builtin_amdgcn_global_load_lds(src, lds, size, offset, aux) ::emu::global_load_lds(src, lds, size, offset)
#define __builtin_amdgcn_s_waitcnt(x) ((void)0)
#define __builtin_amdgcn_s_barrier() ::emu::syncthreads()

#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) \
    ::emu::mfma<1, true>(::emu::one<double>{a}, ::emu::one<double>{b}, c)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) \
    ::emu::mfma<1, false>(::emu::one<float>{a}, ::emu::one<float>{b}, c)
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) ::emu::mfma<8, false>(a, b, c)
