// Internal helpers shared by the tkmi translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tkmi.h"

namespace tkmi {

void set_error(const std::string &msg);

#define TKMI_FAIL(code, msg)          \
    do {                              \
        ::tkmi::set_error(msg);       \
        return (code);                \
    } while (0)

#define TKMI_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            ::tkmi::set_error(std::string(#call) + ": " + hipGetErrorString(_e));        \
            return TKMI_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

// Upload a host array into a fresh device buffer (synchronous; plans only).
int upload(void **dptr, const void *host, size_t bytes);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Per-device set-up of the reductions' static scratch (called by tkmi_init).
int reduce_setup();

// Block one-sided Jacobi SVD of a set of column-major FP64 blocks (tkmi_jacobi.hip), with
// the output packing of tkmi_svd_exec; the workspace is made once per block layout.
struct JacobiWork;
int jacobi_create(JacobiWork **out, int64_t nblocks, const int64_t *m, const int64_t *n, const int64_t *a_off,
                  const int64_t *lda, const int64_t *u_off, const int64_t *s_off, const int64_t *vt_off);
void jacobi_destroy(JacobiWork *w);
// `sweeps` (may be null) receives the number of sweeps the slowest block needed
int jacobi_svd(JacobiWork *w, const double *A, double *U, double *S, double *Vt, hipStream_t s, int *sweeps);

}  // namespace tkmi
