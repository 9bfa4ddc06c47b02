// Runtime part of the tkmi C ABI: errors, device memory, streams, hipGraph capture.
#include "tkmi_internal.h"

#include <mutex>

namespace tkmi {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int upload(void **dptr, const void *host, size_t bytes) {
    *dptr = nullptr;
    if (bytes == 0) bytes = 16;  // keep a valid, non-null pointer for empty tables
    TKMI_HIP(hipMalloc(dptr, bytes));
    if (host) {
        hipError_t e = hipMemcpy(*dptr, host, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(*dptr);
            *dptr = nullptr;
            set_error(std::string("hipMemcpy (plan upload): ") + hipGetErrorString(e));
            return TKMI_ERR_HIP;
        }
    }
    return TKMI_OK;
}

}  // namespace tkmi

struct tkmi_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int tkmi_version(void) { return 100; }

const char *tkmi_last_error(void) { return tkmi::g_last_error.c_str(); }

int tkmi_init(int device) {
    int n = 0;
    TKMI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_init: device index out of range");
    TKMI_HIP(hipSetDevice(device));
    return tkmi::reduce_setup();
}

int tkmi_alloc(void **ptr, size_t bytes) {
    if (!ptr) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_alloc: null output pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *ptr = nullptr;
        TKMI_FAIL(TKMI_ERR_ALLOC, std::string("tkmi_alloc: ") + hipGetErrorString(e));
    }
    return TKMI_OK;
}

int tkmi_free(void *ptr) {
    if (ptr) TKMI_HIP(hipFree(ptr));
    return TKMI_OK;
}

int tkmi_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
    TKMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, tkmi::as_stream(stream)));
    return TKMI_OK;
}

int tkmi_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
    TKMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, tkmi::as_stream(stream)));
    TKMI_HIP(hipStreamSynchronize(tkmi::as_stream(stream)));
    return TKMI_OK;
}

int tkmi_stream_synchronize(void *stream) {
    TKMI_HIP(hipStreamSynchronize(tkmi::as_stream(stream)));
    return TKMI_OK;
}

int tkmi_graph_begin(void *stream) {
    if (!stream) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_graph_begin: capture needs a non-default stream");
    TKMI_HIP(hipStreamBeginCapture(tkmi::as_stream(stream), hipStreamCaptureModeThreadLocal));
    return TKMI_OK;
}

int tkmi_graph_end(void *stream, tkmi_graph *graph) {
    if (!graph) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_graph_end: null output");
    auto *g = new tkmi_graph_s();
    hipError_t e = hipStreamEndCapture(tkmi::as_stream(stream), &g->graph);
    if (e != hipSuccess) {
        delete g;
        TKMI_FAIL(TKMI_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g->graph);
        delete g;
        TKMI_FAIL(TKMI_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *graph = g;
    return TKMI_OK;
}

int tkmi_graph_launch(tkmi_graph graph, void *stream) {
    if (!graph) TKMI_FAIL(TKMI_ERR_ARGUMENT, "tkmi_graph_launch: null graph");
    TKMI_HIP(hipGraphLaunch(graph->exec, tkmi::as_stream(stream)));
    return TKMI_OK;
}

int tkmi_graph_destroy(tkmi_graph graph) {
    if (!graph) return TKMI_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return TKMI_OK;
}

}  // extern "C"
