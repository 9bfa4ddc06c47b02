// Host-thread execution of emulated kernel launches (see hip/hip_runtime.h), and the two
// runtime helpers of libtkmi that the transform and GEMM sources call.
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../tensorkit.jl_amd/csrc/tkmi_internal.h"

thread_local emu_index threadIdx;
emu_index blockIdx, blockDim, gridDim;

namespace {

// A barrier whose party shrinks when a thread leaves the kernel (a thread that returns
// early no longer takes part in later barriers).
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, count = 0;
    std::atomic<uint64_t> gen{0};
    void reset(int n) {
        std::lock_guard<std::mutex> l(m);
        expected = n;
        count = 0;
    }
    void release_locked() {
        count = 0;
        gen.fetch_add(1, std::memory_order_release);
        cv.notify_all();
    }
    void wait() {
        uint64_t g;
        {
            std::lock_guard<std::mutex> l(m);
            g = gen.load(std::memory_order_relaxed);
            if (++count >= expected) {
                release_locked();
                return;
            }
        }
        // yield a while before sleeping: most rendezvous complete within a few time slices
        for (int spin = 0; spin < 64; ++spin) {
            if (gen.load(std::memory_order_acquire) != g) return;
            std::this_thread::yield();
        }
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [&] { return gen.load(std::memory_order_relaxed) != g; });
    }
    void drop() {
        std::lock_guard<std::mutex> l(m);
        --expected;
        if (count > 0 && count >= expected) release_locked();
    }
};

constexpr int MAXT = 1024, WAVE = 64;
Barrier g_block;
Barrier g_wave[MAXT / WAVE];
double g_area[MAXT / WAVE][2][WAVE * 8];

struct Pool {
    std::vector<std::thread> threads;
    std::mutex m;
    std::condition_variable cv_start, cv_done;
    std::atomic<uint64_t> job{0};
    int active = 0, done = 0;
    bool quit = false;
    const std::function<void()> *fn = nullptr;

    Pool() {
        for (int t = 0; t < MAXT; ++t) threads.emplace_back([this, t] { run(t); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> l(m);
            quit = true;
        }
        cv_start.notify_all();
        for (auto &th : threads) th.join();
    }
    void run(int t) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void()> *f;
            int n;
            for (int spin = 0; spin < 64 && job.load(std::memory_order_acquire) == seen; ++spin)
                std::this_thread::yield();
            {
                std::unique_lock<std::mutex> l(m);
                cv_start.wait(l, [&] { return quit || job.load(std::memory_order_relaxed) != seen; });
                if (quit) return;
                seen = job;
                f = fn;
                n = active;
            }
            if (t < n) {
                threadIdx = emu_index{(unsigned)t, 0, 0};
                (*f)();
                g_wave[t / WAVE].drop();
                g_block.drop();
                std::lock_guard<std::mutex> l(m);
                if (++done == n) cv_done.notify_one();
            }
        }
    }
    void block(int n, const std::function<void()> &f) {
        g_block.reset(n);
        for (int w = 0; w < MAXT / WAVE; ++w) g_wave[w].reset(std::max(0, std::min(WAVE, n - w * WAVE)));
        std::unique_lock<std::mutex> l(m);
        fn = &f;
        active = n;
        done = 0;
        ++job;
        cv_start.notify_all();
        cv_done.wait(l, [&] { return done == n; });
    }
};

Pool &pool() {
    static Pool p;
    return p;
}

}  // namespace

namespace emu {

void launch(dim3 grid, dim3 block, const std::function<void()> &fn) {
    if (block.x == 0 || block.x > (unsigned)MAXT || block.y != 1 || block.z != 1) {
        std::fprintf(stderr, "emu: unsupported block shape %u×%u×%u\n", block.x, block.y, block.z);
        std::abort();
    }
    gridDim = emu_index{grid.x, grid.y, grid.z};
    blockDim = emu_index{block.x, 1, 1};
    for (unsigned b = 0; b < grid.x; ++b) {
        blockIdx = emu_index{b, 0, 0};
        pool().block((int)block.x, fn);
    }
}

void syncthreads() { g_block.wait(); }
void wave_sync() { g_wave[threadIdx.x / WAVE].wait(); }
double *wave_area(int which) { return g_area[threadIdx.x / WAVE][which]; }

}  // namespace emu

namespace tkmi {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

int upload(void **dptr, const void *host, size_t bytes) {
    *dptr = nullptr;
    if (bytes == 0) bytes = 16;
    TKMI_HIP(hipMalloc(dptr, bytes));
    if (host) std::memcpy(*dptr, host, bytes);
    return TKMI_OK;
}

const std::string &last_error() { return g_last_error; }

}  // namespace tkmi

extern "C" {

const char *tkmi_last_error(void) { return tkmi::last_error().c_str(); }

int tkmi_alloc(void **ptr, size_t bytes) {
    *ptr = std::malloc(bytes ? bytes : 16);
    return *ptr ? TKMI_OK : TKMI_ERR_ALLOC;
}

int tkmi_free(void *ptr) {
    std::free(ptr);
    return TKMI_OK;
}

}  // extern "C"
