// CPU execution runtime behind tests/cpu_hip/include/hip/hip_runtime.h -- TEST INFRASTRUCTURE ONLY.
#include <hip/hip_runtime.h>

#include <memory>

namespace emu {
thread_local Idx tid, bid;
Idx bdim, gdim;
static thread_local int t_lane = 0, t_wave = 0;

// barrier over the live threads of a group; a thread that leaves the kernel drops out
struct Bar {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, count = 0;
    long gen = 0;
    void reset(int n) {
        std::lock_guard<std::mutex> l(m);
        expected = n;
        count = 0;
    }
    void wait() {
        std::unique_lock<std::mutex> l(m);
        long g = gen;
        if (++count == expected) {
            count = 0;
            ++gen;
            cv.notify_all();
        } else {
            cv.wait(l, [&] { return gen != g; });
        }
    }
    void drop() {
        std::unique_lock<std::mutex> l(m);
        --expected;
        if (expected > 0 && count == expected) {
            count = 0;
            ++gen;
            cv.notify_all();
        }
    }
};

static Bar g_block;
static std::unique_ptr<Bar[]> g_waves;
static std::vector<std::vector<char>> g_wavebuf;  // per wave: 2 slots
static char *g_dyn = nullptr;
static size_t g_dyn_cap = 0, g_dyn_size = 0;
static const size_t kSlot = 64 * 512;  // per lane: up to 512 bytes (the MFMA emulation's operand slots)

void *dyn_lds() { return g_dyn; }
void sync_block() { g_block.wait(); }
void sync_wave() { g_waves[t_wave].wait(); }
void *wave_buf(int slot) { return g_wavebuf[t_wave].data() + slot * kSlot; }
int lane() { return t_lane; }

void run_grid(dim3 grid, dim3 block, size_t shm, void (*body)(void *), void *ctx) {
    const unsigned nt = block.x * block.y * block.z;
    const unsigned nw = (nt + 63) / 64;
    bdim = {block.x, block.y, block.z};
    gdim = {grid.x, grid.y, grid.z};
    g_waves.reset(new Bar[nw]);
    g_wavebuf.assign(nw, std::vector<char>(2 * kSlot));
    // dynamic LDS: exactly `shm` usable bytes, the rest of the buffer poisoned
    size_t need = shm + 4096;
    if (need > g_dyn_cap) {
        if (g_dyn) {
            ASAN_UNPOISON_MEMORY_REGION(g_dyn, g_dyn_cap);
            free(g_dyn);
        }
        g_dyn = (char *)aligned_alloc(256, (need + 255) / 256 * 256);
        g_dyn_cap = (need + 255) / 256 * 256;
    }
    ASAN_UNPOISON_MEMORY_REGION(g_dyn, g_dyn_cap);
    ASAN_POISON_MEMORY_REGION(g_dyn + shm, g_dyn_cap - shm);
    g_dyn_size = shm;
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                g_block.reset((int)nt);
                for (unsigned w = 0; w < nw; ++w) g_waves[w].reset((int)std::min(64u, nt - 64 * w));
                std::vector<std::thread> ts;
                ts.reserve(nt);
                for (unsigned t = 0; t < nt; ++t) {
                    ts.emplace_back([=] {
                        tid = {t % block.x, (t / block.x) % block.y, t / (block.x * block.y)};
                        bid = {bx, by, bz};
                        t_lane = (int)(t % 64);
                        t_wave = (int)(t / 64);
                        body(ctx);
                        g_waves[t_wave].drop();
                        g_block.drop();
                    });
                }
                for (auto &th : ts) th.join();
            }
}
}  // namespace emu

hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated"; }
hipError_t hipMalloc(void **p, size_t n) {
    *p = malloc(n);
    return hipSuccess;
}
// a nominal 64 GiB device, all of it free (the emulation does not track host allocations)
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes) {
    *free_bytes = *total_bytes = (size_t)64 << 30;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950:emulated");
    p->multiProcessorCount = 256;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *, int attr, int value) {
    if (attr == hipFuncAttributeMaxDynamicSharedMemorySize && value > 160 * 1024) return 1;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, int, int) {
    *v = 256;
    return hipSuccess;
}
hipError_t hipFree(void *p) {
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) {
    *e = (hipEvent_t)1;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 0.0f;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = (hipStream_t)2;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
