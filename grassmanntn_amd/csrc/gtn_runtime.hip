// Device memory, copies, fills, events, streams, HIP graphs and synchronisation for the Python host
// layer.
//
// The library owns its HBM through a small caching allocator (exact size classes rounded to
// 256 bytes, reused in stream order on the one stream the host layer uses), so that processes using
// grassmanntn_amd need no other GPU framework and run on the HIP runtime this library was built
// against.
//
// Graph capture and memory pools: while a capture is open (gtn_graph_begin .. gtn_graph_end),
// allocations come from the capture's pool.  A pool block freed at any time returns to that pool
// only, so the addresses a captured graph holds are never handed to work outside it; blocks freed
// inside the capture are reused by later allocations of the same capture in stream order, which is
// also the order of every replay.  gtn_pool_release hands a pool's blocks back to the general cache
// once its graphs are destroyed.
#include "gtn_common.h"
#include <atomic>
#include <map>
#include <mutex>
#include <string.h>
#include <vector>

namespace gtn {

static std::mutex g_mu;
static std::map<int64_t, std::vector<void *>> g_free;  // size class -> cached blocks
struct LiveBlock {
    int64_t cls;
    int pool;
};
static std::map<void *, LiveBlock> g_live;             // pointer -> size class, pool (0: general)
static int64_t g_cached_bytes = 0;
static const int64_t kCacheCap = 64ll << 30;          // release cached blocks beyond 64 GiB
struct Pool {
    std::map<int64_t, std::vector<void *>> free;
    bool released = false;
};
static std::map<int, Pool> g_pools;  // pool id (> 0) -> its free blocks
static int g_cur_pool = 0;           // pool of the open capture, 0 outside captures
static std::atomic<int> g_graphs{0};

// Cached blocks beyond kCacheCap are released (hipFree) from the largest size classes first, never
// from class keep_cls (the one being returned: it is the likeliest to be asked for again), and never
// while a capture is open (hipFree synchronises the device).  Caller holds g_mu.
static void trim_cache(int64_t keep_cls) {
    if (g_cur_pool != 0) return;
    while (g_cached_bytes > kCacheCap) {
        auto victim = g_free.end();
        for (auto it = g_free.begin(); it != g_free.end(); ++it)
            if (it->first != keep_cls && !it->second.empty()) victim = it;  // map order: the last is the largest
        if (victim == g_free.end()) break;
        (void)hipFree(victim->second.back());
        victim->second.pop_back();
        g_cached_bytes -= victim->first;
    }
}

int graphs_alive() { return g_graphs.load(); }
int capture_open() {
    std::lock_guard<std::mutex> l(g_mu);
    return g_cur_pool != 0;
}

template <typename T>
__global__ void k_fill_value(T *p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

template <typename T>
__global__ void k_copy_2d(const T *__restrict__ src, int64_t lds, T *__restrict__ dst, int64_t ldd, int64_t rows,
                          int64_t cols) {
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        dst[r * ldd + c] = src[r * lds + c];
    }
}

}  // namespace gtn

using namespace gtn;

extern "C" {

int gtn_malloc(int64_t bytes, void **out) {
    GTN_REQUIRE(out && bytes >= 0, "gtn_malloc: bad argument");
    int64_t cls = (bytes + 255) / 256 * 256;
    if (cls == 0) cls = 256;
    int pool;
    {
        std::lock_guard<std::mutex> l(g_mu);
        pool = g_cur_pool;
        auto &fl = pool ? g_pools[pool].free : g_free;
        auto it = fl.find(cls);
        if (it != fl.end() && !it->second.empty()) {
            *out = it->second.back();
            it->second.pop_back();
            if (!pool) g_cached_bytes -= cls;
            g_live[*out] = LiveBlock{cls, pool};
            return GTN_OK;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)cls);
    if (e != hipSuccess && !pool) {
        // give the cache back to the device and retry once
        std::lock_guard<std::mutex> l(g_mu);
        for (auto &kv : g_free)
            for (void *q : kv.second) (void)hipFree(q);
        g_free.clear();
        g_cached_bytes = 0;
        e = hipMalloc(&p, (size_t)cls);
    }
    if (e != hipSuccess) {
        set_error("gtn_malloc: hipMalloc(%lld) failed: %s", (long long)cls, hipGetErrorString(e));
        return GTN_ENOMEM;
    }
    std::lock_guard<std::mutex> l(g_mu);
    g_live[p] = LiveBlock{cls, pool};
    *out = p;
    return GTN_OK;
}

int gtn_free(void *p) {
    if (!p) return GTN_OK;
    std::lock_guard<std::mutex> l(g_mu);
    auto it = g_live.find(p);
    GTN_REQUIRE(it != g_live.end(), "gtn_free: pointer not allocated by gtn_malloc");
    int64_t cls = it->second.cls;
    int pool = it->second.pool;
    g_live.erase(it);
    if (pool) {
        auto pit = g_pools.find(pool);
        if (pit != g_pools.end() && !pit->second.released) {
            pit->second.free[cls].push_back(p);
            return GTN_OK;
        }
    }
    // the general cache (graphs hold pool blocks only); over the cap, other size classes make room
    g_free[cls].push_back(p);
    g_cached_bytes += cls;
    if (g_cached_bytes > kCacheCap) trim_cache(cls);
    return GTN_OK;
}

int gtn_mem_info(int64_t *live_bytes, int64_t *cached_bytes, int64_t *device_free, int64_t *device_total) {
    int64_t live = 0, cached = 0;
    {
        std::lock_guard<std::mutex> l(g_mu);
        for (auto &kv : g_live) live += kv.second.cls;
        cached = g_cached_bytes;
        for (auto &pv : g_pools)
            for (auto &kv : pv.second.free) cached += kv.first * (int64_t)kv.second.size();
    }
    size_t f = 0, t = 0;
    GTN_CHECK_HIP(hipMemGetInfo(&f, &t));
    if (live_bytes) *live_bytes = live;
    if (cached_bytes) *cached_bytes = cached;
    if (device_free) *device_free = (int64_t)f;
    if (device_total) *device_total = (int64_t)t;
    return GTN_OK;
}

// kind: 1 host->device, 2 device->host, 3 device->device.  Host<->device copies are synchronous;
// device->device copies are ordered on `stream`.
int gtn_memcpy(void *dst, const void *src, int64_t bytes, int kind, void *stream) {
    if (bytes == 0) return GTN_OK;
    GTN_REQUIRE(dst && src && bytes > 0, "gtn_memcpy: bad argument");
    hipStream_t s = as_stream(stream);
    switch (kind) {
        case 1:
            GTN_CHECK_HIP(hipStreamSynchronize(s));
            GTN_CHECK_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
            break;
        case 2:
            GTN_CHECK_HIP(hipStreamSynchronize(s));
            GTN_CHECK_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
            break;
        case 3: GTN_CHECK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, s)); break;
        default: set_error("gtn_memcpy: unknown kind %d", kind); return GTN_EINVAL;
    }
    return GTN_OK;
}

int gtn_fill(void *p, int64_t n, int dtype, double value, void *stream) {
    if (n == 0) return GTN_OK;
    GTN_REQUIRE(p && n > 0, "gtn_fill: bad argument");
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(n, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_fill_value<double>, g, b, 0, s, (double *)p, n, value); break;
        case GTN_F32: hipLaunchKernelGGL(k_fill_value<float>, g, b, 0, s, (float *)p, n, (float)value); break;
        case GTN_C128:
            GTN_REQUIRE(value == 0.0, "gtn_fill: complex tensors can only be zero-filled");
            hipLaunchKernelGGL(k_fill_value<double>, g, b, 0, s, (double *)p, 2 * n, 0.0);
            break;
        case GTN_I8: hipLaunchKernelGGL(k_fill_value<int8_t>, g, b, 0, s, (int8_t *)p, n, (int8_t)value); break;
        default: set_error("gtn_fill: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_copy2d(void *dst, int64_t ldd, const void *src, int64_t lds, int64_t rows, int64_t cols, int dtype, void *stream) {
    if (rows <= 0 || cols <= 0) return GTN_OK;
    GTN_REQUIRE(dst && src && ldd >= cols && lds >= cols, "gtn_copy2d: bad argument");
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(rows * cols, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_copy_2d<double>, g, b, 0, s, (const double *)src, lds, (double *)dst, ldd, rows, cols); break;
        case GTN_F32: hipLaunchKernelGGL(k_copy_2d<float>, g, b, 0, s, (const float *)src, lds, (float *)dst, ldd, rows, cols); break;
        case GTN_C128:
            hipLaunchKernelGGL(k_copy_2d<double2>, g, b, 0, s, (const double2 *)src, lds, (double2 *)dst, ldd, rows, cols);
            break;
        default: set_error("gtn_copy2d: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_synchronize(void) {
    GTN_CHECK_HIP(hipDeviceSynchronize());
    return GTN_OK;
}

int gtn_stream_create(void **stream) {
    GTN_REQUIRE(stream, "gtn_stream_create: null argument");
    hipStream_t s;
    GTN_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return GTN_OK;
}

int gtn_stream_destroy(void *stream) {
    if (stream) GTN_CHECK_HIP(hipStreamDestroy(as_stream(stream)));
    return GTN_OK;
}

int gtn_stream_synchronize(void *stream) {
    GTN_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
    return GTN_OK;
}

int gtn_graph_begin(void *stream, int pool) {
    GTN_REQUIRE(stream, "gtn_graph_begin: capture needs a created stream, not the null stream");
    GTN_REQUIRE(pool > 0, "gtn_graph_begin: pool ids are positive");
    {
        std::lock_guard<std::mutex> l(g_mu);
        GTN_REQUIRE(g_cur_pool == 0, "gtn_graph_begin: a capture is already open");
        GTN_REQUIRE(!g_pools[pool].released, "gtn_graph_begin: pool %d was released", pool);
        g_cur_pool = pool;
    }
    hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> l(g_mu);
        g_cur_pool = 0;
        set_error("gtn_graph_begin: hipStreamBeginCapture failed: %s", hipGetErrorString(e));
        return GTN_EDEVICE;
    }
    return GTN_OK;
}

int gtn_graph_end(void *stream, void **graph_exec) {
    GTN_REQUIRE(stream, "gtn_graph_end: null stream");
    {
        std::lock_guard<std::mutex> l(g_mu);
        g_cur_pool = 0;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        set_error("gtn_graph_end: hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return GTN_EDEVICE;
    }
    if (!graph_exec) {  // capture abandoned
        (void)hipGraphDestroy(g);
        return GTN_OK;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error("gtn_graph_end: hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return GTN_EDEVICE;
    }
    ++g_graphs;
    *graph_exec = (void *)x;
    return GTN_OK;
}

int gtn_graph_launch(void *graph_exec, void *stream) {
    GTN_REQUIRE(graph_exec && stream, "gtn_graph_launch: null argument");
    GTN_CHECK_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream)));
    return GTN_OK;
}

int gtn_graph_destroy(void *graph_exec) {
    if (!graph_exec) return GTN_OK;
    GTN_CHECK_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    --g_graphs;
    return GTN_OK;
}

int gtn_pool_release(int pool) {
    GTN_REQUIRE(pool > 0, "gtn_pool_release: pool ids are positive");
    std::lock_guard<std::mutex> l(g_mu);
    GTN_REQUIRE(g_cur_pool != pool, "gtn_pool_release: pool %d is in use by an open capture", pool);
    Pool &P = g_pools[pool];
    for (auto &kv : P.free)
        for (void *q : kv.second) {
            g_free[kv.first].push_back(q);
            g_cached_bytes += kv.first;
        }
    P.free.clear();
    P.released = true;
    trim_cache(-1);
    return GTN_OK;
}

int gtn_event_create(void **ev) {
    hipEvent_t e;
    GTN_CHECK_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return GTN_OK;
}

int gtn_event_record(void *ev, void *stream) {
    GTN_CHECK_HIP(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
    return GTN_OK;
}

int gtn_event_elapsed_ms(void *start, void *stop, float *ms) {
    GTN_CHECK_HIP(hipEventSynchronize((hipEvent_t)stop));
    GTN_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GTN_OK;
}

int gtn_event_destroy(void *ev) {
    GTN_CHECK_HIP(hipEventDestroy((hipEvent_t)ev));
    return GTN_OK;
}

}  // extern "C"
