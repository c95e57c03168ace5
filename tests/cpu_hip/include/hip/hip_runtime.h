// CPU execution shim for the HIP kernels of grassmanntn_amd -- TEST INFRASTRUCTURE ONLY.
//
// Compiling the library's .hip sources with this directory first on the include path turns them
// into plain C++: every GPU thread is an OS thread, a workgroup's threads run together (blocks run
// one after another), __syncthreads is a barrier over the live threads of the block, wave-level
// operations (__shfl_down, MFMA) exchange data through a per-wave buffer, __shared__ variables are
// function statics (one copy, reused block after block) and dynamic LDS is a buffer whose bytes
// beyond the requested size are poisoned for AddressSanitizer.  Built with -fsanitize=address,
// every out-of-bounds access to global memory or LDS is reported on the CPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include <sanitizer/asan_interface.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __shared__ static
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu::dyn_lds();

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct double2 {
    double x, y;
};
template <typename T, unsigned N>
struct HIP_vector_type;

typedef int hipError_t;
typedef void *hipStream_t;
#define hipSuccess 0
#define hipFuncAttributeMaxDynamicSharedMemorySize 8
struct hipDeviceProp_t {
    char gcnArchName[256];
    int multiProcessorCount;
};

namespace emu {
struct Idx {
    unsigned x, y, z;
};
extern thread_local Idx tid, bid;
extern Idx bdim, gdim;
void *dyn_lds();
void sync_block();
void sync_wave();
void *wave_buf(int slot);  // per-wave exchange area (64 lanes x 512 bytes per slot)
int lane();
}  // namespace emu

#define threadIdx emu::tid
#define blockIdx emu::bid
#define blockDim emu::bdim
#define gridDim emu::gdim

inline void __syncthreads() { emu::sync_block(); }
// code-motion barrier among the lanes of one wavefront (on the CPU: a wave-wide rendezvous)
inline void __builtin_amdgcn_wave_barrier() { emu::sync_wave(); }

template <typename T>
inline T __shfl_down(T v, unsigned delta, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of a value wider than 8 bytes");
    T *buf = (T *)emu::wave_buf(0);
    int l = emu::lane();
    emu::sync_wave();
    buf[l] = v;
    emu::sync_wave();
    T r = (l + (int)delta < width) ? buf[l + delta] : v;
    emu::sync_wave();
    return r;
}

// butterfly exchange within segments of `width` lanes (width a power of two, mask < width)
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of a value wider than 8 bytes");
    T *buf = (T *)emu::wave_buf(0);
    int l = emu::lane();
    emu::sync_wave();
    buf[l] = v;
    emu::sync_wave();
    int src = l ^ mask;
    T r = ((src / width) == (l / width)) ? buf[src] : v;
    emu::sync_wave();
    return r;
}

// value of v in lane l (l uniform across the wavefront)
inline int __builtin_amdgcn_readlane(int v, int l) {
    int *buf = (int *)emu::wave_buf(0);
    int me = emu::lane();
    emu::sync_wave();
    buf[me] = v;
    emu::sync_wave();
    int r = buf[l];
    emu::sync_wave();
    return r;
}

// value of v in the wavefront's first lane (every lane is active in the emulation)
inline int __builtin_amdgcn_readfirstlane(int v) { return __builtin_amdgcn_readlane(v, 0); }

// raw buffer resource (stride 0): loads at byte offset voffset + soffset from the base; an offset
// at or beyond num_records reads zero, as the hardware's range check does
struct emu_buffer_rsrc {
    const char *base;
    unsigned num_records;
};
typedef emu_buffer_rsrc __amdgpu_buffer_rsrc_t;
inline emu_buffer_rsrc __builtin_amdgcn_make_buffer_rsrc(void *p, short, int num_records, int) {
    return emu_buffer_rsrc{(const char *)p, (unsigned)num_records};
}
template <typename T>
inline T emu_buffer_load(emu_buffer_rsrc r, int voffset, int soffset) {
    const unsigned o = (unsigned)voffset + (unsigned)soffset;
    T v = 0;
    if ((unsigned long long)o + sizeof(T) <= r.num_records) memcpy(&v, r.base + o, sizeof(T));
    return v;
}
inline unsigned long long __builtin_amdgcn_raw_buffer_load_b64(emu_buffer_rsrc r, int vo, int so, int) {
    return emu_buffer_load<unsigned long long>(r, vo, so);
}
inline unsigned __builtin_amdgcn_raw_buffer_load_b32(emu_buffer_rsrc r, int vo, int so, int) {
    return emu_buffer_load<unsigned>(r, vo, so);
}

// value of v in lane src (src may differ per lane; lanes outside [0, width) read their own value)
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of a value wider than 8 bytes");
    T *buf = (T *)emu::wave_buf(0);
    int l = emu::lane();
    emu::sync_wave();
    buf[l] = v;
    emu::sync_wave();
    int base = (l / width) * width;
    T r = (src >= 0 && src < width) ? buf[base + src] : v;
    emu::sync_wave();
    return r;
}

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

// ---- MFMA emulation (wave64; operand/result layouts of the 16x16 instructions) -------------------
typedef double emu_d4 __attribute__((ext_vector_type(4)));
typedef float emu_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 emu_bf8 __attribute__((ext_vector_type(8)));

template <typename TA, typename TC, int KPL, bool ROW_INTERLEAVED = false>
inline TC emu_mfma16(const TA *a, const TA *b, TC c) {
    // lane l supplies A[l%16][KPL*(l/16) + t] and B[KPL*(l/16) + t][l%16], t < KPL;
    // it holds C[4*(l/16) + v][l%16], v < 4 (FP32 / BF16 forms: 4 consecutive rows per lane), or
    // C[l/16 + 4*v][l%16] for the FP64 16x16x4 form (ROW_INTERLEAVED: one row per register group;
    // composable_kernel's mfma_type<mfma_f64_16x16x4f64> has group_size 1, num_groups_per_blk 4)
    struct Slot {
        double a[16], b[16], c[4];
    };
    Slot *buf = (Slot *)emu::wave_buf(1);
    int l = emu::lane();
    emu::sync_wave();
    for (int t = 0; t < KPL; ++t) {
        buf[l].a[t] = (double)a[t];
        buf[l].b[t] = (double)b[t];
    }
    for (int v = 0; v < 4; ++v) buf[l].c[v] = (double)c[v];
    emu::sync_wave();
    TC out;
    for (int v = 0; v < 4; ++v) {
        int row = ROW_INTERLEAVED ? (l / 16) + 4 * v : 4 * (l / 16) + v, col = l % 16;
        double acc = buf[l].c[v];
        for (int kk = 0; kk < 4 * KPL; ++kk) {
            int qa = kk / KPL, ta = kk % KPL;
            double av = buf[row + 16 * qa].a[ta];
            double bv = buf[col + 16 * qa].b[ta];
            acc += av * bv;
        }
        out[v] = acc;
    }
    emu::sync_wave();
    return out;
}
inline emu_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, emu_d4 c, int, int, int) {
    return emu_mfma16<double, emu_d4, 1, true>(&a, &b, c);
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, emu_f4 c, int, int, int) {
    return emu_mfma16<float, emu_f4, 1>(&a, &b, c);
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_bf8 a, emu_bf8 b, emu_f4 c, int, int, int) {
    __bf16 aa[8], bb[8];
    for (int t = 0; t < 8; ++t) {
        aa[t] = a[t];
        bb[t] = b[t];
    }
    return emu_mfma16<__bf16, emu_f4, 8>(aa, bb, c);
}

typedef int emu_i4 __attribute__((ext_vector_type(4)));
// int8 16x16x64: each lane's 16 operand bytes are packed four to an int (little-endian)
inline emu_i4 __builtin_amdgcn_mfma_i32_16x16x64_i8(emu_i4 a, emu_i4 b, emu_i4 c, int, int, int) {
    signed char aa[16], bb[16];
    for (int t = 0; t < 16; ++t) {
        aa[t] = (signed char)((a[t / 4] >> (8 * (t % 4))) & 0xff);
        bb[t] = (signed char)((b[t / 4] >> (8 * (t % 4))) & 0xff);
    }
    return emu_mfma16<signed char, emu_i4, 16>(aa, bb, c);
}

// ---- runtime API ----------------------------------------------------------------------------
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes);
template <typename T>
inline hipError_t hipMalloc(T **p, size_t n) {
    return hipMalloc((void **)p, n);
}
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipFuncSetAttribute(const void *f, int attr, int value);
#define hipDeviceAttributeMultiprocessorCount 63
hipError_t hipDeviceGetAttribute(int *v, int attr, int dev);
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
typedef void *hipEvent_t;
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
#define hipEventDisableTiming 2
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
// streams run their work at once; graph capture is not emulated (hipErrorNotSupported)
#define hipStreamNonBlocking 0x01
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
typedef void *hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipErrorNotSupported 801
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive, hipStreamCaptureStatusInvalidated };
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *st) {
    *st = hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *n, char *log, size_t len);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
// a __device__ variable is a host variable here: its address is the symbol's address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
    *p = const_cast<void *>(sym);
    return hipSuccess;
}
// device-side integer min / max of HIP
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

namespace emu {
void run_grid(dim3 grid, dim3 block, size_t shm, void (*body)(void *), void *ctx);
}

template <typename... KArgs, typename... Args>
inline void hipLaunchKernelGGL(void (*k)(KArgs...), dim3 grid, dim3 block, size_t shm, hipStream_t, Args... args) {
    struct Ctx {
        void (*k)(KArgs...);
        std::tuple<KArgs...> a;
    } ctx{k, std::tuple<KArgs...>((KArgs)args...)};
    emu::run_grid(grid, block, shm, [](void *p) {
        Ctx *c = (Ctx *)p;
        std::apply(c->k, c->a);
    }, &ctx);
}
