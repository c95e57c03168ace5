// Internal helpers shared by the HIP translation units of libgtn_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include "../../include/gtn_hip.h"

namespace gtn {

// Thread-local message behind gtn_last_error().
void set_error(const char *fmt, ...);

#define GTN_CHECK_HIP(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            gtn::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            return GTN_EDEVICE;                                                          \
        }                                                                                \
    } while (0)

#define GTN_REQUIRE(cond, ...)            \
    do {                                  \
        if (!(cond)) {                    \
            gtn::set_error(__VA_ARGS__);  \
            return GTN_EINVAL;            \
        }                                 \
    } while (0)

// Launch-error check after a kernel launch.
#define GTN_CHECK_LAUNCH() GTN_CHECK_HIP(hipGetLastError())

__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }
// eps(i) = i ^ (popcount(i >> 1) & 1): canonical <-> parity-preserving (an involution)
__device__ __forceinline__ uint64_t enc_switch(uint64_t i) { return i ^ (uint64_t)(popc64(i >> 1) & 1u); }
// sigma(i) = -1 iff popcount(i) mod 4 is 2 or 3; returned as a sign bit (1 => negative)
__device__ __forceinline__ uint32_t sigma_bit(uint64_t i) { return (popc64(i) >> 1) & 1u; }

// dry-run validation (host only): registered regions and a range check
bool dry_run();
int check_range(const void *p, int64_t bytes, const char *what);
#define GTN_RANGE(p, bytes, what)                         \
    do {                                                  \
        int rc_ = gtn::check_range((p), (bytes), (what)); \
        if (rc_) return rc_;                              \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// graph bookkeeping of the runtime (gtn_runtime.hip): instantiated graphs that may still be
// launched, and whether a capture is open.  Library-owned device buffers that a graph may hold are
// never freed or moved while either is nonzero.
int graphs_alive();
int capture_open();

inline int grid_for(int64_t n, int block, int max_blocks = 65536) {
    int64_t g = (n + block - 1) / block;
    if (g > max_blocks) g = max_blocks;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace gtn
