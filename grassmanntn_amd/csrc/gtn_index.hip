// Integer index maps and sign factors on the device (param.py:61-83; __init__.py:1011-1154,
// 1944-2126, 2947-3170), and the FP64 reductions / elementwise helpers of the TRG step.
#include "gtn_common.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <utility>
#include <map>
#include <mutex>
#include <string>

namespace gtn {

static thread_local char g_err[1024] = "";
static std::vector<std::pair<uint64_t, uint64_t>> g_regions;
static bool g_dry = false;
bool dry_run() { return g_dry; }
int check_range(const void *p, int64_t bytes, const char *what) {
    if (!g_dry || bytes <= 0) return GTN_OK;
    uint64_t lo = (uint64_t)p, hi = lo + (uint64_t)bytes;
    for (auto &r : g_regions)
        if (r.first <= lo && hi <= r.first + r.second) return GTN_OK;
    set_error("dry run: %s touches [%#llx, %#llx) outside every registered region", what, (unsigned long long)lo,
              (unsigned long long)hi);
    return GTN_EINVAL;
}
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------------------------
// index tables
// ---------------------------------------------------------------------------------------------
__global__ void k_index_tables(int64_t n, int32_t *popc, int8_t *sigma, int32_t *enc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t u = (uint64_t)i;
        if (popc) popc[i] = (int32_t)popc64(u);
        if (sigma) sigma[i] = sigma_bit(u) ? (int8_t)-1 : (int8_t)1;
        if (enc) enc[i] = (int32_t)enc_switch(u);
    }
}

// ---------------------------------------------------------------------------------------------
// signed gather.  The descriptor lives in kernel-argument memory, so every lane reads the same
// words.  Division by an axis dimension is a shift when the dimension is a power of two (all
// fermionic legs are) and a true division otherwise.
// ---------------------------------------------------------------------------------------------
constexpr int AX = GTN_PERM_MAX_AX, NC = GTN_PERM_MAX_COMP;
struct PermArgs {
    int32_t n_out, n_virt, n_comp, n_in;
    int32_t virt_enc_mask, virt_sigma_mask, comp_sigma_mask, in_enc_mask, in_sigma_mask, in_split_mask;
    uint32_t out_dim[AX];
    int8_t out_shift[AX];
    int8_t out_virt[AX][2];
    int64_t out_vmul[AX][2];
    int8_t virt_ncomp[AX];
    uint32_t comp_dim[NC];
    int8_t comp_shift[NC];
    int8_t comp_bit[NC];
    int8_t comp_in[NC][2];
    int64_t comp_mult[NC][2];
    int64_t in_stride[AX];
    int64_t in_stride_lo[AX];
    uint32_t table[2 * (1 << GTN_PERM_MAX_BITS) / 32];
    double alpha_re, alpha_im;
    int64_t total;
};

// Contribution of the virtual axes in `vmask` for virtual-axis values V: their components' parity
// bits and sign flips, and the offset and sign flips of the input axes in `inmask`.  With every
// axis in the masks this is the whole map; a subset is exact when the inputs in `inmask` are fed
// only by components of the virtual axes in `vmask` (inputs fed by nobody hold index 0, whose
// offset, encoding and sigma are all trivial).
__device__ __forceinline__ void perm_core(const PermArgs &a, const uint64_t (&V)[AX], uint32_t vmask,
                                          uint32_t inmask, int64_t &off, uint32_t &bits, uint32_t &neg) {
    uint64_t F[AX];
#pragma unroll
    for (int b = 0; b < AX; ++b) F[b] = 0;
    bits = 0;
    neg = 0;
    int j = a.n_comp;
    for (int v = a.n_virt - 1; v >= 0; --v) {
        if (!((vmask >> v) & 1u)) {
            j -= a.virt_ncomp[v];
            continue;
        }
        uint64_t c = ((a.virt_enc_mask >> v) & 1) ? enc_switch(V[v]) : V[v];
        if ((a.virt_sigma_mask >> v) & 1) neg ^= sigma_bit(c);
        for (int t = 0; t < a.virt_ncomp[v]; ++t) {
            --j;
            uint64_t d = a.comp_dim[j], k;
            if (a.comp_shift[j] >= 0) {
                k = c & (d - 1);
                c >>= a.comp_shift[j];
            } else {
                k = c % d;
                c /= d;
            }
            uint32_t pc = popc64(k);
            if (a.comp_bit[j] >= 0) bits ^= (pc & 1u) << a.comp_bit[j];
            if ((a.comp_sigma_mask >> j) & 1) neg ^= (pc >> 1) & 1u;
            F[a.comp_in[j][0]] += k * (uint64_t)a.comp_mult[j][0];
            if (a.comp_in[j][1] >= 0) F[a.comp_in[j][1]] += k * (uint64_t)a.comp_mult[j][1];
        }
    }
    off = 0;
    for (int b = 0; b < a.n_in; ++b) {
        if (!((inmask >> b) & 1u)) continue;
        uint64_t f = F[b];
        if ((a.in_sigma_mask >> b) & 1) neg ^= sigma_bit(f);
        uint64_t e = ((a.in_enc_mask >> b) & 1) ? enc_switch(f) : f;
        if ((a.in_split_mask >> b) & 1)
            off += (int64_t)(e & 1) * a.in_stride_lo[b] + (int64_t)(e >> 1) * a.in_stride[b];
        else
            off += (int64_t)e * a.in_stride[b];
    }
}

__device__ __forceinline__ uint32_t sign_code(const PermArgs &a, uint32_t bits, uint32_t neg) {
    uint32_t t = (a.table[bits >> 4] >> ((bits & 15u) * 2)) & 3u;
    return (t == 2u) ? 2u : (t ^ neg);
}

// Returns the source offset of output element `lin` and its table entry (0: +1, 1: -1, 2: zero).
__device__ __forceinline__ int64_t perm_map(const PermArgs &a, uint64_t lin, uint32_t &code) {
    uint64_t V[AX];
#pragma unroll
    for (int b = 0; b < AX; ++b) V[b] = 0;
    uint64_t rest = lin;
    for (int ax = a.n_out - 1; ax >= 0; --ax) {
        uint64_t d = a.out_dim[ax], o;
        if (a.out_shift[ax] >= 0) {
            o = rest & (d - 1);
            rest >>= a.out_shift[ax];
        } else {
            o = rest % d;
            rest /= d;
        }
        if (a.out_virt[ax][0] >= 0) V[a.out_virt[ax][0]] += o * (uint64_t)a.out_vmul[ax][0];
        if (a.out_virt[ax][1] >= 0) V[a.out_virt[ax][1]] += o * (uint64_t)a.out_vmul[ax][1];
    }
    int64_t off;
    uint32_t bits, neg;
    perm_core(a, V, 0xffffffffu, 0xffffffffu, off, bits, neg);
    code = sign_code(a, bits, neg);
    return off;
}

// ---------------------------------------------------------------------------------------------
// Table-driven map.  The virtual axes fall into groups: two virtual axes share a group when their
// components feed a common input axis.  A group's contribution (source offset, parity bits, sign
// flip) depends only on the values of its own virtual axes, so it is tabulated once per gather
// shape, indexed by G = sum over the group's axes of V_v * radix_v, itself a linear function of the
// output indices.  Per element the map is then an index decomposition, at most kTabGroups table
// reads, a sum and an XOR, instead of the component walk of perm_map.
// ---------------------------------------------------------------------------------------------
constexpr int kTabGroups = 4;
struct alignas(16) TabEntry {
    int64_t off;
    uint32_t code;  // parity bits | sign flip << 31
    uint32_t pad;
};
struct TabArgs {
    int32_t ng;
    uint32_t gmul[AX][kTabGroups];  // G_g = sum_ax o_ax * gmul[ax][g]
    const TabEntry *tab[kTabGroups];
};
struct TabBuild {
    int32_t ng;
    int64_t start[kTabGroups + 1];  // first entry of each group's table
    uint32_t vmask[kTabGroups], inmask[kTabGroups];
    int32_t nv[kTabGroups];
    int8_t vid[kTabGroups][AX];
    uint32_t vrad[kTabGroups][AX], vext[kTabGroups][AX];
};

__global__ void k_perm_tables(PermArgs a, TabBuild t, TabEntry *out) {
    const int64_t n = t.start[t.ng];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int g = 0;
        while (g + 1 < t.ng && e >= t.start[g + 1]) ++g;
        const uint32_t G = (uint32_t)(e - t.start[g]);
        uint64_t V[AX];
#pragma unroll
        for (int b = 0; b < AX; ++b) V[b] = 0;
        for (int i = 0; i < t.nv[g]; ++i) V[t.vid[g][i]] = (G / t.vrad[g][i]) % t.vext[g][i];
        int64_t off;
        uint32_t bits, neg;
        perm_core(a, V, t.vmask[g], t.inmask[g], off, bits, neg);
        TabEntry r;
        r.off = off;
        r.code = bits | (neg << 31);
        r.pad = 0;
        out[e] = r;
    }
}

// Table indices of the output axes 0 .. hi for the row-major index `lin` over those axes.
__device__ __forceinline__ void tab_index(const PermArgs &a, const TabArgs &t, uint64_t lin, int hi,
                                          uint32_t (&G)[kTabGroups]) {
#pragma unroll
    for (int g = 0; g < kTabGroups; ++g) G[g] = 0;
    uint64_t rest = lin;
    for (int ax = hi; ax >= 0; --ax) {
        uint32_t o;
        if (a.out_shift[ax] >= 0) {
            o = (uint32_t)(rest & (uint64_t)(a.out_dim[ax] - 1));
            rest >>= a.out_shift[ax];
        } else if ((rest >> 32) == 0) {
            const uint32_t r32 = (uint32_t)rest, d = a.out_dim[ax];
            o = r32 % d;
            rest = r32 / d;
        } else {
            o = (uint32_t)(rest % a.out_dim[ax]);
            rest /= a.out_dim[ax];
        }
#pragma unroll
        for (int g = 0; g < kTabGroups; ++g) G[g] += o * t.gmul[ax][g];
    }
}

// Source offset and sign code of the element with table indices G.
__device__ __forceinline__ int64_t tab_lookup(const PermArgs &a, const TabArgs &t, const uint32_t (&G)[kTabGroups],
                                              uint32_t &code) {
    int64_t off = 0;
    uint32_t c = 0;
#pragma unroll
    for (int g = 0; g < kTabGroups; ++g)
        if (g < t.ng) {
            const TabEntry e = t.tab[g][G[g]];
            off += e.off;
            c ^= e.code;
        }
    code = sign_code(a, c & 0x7fffffffu, c >> 31);
    return off;
}

__device__ __forceinline__ int64_t tab_map(const PermArgs &a, const TabArgs &t, uint64_t lin, uint32_t &code) {
    uint32_t G[kTabGroups];
#pragma unroll
    for (int g = 0; g < kTabGroups; ++g) G[g] = 0;
    uint64_t rest = lin;
    for (int ax = a.n_out - 1; ax >= 0; --ax) {
        uint32_t o;
        if (a.out_shift[ax] >= 0) {
            o = (uint32_t)(rest & (uint64_t)(a.out_dim[ax] - 1));
            rest >>= a.out_shift[ax];
        } else if ((rest >> 32) == 0) {
            const uint32_t r32 = (uint32_t)rest, d = a.out_dim[ax];
            o = r32 % d;
            rest = r32 / d;
        } else {
            o = (uint32_t)(rest % a.out_dim[ax]);
            rest /= a.out_dim[ax];
        }
#pragma unroll
        for (int g = 0; g < kTabGroups; ++g) G[g] += o * t.gmul[ax][g];
    }
    int64_t off = 0;
    uint32_t c = 0;
#pragma unroll
    for (int g = 0; g < kTabGroups; ++g)
        if (g < t.ng) {
            const TabEntry e = t.tab[g][G[g]];
            off += e.off;
            c ^= e.code;
        }
    code = sign_code(a, c & 0x7fffffffu, c >> 31);
    return off;
}

template <bool TAB>
__device__ __forceinline__ int64_t map_elem(const PermArgs &a, const TabArgs &t, uint64_t lin, uint32_t &code) {
    if constexpr (TAB)
        return tab_map(a, t, lin, code);
    else
        return perm_map(a, lin, code);
}

// A workgroup takes kPermUnroll * 256 consecutive output elements per pass, a lane every 256th of
// them (each load instruction stays coalesced, and the workgroups sweep the output in order, which
// keeps the source lines they share in L2); a lane issues the index maps and table reads of its
// kPermUnroll elements together, then their source reads, so kPermUnroll dependent load chains are in
// flight per lane instead of one.
constexpr int kPermUnroll = 4;
// Output suffix for the chunk-uniform walk: the fewest trailing output axes (at most two) whose
// extent is a multiple of the chunk and below 2^32 (the in-suffix index is 32-bit), so that every
// chunk shares the indices of the axes before them; 0 when there is none.
__device__ __forceinline__ int chunk_suffix(const PermArgs &a, int64_t chunk) {
    if (a.n_out >= 1 && (int64_t)a.out_dim[a.n_out - 1] % chunk == 0) return 1;
    if (a.n_out >= 2 && a.out_shift[a.n_out - 1] >= 0) {
        const int64_t span = (int64_t)a.out_dim[a.n_out - 1] * a.out_dim[a.n_out - 2];
        if (span % chunk == 0 && span < ((int64_t)1 << 32)) return 2;
    }
    return 0;
}

template <typename T, bool TAB>
__global__ void __launch_bounds__(256) k_permute(PermArgs a, TabArgs tb, const T *__restrict__ src, T *__restrict__ dst) {
    const double al = a.alpha_re;
    const int64_t chunk = (int64_t)kPermUnroll * blockDim.x;
    if constexpr (TAB) {
        // chunk-uniform walk: the leading axes' table indices are formed once per chunk (the same in
        // every lane), each element adds only its trailing one or two indices times their multipliers
        const int ns = chunk_suffix(a, chunk);
        if (ns > 0) {
            const int last = a.n_out - 1;
            const uint64_t span = ns == 1 ? a.out_dim[last] : (uint64_t)a.out_dim[last] * a.out_dim[last - 1];
            for (int64_t i0 = blockIdx.x * chunk; i0 < a.total; i0 += (int64_t)gridDim.x * chunk) {
                const uint64_t row = (uint64_t)i0 / span;
                const uint32_t c0 = (uint32_t)((uint64_t)i0 - row * span);
                uint32_t Gb[kTabGroups];
                tab_index(a, tb, row, last - ns, Gb);
                int64_t off[kPermUnroll];
                uint32_t code[kPermUnroll];
#pragma unroll
                for (int u = 0; u < kPermUnroll; ++u) {
                    const uint32_t c = c0 + threadIdx.x + u * blockDim.x;
                    uint32_t o0 = c, o1 = 0;
                    if (ns == 2) {
                        o0 = c & (a.out_dim[last] - 1);
                        o1 = c >> a.out_shift[last];
                    }
                    uint32_t G[kTabGroups];
#pragma unroll
                    for (int g = 0; g < kTabGroups; ++g)
                        G[g] = Gb[g] + o0 * tb.gmul[last][g] + (ns == 2 ? o1 * tb.gmul[last - 1][g] : 0u);
                    off[u] = tab_lookup(a, tb, G, code[u]);
                }
                double v[kPermUnroll];
#pragma unroll
                for (int u = 0; u < kPermUnroll; ++u) v[u] = (code[u] == 2u) ? 0.0 : (double)src[off[u]];
#pragma unroll
                for (int u = 0; u < kPermUnroll; ++u)
                    dst[i0 + threadIdx.x + u * blockDim.x] = (T)(code[u] == 2u ? 0.0 : v[u] * (code[u] ? -al : al));
            }
            return;
        }
    }
    for (int64_t i0 = blockIdx.x * chunk + threadIdx.x; i0 < a.total; i0 += (int64_t)gridDim.x * chunk) {
        int64_t off[kPermUnroll];
        uint32_t code[kPermUnroll];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            code[u] = 2u;
            off[u] = 0;
            if (i < a.total) off[u] = map_elem<TAB>(a, tb, (uint64_t)i, code[u]);
        }
        double v[kPermUnroll];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) v[u] = (code[u] == 2u) ? 0.0 : (double)src[off[u]];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            if (i < a.total) dst[i] = (T)(code[u] == 2u ? 0.0 : v[u] * (code[u] ? -al : al));
        }
    }
}

// Two gathers in one pass: dst[i] = sign_o(i) sign_n(j) src[off_n(j)] with j = off_o(i), the outer
// gather's source offset, read as the inner gather's output index (the inner gather's output is the
// contiguous tensor the outer one would have read).  Same lane arrangement as k_permute.
template <typename T, bool TABO, bool TABN>
__global__ void __launch_bounds__(256) k_permute2(PermArgs ao, TabArgs to, PermArgs an, TabArgs tn,
                                                  const T *__restrict__ src, T *__restrict__ dst) {
    const double al = ao.alpha_re * an.alpha_re;
    const int64_t chunk = (int64_t)kPermUnroll * blockDim.x;
    for (int64_t i0 = blockIdx.x * chunk + threadIdx.x; i0 < ao.total; i0 += (int64_t)gridDim.x * chunk) {
        int64_t off[kPermUnroll];
        uint32_t code[kPermUnroll];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            code[u] = 2u;
            off[u] = 0;
            if (i < ao.total) {
                uint32_t co, cn;
                const int64_t j = map_elem<TABO>(ao, to, (uint64_t)i, co);
                if (co != 2u) {
                    off[u] = map_elem<TABN>(an, tn, (uint64_t)j, cn);
                    code[u] = cn == 2u ? 2u : (co ^ cn);
                }
            }
        }
        double v[kPermUnroll];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) v[u] = (code[u] == 2u) ? 0.0 : (double)src[off[u]];
#pragma unroll
        for (int u = 0; u < kPermUnroll; ++u) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            if (i < ao.total) dst[i] = (T)(code[u] == 2u ? 0.0 : v[u] * (code[u] ? -al : al));
        }
    }
}

__global__ void k_permute_c128(PermArgs a, const double2 *__restrict__ src, double2 *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.total; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t code;
        int64_t off = perm_map(a, (uint64_t)i, code);
        double2 r = {0.0, 0.0};
        if (code != 2u) {
            double2 v = src[off];
            double s = code ? -1.0 : 1.0;
            r.x = s * (a.alpha_re * v.x - a.alpha_im * v.y);
            r.y = s * (a.alpha_re * v.y + a.alpha_im * v.x);
        }
        dst[i] = r;
    }
}

// Tiled variant for gathers whose output's last axis is far apart in the source: output axis t1
// (the one whose unit step moves the source the least) and the last output axis t0 form 64 x 64
// tiles.  A tile is read with t1 fastest (consecutive source addresses), staged in LDS with its
// signs applied, and written with t0 fastest (consecutive output addresses).
struct TileArgs {
    int32_t t0, t1, n_rest, tiles0, tiles1;
    int32_t rest_ax[AX];
    int64_t ostr[AX];
    int64_t rest_total;
};

template <typename T, bool TAB>
__global__ void __launch_bounds__(256) k_permute_tiled(PermArgs a, TabArgs tb, TileArgs t, const T *__restrict__ src,
                                                       T *__restrict__ dst) {
    __shared__ double tile[64][65];
    const double al = a.alpha_re;
    const int64_t per = (int64_t)t.tiles0 * t.tiles1;
    const int64_t nblk = per * t.rest_total;
    const int d1 = (int)a.out_dim[t.t1], d0 = (int)a.out_dim[t.t0];
    const int64_t s1 = t.ostr[t.t1];
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int64_t r = blk / per;
        const int64_t q = blk - r * per;
        const int i1b = (int)(q / t.tiles0) * 64, i0b = (int)(q % t.tiles0) * 64;
        int64_t base = 0, rr = r;
        for (int k = t.n_rest - 1; k >= 0; --k) {
            const int ax = t.rest_ax[k];
            const int64_t d = a.out_dim[ax];
            base += (rr % d) * t.ostr[ax];
            rr /= d;
        }
        // 16 elements per lane in two batches of 8: index maps and table reads first, then the
        // source reads, so that 8 load chains are in flight per lane
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int64_t off[8];
            uint32_t code[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = threadIdx.x + 256 * (8 * h + u);
                const int i1 = i1b + (e & 63), i0 = i0b + (e >> 6);
                code[u] = 2u;
                off[u] = 0;
                if (i1 < d1 && i0 < d0) off[u] = map_elem<TAB>(a, tb, (uint64_t)(base + i1 * s1 + i0), code[u]);
            }
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = (code[u] == 2u) ? 0.0 : (double)src[off[u]] * (code[u] ? -al : al);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = threadIdx.x + 256 * (8 * h + u);
                tile[e & 63][e >> 6] = v[u];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int l0 = e & 63, l1 = e >> 6;
            const int i1 = i1b + l1, i0 = i0b + l0;
            if (i1 < d1 && i0 < d0) dst[base + i1 * s1 + i0] = (T)tile[l1][l0];
        }
        __syncthreads();
    }
}

static int ilog2_exact(uint64_t d) {
    if (d == 0 || (d & (d - 1))) return -1;
    int s = 0;
    while ((1ull << s) < d) ++s;
    return s;
}

static int build_perm_args(const gtn_perm_desc *d, PermArgs &a) {
    GTN_REQUIRE(d->n_out >= 0 && d->n_out <= AX, "gtn_permute: n_out out of range");
    GTN_REQUIRE(d->n_virt >= 0 && d->n_virt <= AX, "gtn_permute: n_virt out of range");
    GTN_REQUIRE(d->n_comp >= 0 && d->n_comp <= NC, "gtn_permute: n_comp out of range");
    GTN_REQUIRE(d->n_in >= 0 && d->n_in <= AX, "gtn_permute: n_in out of range");
    GTN_REQUIRE(d->n_bits >= 0 && d->n_bits <= GTN_PERM_MAX_BITS, "gtn_permute: n_bits out of range");
    memset(&a, 0, sizeof(a));
    a.n_out = d->n_out;
    a.n_virt = d->n_virt;
    a.n_comp = d->n_comp;
    a.n_in = d->n_in;
    int64_t total = 1;
    int64_t vmax[AX] = {0};
    for (int ax = 0; ax < d->n_out; ++ax) {
        GTN_REQUIRE(d->out_dim[ax] >= 0 && d->out_dim[ax] < (1ll << 31), "gtn_permute: bad out_dim");
        total *= d->out_dim[ax];
        a.out_dim[ax] = (uint32_t)d->out_dim[ax];
        a.out_shift[ax] = (int8_t)ilog2_exact((uint64_t)d->out_dim[ax]);
        for (int t = 0; t < 2; ++t) {
            int v = d->out_virt[ax][t];
            GTN_REQUIRE(v >= -1 && v < d->n_virt, "gtn_permute: out_virt out of range");
            GTN_REQUIRE(d->out_vmul[ax][t] >= 0, "gtn_permute: negative out_vmul");
            a.out_virt[ax][t] = (int8_t)v;
            a.out_vmul[ax][t] = d->out_vmul[ax][t];
            if (v >= 0 && d->out_dim[ax] > 0) vmax[v] += (d->out_dim[ax] - 1) * d->out_vmul[ax][t];
        }
    }
    int jc = 0;
    int64_t fmax[AX] = {0};
    for (int v = 0; v < d->n_virt; ++v) {
        GTN_REQUIRE(d->virt_dim[v] >= 1, "gtn_permute: bad virt_dim");
        GTN_REQUIRE(vmax[v] < d->virt_dim[v], "gtn_permute: virtual axis %d overflows", v);
        if (d->virt_enc[v]) {
            // eps only flips the lowest bit, so it maps [0, d) onto itself for every even d (a fused
            // index of fermionic and bosonic legs is even but need not be a power of two)
            GTN_REQUIRE(d->virt_dim[v] % 2 == 0 || d->virt_dim[v] == 1, "gtn_permute: encoder on an odd axis");
            a.virt_enc_mask |= 1 << v;
        }
        if (d->virt_sigma[v]) a.virt_sigma_mask |= 1 << v;
        a.virt_ncomp[v] = (int8_t)d->virt_ncomp[v];
        int64_t prod = 1;
        for (int t = 0; t < d->virt_ncomp[v]; ++t, ++jc) {
            GTN_REQUIRE(jc < d->n_comp, "gtn_permute: components do not add up");
            prod *= d->comp_dim[jc];
        }
        GTN_REQUIRE(prod == d->virt_dim[v], "gtn_permute: components of virtual axis %d do not multiply to its dim", v);
    }
    GTN_REQUIRE(jc == d->n_comp, "gtn_permute: components do not add up");
    for (int j = 0; j < d->n_comp; ++j) {
        GTN_REQUIRE(d->comp_dim[j] >= 1 && d->comp_dim[j] < (1ll << 31), "gtn_permute: bad comp_dim");
        a.comp_dim[j] = (uint32_t)d->comp_dim[j];
        a.comp_shift[j] = (int8_t)ilog2_exact((uint64_t)d->comp_dim[j]);
        GTN_REQUIRE(d->comp_bit[j] >= -1 && d->comp_bit[j] < d->n_bits, "gtn_permute: comp_bit beyond n_bits");
        a.comp_bit[j] = (int8_t)d->comp_bit[j];
        if (d->comp_sigma[j]) a.comp_sigma_mask |= 1 << j;
        GTN_REQUIRE(d->comp_in[j][0] >= 0 && d->comp_in[j][0] < d->n_in, "gtn_permute: comp_in out of range");
        GTN_REQUIRE(d->comp_in[j][1] >= -1 && d->comp_in[j][1] < d->n_in, "gtn_permute: comp_in out of range");
        for (int t = 0; t < 2; ++t) {
            a.comp_in[j][t] = (int8_t)d->comp_in[j][t];
            a.comp_mult[j][t] = d->comp_mult[j][t];
            GTN_REQUIRE(d->comp_mult[j][t] >= 0, "gtn_permute: negative comp_mult");
            if (d->comp_in[j][t] >= 0) fmax[d->comp_in[j][t]] += (d->comp_dim[j] - 1) * d->comp_mult[j][t];
        }
    }
    // bounds: every fused input index stays inside its axis, the furthest element inside src
    int64_t far = 0;
    for (int b = 0; b < d->n_in; ++b) {
        GTN_REQUIRE(fmax[b] < d->in_dim[b], "gtn_permute: input axis %d overflows (%lld >= %lld)", b, (long long)fmax[b],
                    (long long)d->in_dim[b]);
        if (d->in_enc[b]) {
            GTN_REQUIRE(d->in_dim[b] % 2 == 0 || d->in_dim[b] == 1, "gtn_permute: encoder on an odd input axis");
            a.in_enc_mask |= 1 << b;
        }
        if (d->in_sigma[b]) a.in_sigma_mask |= 1 << b;
        GTN_REQUIRE(d->in_stride[b] >= 0 && d->in_stride_lo[b] >= 0, "gtn_permute: negative stride");
        a.in_stride[b] = d->in_stride[b];
        a.in_stride_lo[b] = d->in_stride_lo[b];
        // the largest encoded index is in_dim - 1 (eps maps [0, d) onto itself for even d)
        int64_t emax = d->in_dim[b] - 1;
        if (d->in_split[b]) {
            a.in_split_mask |= 1 << b;
            GTN_REQUIRE(d->in_dim[b] % 2 == 0 || d->in_dim[b] == 1, "gtn_permute: split axis needs an even dim");
            far += (emax >= 1 ? d->in_stride_lo[b] : 0) + (emax >> 1) * d->in_stride[b];
        } else {
            far += emax * d->in_stride[b];
        }
    }
    GTN_REQUIRE(total == 0 || far < d->src_elems, "gtn_permute: source too small (%lld needed, %lld given)",
                (long long)far + 1, (long long)d->src_elems);
    memcpy(a.table, d->table, sizeof(a.table));
    a.alpha_re = d->alpha_re;
    a.alpha_im = d->alpha_im;
    a.total = total;
    return GTN_OK;
}

// ---------------------------------------------------------------------------------------------
// sign tensor (diagnostic output of the reference's S1 / S3)
// ---------------------------------------------------------------------------------------------
struct SignArgs {
    int32_t ndim;
    uint32_t dims[GTN_PERM_MAX_BITS];
    uint32_t table[(1 << GTN_PERM_MAX_BITS) / 32];
    int64_t total;
};

__global__ void k_sign_tensor(SignArgs a, int8_t *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.total; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t r = (uint64_t)i;
        uint32_t bits = 0;
        for (int ax = a.ndim - 1; ax >= 0; --ax) {
            uint64_t k = r % a.dims[ax];
            r /= a.dims[ax];
            bits |= (popc64(k) & 1u) << ax;
        }
        out[i] = ((a.table[bits >> 5] >> (bits & 31u)) & 1u) ? (int8_t)-1 : (int8_t)1;
    }
}

// ---------------------------------------------------------------------------------------------
// reductions (FP64 accumulation, deterministic two-pass: per-block partials then one block)
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double sq(T v) { return (double)v * (double)v; }
template <>
__device__ __forceinline__ double sq<double2>(double2 v) { return v.x * v.x + v.y * v.y; }

__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    }
    __syncthreads();
    return t;
}

template <typename T>
__global__ void k_sumsq_partial(const T *__restrict__ x, int64_t n, double *part) {
    __shared__ double sh[16];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += sq(x[i]);
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_sum_final(const double *part, int np, double *out) {
    __shared__ double sh[16];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = t;
}

template <typename T>
__global__ void k_scale_by_norm(T *x, int64_t n, const double *nrm2) {
    double inv = 1.0 / sqrt(nrm2[0]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)((double)x[i] * inv);
}
__global__ void k_scale_by_norm_c(double2 *x, int64_t n, const double *nrm2) {
    double inv = 1.0 / sqrt(nrm2[0]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = x[i];
        v.x *= inv;
        v.y *= inv;
        x[i] = v;
    }
}

template <typename T>
__global__ void k_reduce_last(const T *__restrict__ src, T *__restrict__ dst, int64_t rows, int64_t inner) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0;
        for (int64_t t = 0; t < inner; ++t) acc += (double)src[r * inner + t];
        dst[r] = (T)acc;
    }
}
__global__ void k_reduce_last_c(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t rows, int64_t inner) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = {0, 0};
        for (int64_t t = 0; t < inner; ++t) {
            acc.x += src[r * inner + t].x;
            acc.y += src[r * inner + t].y;
        }
        dst[r] = acc;
    }
}

__global__ void k_lognorm(double *ln, const double *nrm2) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ln[0] = 2.0 * ln[0] + 0.5 * log(nrm2[0]);
}

// ---- batched normalisation: every point b of a batch of nb tensors (n elements each) --------------
// pass 1: per-point partial sums of squares, grid (G, nb); part[b * G + g]
template <typename T>
__global__ void k_sumsq_batch_partial(const T *__restrict__ x, int64_t n, double *part) {
    __shared__ double sh[16];
    const T *xb = x + (int64_t)blockIdx.y * n;
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += sq(xb[i]);
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}
// pass 2, one workgroup per point: nrm2[b]; the driver's log-norm rule (example.py:192)
// lognorm[b] = 2 lognorm[b] + log ||T_b||; hist[b * hist_len + step] = log ||T_b|| while step < hist_len
__global__ void k_norm_batch_final(const double *part, int G, double *nrm2, double *lognorm, double *hist,
                                   int hist_len, const double *step) {
    __shared__ double sh[16];
    const int b = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < G; i += blockDim.x) acc += part[(int64_t)b * G + i];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        nrm2[b] = t;
        const double ln = 0.5 * log(t);
        if (lognorm) lognorm[b] = 2.0 * lognorm[b] + ln;
        if (hist && step) {
            const int k = (int)step[0];
            if (k >= 0 && k < hist_len) hist[(int64_t)b * hist_len + k] = ln;
        }
    }
}
// pass 3: x_b /= ||x_b||, grid (G, nb)
template <typename T>
__global__ void k_scale_batch(const T *x, T *dst, int64_t n, const double *nrm2) {
    const double inv = 1.0 / sqrt(nrm2[blockIdx.y]);
    const T *xb = x + (int64_t)blockIdx.y * n;
    T *db = dst + (int64_t)blockIdx.y * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        db[i] = (T)((double)xb[i] * inv);
}
__global__ void k_step_increment(double *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] += 1.0;
}
// dst (nb, 2d, 2d) = diagonal matrices whose entry (r, r) is src[b][r % 2][r / 2] (src (nb, 2, d)):
// the singular values of the even and odd parity blocks interleaved by the parity-preserving index
template <typename T>
__global__ void k_diag_interleave(T *dst, int64_t nb, int64_t d, const T *src) {
    const int64_t w = 2 * d, total = nb * w * w;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / (w * w), rc = e - b * w * w, r = rc / w, c = rc - r * w;
        dst[e] = (r == c) ? src[b * w + (r & 1) * d + (r >> 1)] : (T)0;
    }
}

template <typename T>
__global__ void k_power_diag(T *x, int64_t n, int64_t stride, double p, double rcond) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = (double)x[i * stride];
        x[i * stride] = (T)(fabs(v) > rcond ? pow(v, p) : 0.0);
    }
}

template <typename T>
__global__ void k_elementwise(T *x, int64_t n, int op, double p, double rcond) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)x[i];
        double r = 0.0;
        if (fabs(v) > rcond) r = (op == GTN_EW_INV) ? 1.0 / v : (op == GTN_EW_EXP ? exp(v) : pow(v, p));
        x[i] = (T)r;
    }
}

struct OddArgs {
    int32_t ndim;
    uint32_t dims[GTN_PERM_MAX_AX];
    int32_t ferm_mask, pp_mask;
};
template <typename T>
__global__ void k_odd_partial(OddArgs a, const T *__restrict__ x, int64_t n, double *part) {
    __shared__ double sh[16];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t r = (uint64_t)i;
        uint32_t par = 0;
        for (int ax = a.ndim - 1; ax >= 0; --ax) {
            uint64_t k = r % a.dims[ax];
            r /= a.dims[ax];
            if ((a.ferm_mask >> ax) & 1) par ^= ((a.pp_mask >> ax) & 1) ? (uint32_t)(k & 1) : (popc64(k) & 1u);
        }
        if (par) acc += sq(x[i]);
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename T>
__global__ void k_axpby(const T *x, const T *y, T *out, int64_t n, double a, double b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (T)(a * (double)x[i] + (y ? b * (double)y[i] : 0.0));
}
__global__ void k_axpby_c(const double2 *x, const double2 *y, double2 *out, int64_t n, double ar, double ai, double br,
                          double bi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 u = x[i], r;
        r.x = ar * u.x - ai * u.y;
        r.y = ar * u.y + ai * u.x;
        if (y) {
            double2 v = y[i];
            r.x += br * v.x - bi * v.y;
            r.y += br * v.y + bi * v.x;
        }
        out[i] = r;
    }
}

// Partial sums of the two-pass reductions: 1024 doubles per (device, stream), so that reductions
// queued on different streams never share a buffer.  Kept for the life of the process.
static std::mutex g_part_mu;
static std::map<std::pair<int, hipStream_t>, double *> g_partials;
static int partial_buffer(hipStream_t s, double **p) {
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> l(g_part_mu);
    double *&b = g_partials[std::make_pair(dev, s)];
    if (!b) GTN_CHECK_HIP(hipMalloc(&b, 1024 * sizeof(double)));
    *p = b;
    return GTN_OK;
}

}  // namespace gtn

using namespace gtn;

extern "C" {

int gtn_version(void) { return 100; }

int gtn_set_dry_run(int n, const uint64_t *bases, const int64_t *bytes) {
    g_regions.clear();
    for (int i = 0; i < n; ++i) g_regions.push_back({bases[i], (uint64_t)bytes[i]});
    g_dry = n > 0;
    return GTN_OK;
}
const char *gtn_last_error(void) { return gtn::g_err; }

int gtn_device_info(char *name, int name_len, int *n_cu) {
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    GTN_CHECK_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) snprintf(name, name_len, "%s", p.gcnArchName);
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        set_error("gtn_device_info: device is %s, this library is built for gfx950", p.gcnArchName);
        return GTN_ENOTSUP;
    }
    return GTN_OK;
}

int gtn_index_tables(int64_t n, int32_t *popc, int8_t *sigma, int32_t *enc, void *stream) {
    GTN_REQUIRE(n >= 0, "gtn_index_tables: n < 0");
    if (n == 0) return GTN_OK;
    GTN_RANGE(popc, popc ? n * 4 : 0, "index_tables popc");
    GTN_RANGE(sigma, sigma ? n : 0, "index_tables sigma");
    GTN_RANGE(enc, enc ? n * 4 : 0, "index_tables enc");
    if (dry_run()) return GTN_OK;
    hipLaunchKernelGGL(k_index_tables, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), n, popc, sigma, enc);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

// Source stride of a unit step of output axis `ax`: through its virtual axis, the least significant
// component of that axis and the input axis it feeds (0 when it cannot be told: a trace, two virtual
// axes, a split input).
static int64_t unit_source_stride(const gtn_perm_desc *d, int ax) {
    if (d->out_virt[ax][1] >= 0 || d->out_virt[ax][0] < 0) return 0;
    const int v = d->out_virt[ax][0];
    int j = 0;
    for (int u = 0; u <= v; ++u) j += d->virt_ncomp[u];
    j -= 1;  // the last (least significant) component of v
    if (j < 0 || d->comp_in[j][1] >= 0) return 0;
    const int b = d->comp_in[j][0];
    if (d->in_split[b]) return 0;
    return d->in_stride[b] * d->comp_mult[j][0] * d->out_vmul[ax][0];
}

// The tiled kernel pays off when the output's last axis is strided in the source and another output
// axis of at least 16 entries steps the source by one or two elements.
static bool plan_tiles(const gtn_perm_desc *d, TileArgs &t) {
    const int n = d->n_out;
    if (n < 2) return false;
    const int t0 = n - 1;
    if (d->out_dim[t0] < 16) return false;
    const int64_t s0 = unit_source_stride(d, t0);
    if (s0 >= 1 && s0 <= 2) return false;  // already coalesced
    int t1 = -1;
    for (int ax = 0; ax < t0; ++ax) {
        const int64_t sa = unit_source_stride(d, ax);
        if (d->out_dim[ax] >= 16 && sa >= 1 && sa <= 2) {
            t1 = ax;
            break;
        }
    }
    if (t1 < 0) return false;
    memset(&t, 0, sizeof(t));
    t.t0 = t0;
    t.t1 = t1;
    int64_t acc = 1;
    for (int ax = n - 1; ax >= 0; --ax) {
        t.ostr[ax] = acc;
        acc *= d->out_dim[ax];
    }
    t.rest_total = 1;
    for (int ax = 0; ax < n; ++ax)
        if (ax != t0 && ax != t1) {
            t.rest_ax[t.n_rest++] = ax;
            t.rest_total *= d->out_dim[ax];
        }
    t.tiles0 = (int32_t)((d->out_dim[t0] + 63) / 64);
    t.tiles1 = (int32_t)((d->out_dim[t1] + 63) / 64);
    return true;
}

// Group the virtual axes for the table-driven map (k_perm_tables).  Returns false when the gather
// is too small to repay a table or a group's table would be large.
static bool plan_groups(const gtn_perm_desc *d, int64_t total, TabBuild &tb, TabArgs &ta) {
    if (total < (1 << 16) || d->n_virt == 0) return false;
    const int nv = d->n_virt;
    int64_t ext[AX];
    for (int v = 0; v < nv; ++v) ext[v] = 1;
    for (int ax = 0; ax < d->n_out; ++ax)
        for (int t = 0; t < 2; ++t)
            if (d->out_virt[ax][t] >= 0 && d->out_dim[ax] > 0)
                ext[d->out_virt[ax][t]] += (d->out_dim[ax] - 1) * d->out_vmul[ax][t];
    // union virtual axes whose components feed a common input axis
    int parent[AX];
    for (int v = 0; v < nv; ++v) parent[v] = v;
    auto root = [&](int v) {
        while (parent[v] != v) v = parent[v];
        return v;
    };
    int owner[AX];
    uint32_t feeds[AX] = {0};  // input axes fed by each virtual axis
    for (int b = 0; b < d->n_in; ++b) owner[b] = -1;
    int j = 0;
    for (int v = 0; v < nv; ++v)
        for (int t = 0; t < d->virt_ncomp[v]; ++t, ++j)
            for (int u = 0; u < 2; ++u) {
                const int b = d->comp_in[j][u];
                if (b < 0) continue;
                feeds[v] |= 1u << b;
                if (owner[b] < 0)
                    owner[b] = v;
                else
                    parent[root(v)] = root(owner[b]);
            }
    // groups in order of their smallest virtual axis
    int gid[AX], ng = 0;
    int64_t gsize[AX];
    int members[AX][AX], nmem[AX];
    for (int v = 0; v < nv; ++v) gid[v] = -1;
    for (int v = 0; v < nv; ++v) {
        const int r = root(v);
        if (gid[r] < 0) {
            gid[r] = ng;
            gsize[ng] = 1;
            nmem[ng] = 0;
            ++ng;
        }
        const int g = gid[r];
        members[g][nmem[g]++] = v;
        gsize[g] *= ext[v];
        if (gsize[g] > (1 << 22)) return false;
    }
    // merge the two smallest groups while there are too many
    while (ng > kTabGroups) {
        int a0 = 0, a1 = 1;
        if (gsize[a1] < gsize[a0]) std::swap(a0, a1);
        for (int g = 2; g < ng; ++g) {
            if (gsize[g] < gsize[a0]) {
                a1 = a0;
                a0 = g;
            } else if (gsize[g] < gsize[a1]) {
                a1 = g;
            }
        }
        if (a1 < a0) std::swap(a0, a1);
        gsize[a0] *= gsize[a1];
        if (gsize[a0] > (1 << 22)) return false;
        for (int i = 0; i < nmem[a1]; ++i) members[a0][nmem[a0]++] = members[a1][i];
        for (int g = a1; g + 1 < ng; ++g) {
            gsize[g] = gsize[g + 1];
            nmem[g] = nmem[g + 1];
            for (int i = 0; i < nmem[g]; ++i) members[g][i] = members[g + 1][i];
        }
        --ng;
    }
    int64_t entries = 0;
    for (int g = 0; g < ng; ++g) entries += gsize[g];
    if (entries * 4 > total) return false;
    memset(&tb, 0, sizeof(tb));
    memset(&ta, 0, sizeof(ta));
    tb.ng = ta.ng = ng;
    int64_t rad_of[AX];
    int grp_of[AX];
    for (int g = 0; g < ng; ++g) {
        tb.start[g + 1] = tb.start[g] + gsize[g];
        tb.nv[g] = nmem[g];
        int64_t rad = 1;
        for (int i = nmem[g] - 1; i >= 0; --i) {
            const int v = members[g][i];
            tb.vid[g][i] = (int8_t)v;
            tb.vrad[g][i] = (uint32_t)rad;
            tb.vext[g][i] = (uint32_t)ext[v];
            rad_of[v] = rad;
            grp_of[v] = g;
            rad *= ext[v];
            tb.vmask[g] |= 1u << v;
            tb.inmask[g] |= feeds[v];
        }
    }
    for (int ax = 0; ax < d->n_out; ++ax)
        for (int t = 0; t < 2; ++t) {
            const int v = d->out_virt[ax][t];
            if (v < 0) continue;
            ta.gmul[ax][grp_of[v]] += (uint32_t)(d->out_vmul[ax][t] * rad_of[v]);
        }
    return true;
}

// Device tables of the table-driven map, built once per gather shape and device and kept until
// the cache outgrows kTabCacheBytes (then every table is released after a device synchronisation).
struct TabCacheEntry {
    TabEntry *dev;
    int64_t entries;
};
static std::mutex g_tab_mu;
static std::map<std::string, TabCacheEntry> g_tab_cache;
static int64_t g_tab_bytes = 0;
constexpr int64_t kTabCacheBytes = 64ll << 20;

static int tab_tables(const PermArgs &a, const TabBuild &tb, hipStream_t s, TabArgs &ta) {
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    PermArgs k;
    memcpy(&k, &a, sizeof(k));
    memset(k.table, 0, sizeof(k.table));
    k.alpha_re = k.alpha_im = 0.0;
    std::string key((const char *)&dev, sizeof(dev));
    key.append((const char *)&k, sizeof(k));
    key.append((const char *)&tb, sizeof(tb));
    std::lock_guard<std::mutex> l(g_tab_mu);
    auto it = g_tab_cache.find(key);
    if (it == g_tab_cache.end()) {
        const int64_t n = tb.start[tb.ng];
        // over budget: drop the cache, unless a graph may hold some of its tables
        if (g_tab_bytes + n * (int64_t)sizeof(TabEntry) > kTabCacheBytes && !g_tab_cache.empty() &&
            !graphs_alive() && !capture_open()) {
            GTN_CHECK_HIP(hipDeviceSynchronize());
            for (auto &e : g_tab_cache) GTN_CHECK_HIP(hipFree(e.second.dev));
            g_tab_cache.clear();
            g_tab_bytes = 0;
        }
        TabEntry *p = nullptr;
        GTN_CHECK_HIP(hipMalloc(&p, n * sizeof(TabEntry)));
        // built on a stream of their own and finished here: the tables are shared by later gathers
        // on any stream, and the caller's stream may be capturing a graph
        hipStream_t bs = nullptr;
        GTN_CHECK_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
        hipLaunchKernelGGL(k_perm_tables, dim3(grid_for(n, 256, 4096)), dim3(256), 0, bs, a, tb, p);
        hipError_t le = hipGetLastError();
        hipError_t se = hipStreamSynchronize(bs);
        (void)hipStreamDestroy(bs);
        GTN_CHECK_HIP(le);
        GTN_CHECK_HIP(se);
        (void)s;
        it = g_tab_cache.emplace(key, TabCacheEntry{p, n}).first;
        g_tab_bytes += n * (int64_t)sizeof(TabEntry);
    }
    for (int g = 0; g < tb.ng; ++g) ta.tab[g] = it->second.dev + tb.start[g];
    return GTN_OK;
}

// GTN_PERM_TILED=0 in the environment turns the LDS-tiled gather off (tuning and tests).
static bool tiles_enabled() {
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("GTN_PERM_TILED");
        on = (e && atoi(e) == 0) ? 0 : 1;
    }
    return on == 1;
}

int gtn_permute(const gtn_perm_desc *d, const void *src, void *dst, void *stream) {
    GTN_REQUIRE(d && src && dst, "gtn_permute: null argument");
    PermArgs a;
    int rc = build_perm_args(d, a);
    if (rc) return rc;
    if (a.total == 0) return GTN_OK;
    if (dry_run()) {
        int es = d->dtype == GTN_C128 ? 16 : (d->dtype == GTN_F32 ? 4 : 8);
        GTN_RANGE(src, d->src_elems * es, "permute src");
        GTN_RANGE(dst, a.total * es, "permute dst");
        return GTN_OK;
    }
    hipStream_t s = as_stream(stream);
    TabArgs tb;
    memset(&tb, 0, sizeof(tb));
    TabBuild build;
    const bool tab = d->dtype != GTN_C128 && getenv("GTN_PERM_NO_TABLES") == nullptr && plan_groups(d, a.total, build, tb);
    if (tab) {
        rc = tab_tables(a, build, s, tb);
        if (rc) return rc;
    }
    TileArgs t;
    if (d->dtype != GTN_C128 && tiles_enabled() && plan_tiles(d, t)) {
        const int64_t nblk = (int64_t)t.tiles0 * t.tiles1 * t.rest_total;
        dim3 gt((unsigned)(nblk < (1 << 20) ? nblk : (1 << 20))), bt(256);
        if (d->dtype == GTN_F64) {
            if (tab)
                hipLaunchKernelGGL((k_permute_tiled<double, true>), gt, bt, 0, s, a, tb, t, (const double *)src, (double *)dst);
            else
                hipLaunchKernelGGL((k_permute_tiled<double, false>), gt, bt, 0, s, a, tb, t, (const double *)src, (double *)dst);
        } else {
            if (tab)
                hipLaunchKernelGGL((k_permute_tiled<float, true>), gt, bt, 0, s, a, tb, t, (const float *)src, (float *)dst);
            else
                hipLaunchKernelGGL((k_permute_tiled<float, false>), gt, bt, 0, s, a, tb, t, (const float *)src, (float *)dst);
        }
        GTN_CHECK_LAUNCH();
        return GTN_OK;
    }
    dim3 g(grid_for((a.total + kPermUnroll - 1) / kPermUnroll, 256, 1 << 20)), b(256);
    switch (d->dtype) {
        case GTN_F64:
            if (tab)
                hipLaunchKernelGGL((k_permute<double, true>), g, b, 0, s, a, tb, (const double *)src, (double *)dst);
            else
                hipLaunchKernelGGL((k_permute<double, false>), g, b, 0, s, a, tb, (const double *)src, (double *)dst);
            break;
        case GTN_F32:
            if (tab)
                hipLaunchKernelGGL((k_permute<float, true>), g, b, 0, s, a, tb, (const float *)src, (float *)dst);
            else
                hipLaunchKernelGGL((k_permute<float, false>), g, b, 0, s, a, tb, (const float *)src, (float *)dst);
            break;
        case GTN_C128: hipLaunchKernelGGL(k_permute_c128, g, b, 0, s, a, (const double2 *)src, (double2 *)dst); break;
        default: set_error("gtn_permute: unsupported dtype %d", d->dtype); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_permute2(const gtn_perm_desc *outer, const gtn_perm_desc *inner, const void *src, void *dst, void *stream) {
    GTN_REQUIRE(outer && inner && src && dst, "gtn_permute2: null argument");
    GTN_REQUIRE(outer->dtype == inner->dtype && (outer->dtype == GTN_F64 || outer->dtype == GTN_F32),
                "gtn_permute2: both gathers must be F64 or both F32");
    PermArgs ao, an;
    int rc = build_perm_args(outer, ao);
    if (rc) return rc;
    rc = build_perm_args(inner, an);
    if (rc) return rc;
    GTN_REQUIRE(outer->src_elems == an.total, "gtn_permute2: the outer gather reads %lld elements, the inner one makes %lld",
                (long long)outer->src_elems, (long long)an.total);
    if (ao.total == 0) return GTN_OK;
    if (dry_run()) {
        const int es = outer->dtype == GTN_F32 ? 4 : 8;
        GTN_RANGE(src, inner->src_elems * es, "permute2 src");
        GTN_RANGE(dst, ao.total * es, "permute2 dst");
        return GTN_OK;
    }
    hipStream_t s = as_stream(stream);
    TabArgs to, tn;
    memset(&to, 0, sizeof(to));
    memset(&tn, 0, sizeof(tn));
    TabBuild bo, bn;
    const bool tabs = getenv("GTN_PERM_NO_TABLES") == nullptr;
    const bool tabo = tabs && plan_groups(outer, ao.total, bo, to);
    const bool tabn = tabs && plan_groups(inner, an.total, bn, tn);
    if (tabo && (rc = tab_tables(ao, bo, s, to))) return rc;
    if (tabn && (rc = tab_tables(an, bn, s, tn))) return rc;
    dim3 g(grid_for((ao.total + kPermUnroll - 1) / kPermUnroll, 256, 1 << 20)), b(256);
#define GTN_P2(T, A, B) hipLaunchKernelGGL((k_permute2<T, A, B>), g, b, 0, s, ao, to, an, tn, (const T *)src, (T *)dst)
    if (outer->dtype == GTN_F64) {
        if (tabo && tabn) GTN_P2(double, true, true);
        else if (tabo) GTN_P2(double, true, false);
        else if (tabn) GTN_P2(double, false, true);
        else GTN_P2(double, false, false);
    } else {
        if (tabo && tabn) GTN_P2(float, true, true);
        else if (tabo) GTN_P2(float, true, false);
        else if (tabn) GTN_P2(float, false, true);
        else GTN_P2(float, false, false);
    }
#undef GTN_P2
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_sign_tensor(int ndim, const int64_t *dims, const uint32_t *table, int8_t *out, void *stream) {
    GTN_REQUIRE(ndim >= 0 && ndim <= GTN_PERM_MAX_BITS, "gtn_sign_tensor: ndim out of range");
    SignArgs a;
    memset(&a, 0, sizeof(a));
    a.ndim = ndim;
    int64_t total = 1;
    for (int i = 0; i < ndim; ++i) {
        GTN_REQUIRE(dims[i] >= 1 && dims[i] < (1ll << 31), "gtn_sign_tensor: bad dim");
        a.dims[i] = (uint32_t)dims[i];
        total *= dims[i];
    }
    a.total = total;
    memcpy(a.table, table, ((1u << ndim) + 31) / 32 * sizeof(uint32_t));
    hipLaunchKernelGGL(k_sign_tensor, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0, as_stream(stream), a, out);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_sumsq(const void *x, int64_t n, int dtype, double *out, void *stream) {
    hipStream_t s = as_stream(stream);
    double *part;
    int rc = partial_buffer(s, &part);
    if (rc) return rc;
    int g = grid_for(n, 256, 1024);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_sumsq_partial<double>, dim3(g), dim3(256), 0, s, (const double *)x, n, part); break;
        case GTN_F32: hipLaunchKernelGGL(k_sumsq_partial<float>, dim3(g), dim3(256), 0, s, (const float *)x, n, part); break;
        case GTN_C128: hipLaunchKernelGGL(k_sumsq_partial<double2>, dim3(g), dim3(256), 0, s, (const double2 *)x, n, part); break;
        default: set_error("gtn_sumsq: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1024), 0, s, (const double *)part, g, out);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_scale_by_norm(void *x, int64_t n, int dtype, const double *nrm2, void *stream) {
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(n, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_scale_by_norm<double>, g, b, 0, s, (double *)x, n, nrm2); break;
        case GTN_F32: hipLaunchKernelGGL(k_scale_by_norm<float>, g, b, 0, s, (float *)x, n, nrm2); break;
        case GTN_C128: hipLaunchKernelGGL(k_scale_by_norm_c, g, b, 0, s, (double2 *)x, n, nrm2); break;
        default: set_error("gtn_scale_by_norm: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_reduce_last(const void *src, void *dst, int64_t rows, int64_t inner, int dtype, void *stream) {
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(rows, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_reduce_last<double>, g, b, 0, s, (const double *)src, (double *)dst, rows, inner); break;
        case GTN_F32: hipLaunchKernelGGL(k_reduce_last<float>, g, b, 0, s, (const float *)src, (float *)dst, rows, inner); break;
        case GTN_C128: hipLaunchKernelGGL(k_reduce_last_c, g, b, 0, s, (const double2 *)src, (double2 *)dst, rows, inner); break;
        default: set_error("gtn_reduce_last: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_lognorm_update(double *ln, const double *nrm2, void *stream) {
    hipLaunchKernelGGL(k_lognorm, dim3(1), dim3(64), 0, as_stream(stream), ln, nrm2);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_power_diag(void *x, int64_t n, int64_t stride, int dtype, double p, double rcond, void *stream) {
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(n, 256)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_power_diag<double>, g, b, 0, s, (double *)x, n, stride, p, rcond); break;
        case GTN_F32: hipLaunchKernelGGL(k_power_diag<float>, g, b, 0, s, (float *)x, n, stride, p, rcond); break;
        default: set_error("gtn_power_diag: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_elementwise(void *x, int64_t n, int dtype, int op, double p, double rcond, void *stream) {
    GTN_REQUIRE(n >= 0 && (n == 0 || x), "gtn_elementwise: bad buffer");
    GTN_REQUIRE(op == GTN_EW_POWER || op == GTN_EW_INV || op == GTN_EW_EXP, "gtn_elementwise: unknown op %d", op);
    if (n == 0) return GTN_OK;
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(n, 256)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_elementwise<double>, g, b, 0, s, (double *)x, n, op, p, rcond); break;
        case GTN_F32: hipLaunchKernelGGL(k_elementwise<float>, g, b, 0, s, (float *)x, n, op, p, rcond); break;
        default: set_error("gtn_elementwise: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_axpby(const void *x, const void *y, void *out, int64_t n, int dtype, double ar, double ai, double br, double bi,
              void *stream) {
    if (n == 0) return GTN_OK;
    hipStream_t s = as_stream(stream);
    dim3 g(grid_for(n, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_axpby<double>, g, b, 0, s, (const double *)x, (const double *)y, (double *)out, n, ar, br); break;
        case GTN_F32: hipLaunchKernelGGL(k_axpby<float>, g, b, 0, s, (const float *)x, (const float *)y, (float *)out, n, ar, br); break;
        case GTN_C128:
            hipLaunchKernelGGL(k_axpby_c, g, b, 0, s, (const double2 *)x, (const double2 *)y, (double2 *)out, n, ar, ai, br, bi);
            break;
        default: set_error("gtn_axpby: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_normalize_batch(void *x, void *dst, int64_t nbatch, int64_t n, int dtype, double *nrm2, double *lognorm,
                        double *hist, int hist_len, double *step, void *stream) {
    GTN_REQUIRE(nbatch >= 1 && nbatch <= 1024 && n >= 1, "gtn_normalize_batch: need 1 <= nbatch <= 1024 and n >= 1");
    GTN_REQUIRE(x && nrm2, "gtn_normalize_batch: null buffer");
    GTN_REQUIRE(dtype == GTN_F64 || dtype == GTN_F32, "gtn_normalize_batch: dtype must be F64 or F32");
    hipStream_t s = as_stream(stream);
    double *part;
    int rc = partial_buffer(s, &part);
    if (rc) return rc;
    const int G = (int)(1024 / nbatch) < grid_for(n, 256, 1024) ? (int)(1024 / nbatch) : grid_for(n, 256, 1024);
    const dim3 g2((unsigned)G, (unsigned)nbatch), b(256);
    if (dtype == GTN_F64)
        hipLaunchKernelGGL(k_sumsq_batch_partial<double>, g2, b, 0, s, (const double *)x, n, part);
    else
        hipLaunchKernelGGL(k_sumsq_batch_partial<float>, g2, b, 0, s, (const float *)x, n, part);
    GTN_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_norm_batch_final, dim3((unsigned)nbatch), dim3(256), 0, s, (const double *)part, G, nrm2,
                       lognorm, hist, hist_len, (const double *)step);
    GTN_CHECK_LAUNCH();
    const dim3 g3((unsigned)grid_for(n, 256, (int)(65536 / nbatch) > 0 ? (int)(65536 / nbatch) : 1), (unsigned)nbatch);
    if (!dst) dst = x;
    if (dtype == GTN_F64)
        hipLaunchKernelGGL(k_scale_batch<double>, g3, b, 0, s, (const double *)x, (double *)dst, n, (const double *)nrm2);
    else
        hipLaunchKernelGGL(k_scale_batch<float>, g3, b, 0, s, (const float *)x, (float *)dst, n, (const double *)nrm2);
    GTN_CHECK_LAUNCH();
    if (step) {
        hipLaunchKernelGGL(k_step_increment, dim3(1), dim3(64), 0, s, step);
        GTN_CHECK_LAUNCH();
    }
    return GTN_OK;
}

int gtn_diag_interleave(void *dst, int64_t nbatch, int64_t d, const void *src, int dtype, void *stream) {
    GTN_REQUIRE(nbatch >= 0 && d >= 0 && (nbatch * d == 0 || (dst && src)), "gtn_diag_interleave: bad argument");
    if (nbatch * d == 0) return GTN_OK;
    hipStream_t s = as_stream(stream);
    const int64_t total = nbatch * 4 * d * d;
    dim3 g(grid_for(total, 256, 1 << 16)), b(256);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_diag_interleave<double>, g, b, 0, s, (double *)dst, nbatch, d, (const double *)src); break;
        case GTN_F32: hipLaunchKernelGGL(k_diag_interleave<float>, g, b, 0, s, (float *)dst, nbatch, d, (const float *)src); break;
        default: set_error("gtn_diag_interleave: dtype must be F64 or F32"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_odd_norm(const void *x, int ndim, const int64_t *dims, const int32_t *ferm, const int32_t *pp, int dtype,
                 double *out, void *stream) {
    GTN_REQUIRE(ndim >= 0 && ndim <= GTN_PERM_MAX_AX, "gtn_odd_norm: ndim out of range");
    OddArgs a;
    memset(&a, 0, sizeof(a));
    a.ndim = ndim;
    int64_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        a.dims[i] = (uint32_t)dims[i];
        n *= dims[i];
        if (ferm[i]) a.ferm_mask |= 1 << i;
        if (pp[i]) a.pp_mask |= 1 << i;
    }
    hipStream_t s = as_stream(stream);
    double *part;
    int rc = partial_buffer(s, &part);
    if (rc) return rc;
    int g = grid_for(n, 256, 1024);
    switch (dtype) {
        case GTN_F64: hipLaunchKernelGGL(k_odd_partial<double>, dim3(g), dim3(256), 0, s, a, (const double *)x, n, part); break;
        case GTN_F32: hipLaunchKernelGGL(k_odd_partial<float>, dim3(g), dim3(256), 0, s, a, (const float *)x, n, part); break;
        case GTN_C128: hipLaunchKernelGGL(k_odd_partial<double2>, dim3(g), dim3(256), 0, s, a, (const double2 *)x, n, part); break;
        default: set_error("gtn_odd_norm: unsupported dtype"); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1024), 0, s, (const double *)part, g, out);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

}  // extern "C"
