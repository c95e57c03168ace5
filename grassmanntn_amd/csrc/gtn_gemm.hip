// Grouped, batched MFMA GEMM for the parity-block contractions (replaces oe.contract at
// __init__.py:2295 and einsum_block's per-sector np.einsum at :2781).
//
// One launch covers up to kMaxGroups independent GEMMs (the nonzero parity sectors of a
// contraction, or the parity blocks of a decomposition).  Workgroups of 256 lanes (4 wave64s in a
// 2x2 arrangement) compute a BM x BN tile; operand tiles are staged through LDS.  Precision is a
// template parameter:
//   F64    : v_mfma_f64_16x16x4f64, FP64 accumulation
//   F32    : operands rounded to FP32 on their way into LDS, v_mfma_f32_16x16x4f32
//   BF16Xn : operands split into n BF16 terms on their way into LDS (hi/lo for n=2, hi/mid/lo for
//            n=3), v_mfma_f32_16x16x32_bf16 on the 3 (n=2) or 6 (n=3) leading cross products,
//            FP32 accumulation.
// The element type in HBM stays FP64 (or FP32); the reduced precision lives only between LDS and
// the matrix cores.
#include "gtn_common.h"
#include <stdlib.h>
#include <string.h>
#include <map>
#include <mutex>
#include <vector>

namespace gtn {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kMaxGroups = 12;

// splits > 1: the reduction dimension is cut into chunks of kchunk; the tile enumeration runs over
// batch * splits, and chunk sp of batch entry b writes its partial product to C + b * scb + sp * scs
// (a workspace summed afterwards by k_split_sum).
struct GemmGroup {
    const void *A;
    const void *B;
    void *C;
    int32_t M, N, K, batch;
    int32_t tiles_m, tiles_n;
    int32_t splits, kchunk;
    int64_t sam, sak, sbk, sbn, scm, scn, sab, sbb, scb, scs;
    double alpha, beta;
};

struct GemmArgs {
    int32_t n_groups;
    int32_t tile_start[kMaxGroups + 1];  // prefix sum of tiles (incl. batch) per group
    GemmGroup g[kMaxGroups];
};

// Locate this workgroup's (group, batch, tile_m, tile_n).  Tiles are enumerated per XCD in
// contiguous ranges so that workgroups sharing operand panels share an L2.
__device__ __forceinline__ bool locate(const GemmArgs &a, int &gi, int &b, int &tm, int &tn) {
    int total = a.tile_start[a.n_groups];
    int bid = blockIdx.x;
    int per = (total + 7) / 8;
    int xcd = bid % 8, slot = bid / 8;
    int t = xcd * per + slot;
    // blocks beyond `total` are launched only to round the grid to a multiple of 8
    if (t >= total) return false;
    gi = 0;
    while (gi + 1 < a.n_groups && t >= a.tile_start[gi + 1]) ++gi;
    t -= a.tile_start[gi];
    const GemmGroup &g = a.g[gi];
    int per_b = g.tiles_m * g.tiles_n;
    b = t / per_b;
    t -= b * per_b;
    tm = t / g.tiles_n;
    tn = t - tm * g.tiles_n;
    return true;
}

template <typename T>
__device__ __forceinline__ double ld(const T *p, int64_t i) { return (double)p[i]; }

// Load a (rows x BK) panel of an operand whose element (r, k) is at p[r*sr + k*sk] into
// registers, one element per slot; out-of-range elements read as zero.  `kc` selects the mapping
// of lanes to elements so that consecutive lanes read consecutive addresses.
template <typename T, int ROWS, int BK>
__device__ __forceinline__ void load_panel(const T *p, int64_t sr, int64_t sk, int r0, int k0, int R, int K,
                                           double (&reg)[ROWS * BK / 256]) {
    constexpr int N = ROWS * BK / 256;
    const bool kc = (sk == 1);
#pragma unroll
    for (int s = 0; s < N; ++s) {
        int e = threadIdx.x + 256 * s;
        int r, k;
        if (kc) {
            r = e / BK;
            k = e % BK;
        } else {
            k = e / ROWS;
            r = e % ROWS;
        }
        const int gr = r0 + r, gk = k0 + k;
        // branch-free: an out-of-range slot reads element (0, 0) of the panel's operand (valid for
        // R, K >= 1) and is replaced by zero, so every slot's load is in flight at once
        const bool ok = gr < R && gk < K;
        const double v = ld(p, ok ? (int64_t)gr * sr + (int64_t)gk * sk : (int64_t)0);
        reg[s] = ok ? v : 0.0;
    }
}
// Range-checked panel walk: load_panel at every panel.
template <typename T, int ROWS, int BK>
struct CheckedPanels {
    const T *p;
    int64_t sr, sk;
    int R0, R, K;
    __device__ __forceinline__ void load(int k0, double (&reg)[ROWS * BK / 256]) const {
        load_panel<T, ROWS, BK>(p, sr, sk, R0, k0, R, K, reg);
    }
};

// Panel walk over an operand whose ROWS x K slab (rows R0.., all of the reduction) lies wholly in
// range, every byte of it within 2^31 of the slab's first element.  The lane -> element map of
// load_panel is affine in the slot (256 is a multiple of BK and of ROWS): slot s holds element
// (r + s dr, k + s dk), so each slot is one buffer load from a resource at the slab's origin, this
// lane's 32-bit byte offset (advanced by BK elements of the reduction per panel) and a uniform
// per-slot byte offset -- no per-slot index arithmetic or range test.
template <typename T, int ROWS, int BK>
struct InnerPanels {
    static_assert(256 % BK == 0 && 256 % ROWS == 0, "the slot map needs 256 to be a multiple of BK and ROWS");
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff0, kbytes, sbytes;
    __device__ __forceinline__ InnerPanels(const T *p, int64_t sr, int64_t sk, int R0) {
        const bool kc = (sk == 1);
        const int t = threadIdx.x;
        const int r = kc ? t / BK : t % ROWS, k = kc ? t % BK : t / ROWS;
        // raw buffer (stride 0), num_records in bytes; word 3 = 0x00020000 for the gfx9 family.  The
        // origin and the strides are the same in every lane: readfirstlane states it, so that the
        // resource and the slot offsets live in scalar registers.
        const uint64_t o = (uint64_t)(p + (int64_t)R0 * sr);
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)o), hi = __builtin_amdgcn_readfirstlane((uint32_t)(o >> 32));
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), (short)0, 0x7fffffff, 0x00020000);
        voff0 = (uint32_t)(((int64_t)r * sr + (int64_t)k * sk) * (int64_t)sizeof(T));
        kbytes = __builtin_amdgcn_readfirstlane((uint32_t)(sk * (int64_t)sizeof(T)));
        sbytes = __builtin_amdgcn_readfirstlane((uint32_t)((kc ? (256 / BK) * sr : (256 / ROWS) * sk) * (int64_t)sizeof(T)));
    }
    // the slab fits the resource: rows R0 .. R0 + ROWS - 1 exist and all offsets stay below 2^31
    static __device__ __forceinline__ bool fits(int64_t sr, int64_t sk, int R0, int R, int K) {
        return R0 + ROWS <= R && sr >= 0 && sk >= 0 &&
               ((int64_t)(ROWS - 1) * sr + (int64_t)(K - 1) * sk) * (int64_t)sizeof(T) < ((int64_t)1 << 31);
    }
    __device__ __forceinline__ void load(int k0, double (&reg)[ROWS * BK / 256]) const {
        const uint32_t vo = voff0 + (uint32_t)k0 * kbytes;
#pragma unroll
        for (int s = 0; s < ROWS * BK / 256; ++s) {
            if constexpr (sizeof(T) == 8)
                reg[s] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)vo, (int)(s * sbytes), 0));
            else
                reg[s] = (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)vo, (int)(s * sbytes), 0));
        }
    }
};

template <int ROWS, int BK>
__device__ __forceinline__ void panel_coord(int s, bool kc, int &r, int &k) {
    int e = threadIdx.x + 256 * s;
    if (kc) {
        r = e / BK;
        k = e % BK;
    } else {
        k = e / ROWS;
        r = e % ROWS;
    }
}

// ------------------------------------------------------------------------------------------------
// FP64 / FP32 kernels: 16x16x4 MFMA, LDS panels stored k-major [BK][ROWS + pad].
// MFMA 16x16x4 operand layout (per lane l of the wave): A value = A[i = l%16][k = l/16],
// B value = B[k = l/16][j = l%16].  Result registers v = 0..3 of lane l hold C[row][l%16] with
// row = 4*(l/16) + v for the FP32 form and row = l/16 + 4*v for the FP64 form.
// ------------------------------------------------------------------------------------------------
template <typename TIN, typename TC, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) k_gemm_fp(GemmArgs args) {
    int gi, b, tm, tn;
    if (!locate(args, gi, b, tm, tn)) return;
    const GemmGroup &g = args.g[gi];
    const int sp = b % g.splits;
    b /= g.splits;
    const int kb = sp * g.kchunk;
    const int K = min(g.K - kb, g.kchunk);
    const TIN *A = (const TIN *)g.A + b * g.sab + kb * g.sak;
    const TIN *B = (const TIN *)g.B + b * g.sbb + kb * g.sbk;
    TIN *C = (TIN *)g.C + b * g.scb + sp * g.scs;
    constexpr int PAD = 1;  // rows of BM + 1: lanes storing one k-column of a row-major panel hit distinct banks
    // FP32 tiles double-buffer their panels (one barrier per panel: 4 % faster on the 2048^3 FP32
    // contraction); the FP64 kernel keeps one buffer, which leaves it the registers for two waves
    // per SIMD
    constexpr int NBUF = sizeof(TC) == 4 ? 2 : 1;
    __shared__ TC As[NBUF][BK][BM + PAD];
    __shared__ TC Bs[NBUF][BK][BN + PAD];
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    constexpr int WM = BM / 2, WN = BN / 2, TI = WM / 16, TJ = WN / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int m0 = tm * BM, n0 = tn * BN;
    using acc_t = typename std::conditional<sizeof(TC) == 8, d4, f4>::type;
    acc_t acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    const bool akc = (g.sak == 1), bkc = (g.sbk == 1);
    // slot s of this lane holds panel element (r + s dr, k + s dk) and goes to Xs[k + s dk][r + s dr]
    const int t = threadIdx.x;
    const int lda0 = akc ? (t % BK) * (BM + PAD) + t / BK : (t / BM) * (BM + PAD) + t % BM;
    const int ldb0 = bkc ? (t % BK) * (BN + PAD) + t / BK : (t / BN) * (BN + PAD) + t % BN;
    const int lda_step = akc ? 256 / BK : (256 / BM) * (BM + PAD);
    const int ldb_step = bkc ? 256 / BK : (256 / BN) * (BN + PAD);

    // the matrix-core work on the panels in buffer `b`
    auto panel_mfma = [&](int b) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            TC af[TI], bf[TJ];
#pragma unroll
            for (int i = 0; i < TI; ++i) af[i] = As[b][kk + lane / 16][wm + 16 * i + lane % 16];
#pragma unroll
            for (int j = 0; j < TJ; ++j) bf[j] = Bs[b][kk + lane / 16][wn + 16 * j + lane % 16];
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    if constexpr (sizeof(TC) == 8)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
                }
        }
    };
    auto mainloop = [&](const auto &la, const auto &lb) {
        double ra[NA], rb[NB];
        if constexpr (NBUF == 1) {
            if (K > 0) {
                la.load(0, ra);
                lb.load(0, rb);
            }
            for (int k0 = 0; k0 < K; k0 += BK) {
                __syncthreads();
#pragma unroll
                for (int s = 0; s < NA; ++s) (&As[0][0][0])[lda0 + s * lda_step] = (TC)ra[s];
#pragma unroll
                for (int s = 0; s < NB; ++s) (&Bs[0][0][0])[ldb0 + s * ldb_step] = (TC)rb[s];
                __syncthreads();
                if (k0 + BK < K) {  // prefetch the next panels while the matrix cores work
                    la.load(k0 + BK, ra);
                    lb.load(k0 + BK, rb);
                }
#pragma unroll
                for (int kk = 0; kk < BK; kk += 4) {
                    TC af[TI], bf[TJ];
#pragma unroll
                    for (int i = 0; i < TI; ++i) af[i] = As[0][kk + lane / 16][wm + 16 * i + lane % 16];
#pragma unroll
                    for (int j = 0; j < TJ; ++j) bf[j] = Bs[0][kk + lane / 16][wn + 16 * j + lane % 16];
#pragma unroll
                    for (int i = 0; i < TI; ++i)
#pragma unroll
                        for (int j = 0; j < TJ; ++j) {
                            if constexpr (sizeof(TC) == 8)
                                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                            else
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
                        }
                }
            }
        } else {
            if (K <= 0) return;
            la.load(0, ra);
            lb.load(0, rb);
#pragma unroll
            for (int s = 0; s < NA; ++s) (&As[0][0][0])[lda0 + s * lda_step] = (TC)ra[s];
#pragma unroll
            for (int s = 0; s < NB; ++s) (&Bs[0][0][0])[ldb0 + s * ldb_step] = (TC)rb[s];
            if (BK < K) {
                la.load(BK, ra);
                lb.load(BK, rb);
            }
            __syncthreads();
            int cur = 0;
            for (int k0 = 0; k0 < K; k0 += BK) {
                // the next panels go to the other buffer (its last readers passed the barrier that
                // ended the previous panel), the ones after them into registers
                if (k0 + BK < K) {
#pragma unroll
                    for (int s = 0; s < NA; ++s) (&As[cur ^ 1][0][0])[lda0 + s * lda_step] = (TC)ra[s];
#pragma unroll
                    for (int s = 0; s < NB; ++s) (&Bs[cur ^ 1][0][0])[ldb0 + s * ldb_step] = (TC)rb[s];
                    if (k0 + 2 * BK < K) {
                        la.load(k0 + 2 * BK, ra);
                        lb.load(k0 + 2 * BK, rb);
                    }
                }
                panel_mfma(cur);
                __syncthreads();
                cur ^= 1;
            }
        }
    };
    // interior tiles with whole panels take the buffer-load walk, the others the range-checked one
    if (K % BK == 0 && InnerPanels<TIN, BM, BK>::fits(g.sam, g.sak, m0, g.M, K) &&
        InnerPanels<TIN, BN, BK>::fits(g.sbn, g.sbk, n0, g.N, K))
        mainloop(InnerPanels<TIN, BM, BK>(A, g.sam, g.sak, m0), InnerPanels<TIN, BN, BK>(B, g.sbn, g.sbk, n0));
    else
        mainloop(CheckedPanels<TIN, BM, BK>{A, g.sam, g.sak, m0, g.M, K},
                 CheckedPanels<TIN, BN, BK>{B, g.sbn, g.sbk, n0, g.N, K});
    const double alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = sizeof(TC) == 8 ? (lane / 16) + 4 * v : 4 * (lane / 16) + v;
                int r = m0 + wm + 16 * i + row;
                int c = n0 + wn + 16 * j + lane % 16;
                if (r < g.M && c < g.N) {
                    int64_t o = r * g.scm + c * g.scn;
                    double val = alpha * (double)acc[i][j][v];
                    if (beta != 0.0) val += beta * (double)C[o];
                    C[o] = (TIN)val;
                }
            }
}

// ------------------------------------------------------------------------------------------------
// BF16 split-product kernel: 16x16x32 MFMA.  LDS panels stored row-major [ROWS][BK + pad] so that a
// lane's 8 consecutive k values are one 16-byte read.  Operand layout (lane l): A values
// A[i = l%16][k = 8*(l/16) + t], t = 0..7; B values B[k = 8*(l/16) + t][j = l%16]; the result layout
// is that of the 16x16x4 instructions.
// ------------------------------------------------------------------------------------------------
// SA / SB: BF16 terms each A / B element is split into (hi, mid, lo); the products of term t of A
// and term u of B with t + u < max(SA, SB) are accumulated, smallest first.
template <typename TIN, int SA, int SB, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs args) {
    constexpr int NSPLIT = SA > SB ? SA : SB;
    int gi, b, tm, tn;
    if (!locate(args, gi, b, tm, tn)) return;
    const GemmGroup &g = args.g[gi];
    const int sp = b % g.splits;
    b /= g.splits;
    const int kb = sp * g.kchunk;
    const int K = min(g.K - kb, g.kchunk);
    const TIN *A = (const TIN *)g.A + b * g.sab + kb * g.sak;
    const TIN *B = (const TIN *)g.B + b * g.sbb + kb * g.sbk;
    TIN *C = (TIN *)g.C + b * g.scb + sp * g.scs;
    constexpr int PAD = 8;
    __shared__ __attribute__((aligned(16))) __bf16 As[SA][BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[SB][BN][BK + PAD];
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    constexpr int WM = BM / 2, WN = BN / 2, TI = WM / 16, TJ = WN / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int m0 = tm * BM, n0 = tn * BN;
    f4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f4{0, 0, 0, 0};
    double ra[NA], rb[NB];
    const bool akc = (g.sak == 1), bkc = (g.sbk == 1);
    if (K > 0) {
        load_panel<TIN, BM, BK>(A, g.sam, g.sak, m0, 0, g.M, K, ra);
        load_panel<TIN, BN, BK>(B, g.sbn, g.sbk, n0, 0, g.N, K, rb);
    }
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NA; ++s) {
            int r, k;
            panel_coord<BM, BK>(s, akc, r, k);
            double x = ra[s];
#pragma unroll
            for (int t = 0; t < SA; ++t) {
                __bf16 h = (__bf16)(float)x;
                As[t][r][k] = h;
                x -= (double)(float)h;
            }
        }
#pragma unroll
        for (int s = 0; s < NB; ++s) {
            int r, k;
            panel_coord<BN, BK>(s, bkc, r, k);
            double x = rb[s];
#pragma unroll
            for (int t = 0; t < SB; ++t) {
                __bf16 h = (__bf16)(float)x;
                Bs[t][r][k] = h;
                x -= (double)(float)h;
            }
        }
        __syncthreads();
        if (k0 + BK < K) {
            load_panel<TIN, BM, BK>(A, g.sam, g.sak, m0, k0 + BK, g.M, K, ra);
            load_panel<TIN, BN, BK>(B, g.sbn, g.sbk, n0, k0 + BK, g.N, K, rb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            bf8 af[SA][TI], bfr[SB][TJ];
#pragma unroll
            for (int t = 0; t < SA; ++t)
#pragma unroll
                for (int i = 0; i < TI; ++i)
                    af[t][i] = *(const bf8 *)&As[t][wm + 16 * i + lane % 16][kk + 8 * (lane / 16)];
#pragma unroll
            for (int t = 0; t < SB; ++t)
#pragma unroll
                for (int j = 0; j < TJ; ++j)
                    bfr[t][j] = *(const bf8 *)&Bs[t][wn + 16 * j + lane % 16][kk + 8 * (lane / 16)];
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    // small cross terms first, the hi*hi product last
                    if constexpr (NSPLIT == 3) {
                        if constexpr (SA > 2) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2][i], bfr[0][j], acc[i][j], 0, 0, 0);
                        if constexpr (SA > 1 && SB > 1) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][i], bfr[1][j], acc[i][j], 0, 0, 0);
                        if constexpr (SB > 2) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfr[2][j], acc[i][j], 0, 0, 0);
                    }
                    if constexpr (NSPLIT >= 2) {
                        if constexpr (SA > 1) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][i], bfr[0][j], acc[i][j], 0, 0, 0);
                        if constexpr (SB > 1) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfr[1][j], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][i], bfr[0][j], acc[i][j], 0, 0, 0);
                }
        }
    }
    const double alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int r = m0 + wm + 16 * i + 4 * (lane / 16) + v;
                int c = n0 + wn + 16 * j + lane % 16;
                if (r < g.M && c < g.N) {
                    int64_t o = r * g.scm + c * g.scn;
                    double val = alpha * (double)acc[i][j][v];
                    if (beta != 0.0) val += beta * (double)C[o];
                    C[o] = (TIN)val;
                }
            }
}

// Lane-layout self-test: A[i][0] = 1, A[i][1] = i, B[0][j] = j + 1, B[1][j] = 16 (other k zero), so
// D[i][j] = 16 i + j + 1.  Each lane feeds the operand elements the GEMM kernels assume it holds and
// stores its four result registers.  One wavefront; writes exactly 768 doubles.
__device__ __forceinline__ double probe_a(int i, int k) { return k == 0 ? 1.0 : (k == 1 ? (double)i : 0.0); }
__device__ __forceinline__ double probe_b(int k, int j) { return k == 0 ? (double)(j + 1) : (k == 1 ? 16.0 : 0.0); }

__global__ void __launch_bounds__(64) k_mfma_selftest(double *out) {
    const int l = threadIdx.x;
    if (l >= 64) return;
    const int r = l % 16, q = l / 16;
    d4 a64 = {0, 0, 0, 0};
    a64 = __builtin_amdgcn_mfma_f64_16x16x4f64(probe_a(r, q), probe_b(q, r), a64, 0, 0, 0);
    f4 a32 = {0, 0, 0, 0};
    a32 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)probe_a(r, q), (float)probe_b(q, r), a32, 0, 0, 0);
    bf8 av, bv;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        av[t] = (__bf16)(float)probe_a(r, 8 * q + t);
        bv[t] = (__bf16)(float)probe_b(8 * q + t, r);
    }
    f4 ab = {0, 0, 0, 0};
    ab = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, ab, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        out[4 * l + v] = a64[v];
        out[256 + 4 * l + v] = (double)a32[v];
        out[512 + 4 * l + v] = (double)ab[v];
    }
}

// C[b](r, c) = alpha * sum_sp W[b][sp](r, c) + beta * C[b](r, c) over the split partials W
// (batch x splits consecutive M x N matrices), in a fixed order.
template <typename T>
__global__ void k_split_sum(const T *__restrict__ W, int splits, int M, int N, int batch, T *C, int64_t scm,
                            int64_t scn, int64_t scb, double alpha, double beta) {
    const int64_t MN = (int64_t)M * N, total = MN * batch;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / MN, e = i - b * MN;
        const int64_t r = e / N, c = e - r * N;
        const T *w = W + b * splits * MN + e;
        double acc = 0.0;
        for (int sp = 0; sp < splits; ++sp) acc += (double)w[sp * MN];
        const int64_t o = b * scb + r * scm + c * scn;
        double v = alpha * acc;
        if (beta != 0.0) v += beta * (double)C[o];
        C[o] = (T)v;
    }
}

// Split-K workspace: one grow-only buffer per device.  Launches on one stream are ordered; a launch
// on another stream first waits for the event recorded after the buffer's last use.
struct SplitWs {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t last = nullptr;
    hipEvent_t done = nullptr;
    bool used = false;
};
static std::mutex g_ws_mu;
static std::map<int, SplitWs> g_ws;

static std::vector<void *> g_ws_retired;  // outgrown workspaces a graph may still hold

static int split_workspace(SplitWs &w, size_t bytes, hipStream_t s) {
    if (w.used && w.last != s) GTN_CHECK_HIP(hipStreamWaitEvent(s, w.done, 0));
    if (w.bytes < bytes) {
        if (w.p) {
            if (graphs_alive() || capture_open()) {
                g_ws_retired.push_back(w.p);  // kept: a captured graph may address it
            } else {
                GTN_CHECK_HIP(hipDeviceSynchronize());
                GTN_CHECK_HIP(hipFree(w.p));
            }
            w.p = nullptr;
            w.bytes = 0;
        }
        GTN_CHECK_HIP(hipMalloc(&w.p, bytes));
        w.bytes = bytes;
    }
    if (!w.done) GTN_CHECK_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
    return GTN_OK;
}

// tile: 64 (64 x 64 tiles) or 128 (128 x 128 tiles; each of the 4 wavefronts owns 64 x 64)
template <typename TIN>
static int launch_gemm(GemmArgs &a, int precision, int tile, hipStream_t s) {
    int total = a.tile_start[a.n_groups];
    if (total == 0) return GTN_OK;
    int grid = (total + 7) / 8 * 8;
    const bool big = tile == 128;
    switch (precision) {
        case GTN_PREC_F64:
            if (big)
                hipLaunchKernelGGL((k_gemm_fp<TIN, double, 128, 128, 16>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_fp<TIN, double, 64, 64, 16>), dim3(grid), dim3(256), 0, s, a);
            break;
        case GTN_PREC_F32:
            if (big)
                hipLaunchKernelGGL((k_gemm_fp<TIN, float, 128, 128, 16>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_fp<TIN, float, 64, 64, 16>), dim3(grid), dim3(256), 0, s, a);
            break;
        case GTN_PREC_BF16X3:
            if (big)
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 2, 2, 128, 128, 32>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 2, 2, 64, 64, 32>), dim3(grid), dim3(256), 0, s, a);
            break;
        case GTN_PREC_BF16X6:
            if (big)
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 3, 3, 128, 128, 32>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 3, 3, 64, 64, 32>), dim3(grid), dim3(256), 0, s, a);
            break;
        case GTN_PREC_BF16:
            if (big)
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 1, 1, 128, 128, 32>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 1, 1, 64, 64, 32>), dim3(grid), dim3(256), 0, s, a);
            break;
        case GTN_PREC_BF16_SPLIT_A:
            if (big)
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 2, 1, 128, 128, 32>), dim3(grid), dim3(256), 0, s, a);
            else
                hipLaunchKernelGGL((k_gemm_bf16<TIN, 2, 1, 64, 64, 32>), dim3(grid), dim3(256), 0, s, a);
            break;
        default: set_error("gtn_gemm_grouped: unknown precision %d", precision); return GTN_ENOTSUP;
    }
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

// 128 x 128 tiles when every group is at least 128 x 128 and the launch still has a tile for every
// compute unit; 64 x 64 otherwise.  GTN_GEMM_TILE=64|128 in the environment forces a size (128 only
// where every group is at least 128 x 128) for tuning and tests.
static int pick_tile(const gtn_gemm_desc *groups, int n) {
    static int forced = -1;
    if (forced < 0) {
        const char *e = getenv("GTN_GEMM_TILE");
        forced = e ? atoi(e) : 0;
    }
    int64_t t128 = 0;
    for (int i = 0; i < n; ++i) {
        const gtn_gemm_desc &d = groups[i];
        if (d.M == 0 || d.N == 0 || d.batch == 0) continue;
        if (d.M < 128 || d.N < 128) return 64;
        t128 += ((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch;
    }
    if (forced == 64 || forced == 128) return forced;
    return t128 >= 256 ? 128 : 64;
}

// GTN_GEMM_SPLIT=0 in the environment turns the split reduction off; GTN_GEMM_SPLIT=n > 1 sets the
// workgroup count a split launch aims for (default 1024) (tuning and tests).
static int split_target() {
    static int target = -1;
    if (target < 0) {
        const char *e = getenv("GTN_GEMM_SPLIT");
        const int v = e ? atoi(e) : 1024;
        target = v == 0 ? 0 : (v > 1 ? v : 1024);
    }
    return target;
}

}  // namespace gtn

using namespace gtn;

// Split-reduction scratch of gtn_gemm_grouped_noalloc: one fixed array per device, held for the life
// of the process.  In host code `scratch` only names the symbol; its device address on the current
// device comes from hipGetSymbolAddress (scratch_ptr below) and is what kernels are given.
constexpr int64_t kNoAllocScratchBytes = 64ll << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// device address of `scratch` on the current device, looked up once per device
static int scratch_ptr(void **out) {
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    GTN_REQUIRE(dev >= 0 && dev < kMaxDevices, "gtn_gemm_grouped_noalloc: device index %d out of range", dev);
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("gtn_gemm_grouped_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return GTN_EDEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GTN_OK;
}

// The grouped GEMM.  fixed_ws == nullptr: split-reduction partials go to the library's per-device
// workspace (allocated and grown here).  Otherwise they go to fixed_ws (fixed_bytes long) and a
// split plan that needs more is an error; nothing is allocated.
// tile <= 0 / target < 0: the heuristics (pick_tile, split_target)
static int gemm_grouped_impl(const gtn_gemm_desc *groups, int n_groups, int dtype, int precision, void *stream,
                             void *fixed_ws, int64_t fixed_bytes, const char *name, int tile_cfg = 0,
                             int target_cfg = -1) {
    GTN_REQUIRE(n_groups >= 0 && (n_groups == 0 || groups), "%s: bad groups", name);
    GTN_REQUIRE(dtype == GTN_F64 || dtype == GTN_F32, "gtn_gemm_grouped: dtype must be F64 or F32");
    const int tile = tile_cfg > 0 ? tile_cfg : pick_tile(groups, n_groups);
    const int BM = tile, BN = tile;
    hipStream_t s = as_stream(stream);
    for (int base = 0; base < n_groups; base += kMaxGroups) {
        GemmArgs a;
        memset(&a, 0, sizeof(a));
        int n = n_groups - base < kMaxGroups ? n_groups - base : kMaxGroups;
        int tiles = 0;
        int used = 0;
        for (int i = 0; i < n; ++i) {
            const gtn_gemm_desc &d = groups[base + i];
            GTN_REQUIRE(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 0, "gtn_gemm_grouped: negative size");
            GTN_REQUIRE(d.M < (1 << 30) && d.N < (1 << 30) && d.K < (1 << 30), "gtn_gemm_grouped: size too large");
            GTN_REQUIRE(d.sak == 1 || d.sam == 1 || d.M <= 1 || d.K <= 1, "gtn_gemm_grouped: A needs a unit stride");
            GTN_REQUIRE(d.sbk == 1 || d.sbn == 1 || d.N <= 1 || d.K <= 1, "gtn_gemm_grouped: B needs a unit stride");
            if (d.M == 0 || d.N == 0 || d.batch == 0) continue;
            GTN_REQUIRE(d.A && d.B && d.C, "gtn_gemm_grouped: null operand");
            GemmGroup &g = a.g[used];
            g.A = d.A;
            g.B = d.B;
            g.C = d.C;
            g.M = (int32_t)d.M;
            g.N = (int32_t)d.N;
            g.K = (int32_t)d.K;
            g.batch = (int32_t)d.batch;
            g.tiles_m = (int32_t)((d.M + BM - 1) / BM);
            g.tiles_n = (int32_t)((d.N + BN - 1) / BN);
            // a unit stride on a length-1 dimension is ambiguous; normalise so the loader picks
            // the contiguous mapping
            g.sam = d.sam;
            g.sak = d.sak;
            if (d.K <= 1) g.sak = 1;
            g.sbk = d.sbk;
            g.sbn = d.sbn;
            if (d.K <= 1) g.sbk = 1;
            g.scm = d.scm;
            g.scn = d.scn;
            g.sab = d.sab;
            g.sbb = d.sbb;
            g.scb = d.scb;
            g.alpha = d.alpha;
            g.beta = d.beta;
            g.splits = 1;
            g.kchunk = g.K;
            g.scs = 0;
            a.tile_start[used] = tiles;
            int64_t t = (int64_t)g.tiles_m * g.tiles_n * g.batch;
            GTN_REQUIRE(tiles + t < (1ll << 31), "gtn_gemm_grouped: too many tiles");
            tiles += (int)t;
            ++used;
        }
        a.n_groups = used;
        a.tile_start[used] = tiles;
        if (used == 0) continue;
        if (dry_run()) {
            const int es = dtype == GTN_F64 ? 8 : 4;
            for (int i = 0; i < used; ++i) {
                const GemmGroup &q = a.g[i];
                int64_t lastb = q.batch - 1;
                int64_t ea = q.K > 0 ? lastb * q.sab + (int64_t)(q.M - 1) * q.sam + (int64_t)(q.K - 1) * q.sak + 1 : 0;
                int64_t eb = q.K > 0 ? lastb * q.sbb + (int64_t)(q.K - 1) * q.sbk + (int64_t)(q.N - 1) * q.sbn + 1 : 0;
                int64_t ec = lastb * q.scb + (int64_t)(q.M - 1) * q.scm + (int64_t)(q.N - 1) * q.scn + 1;
                GTN_RANGE(q.A, ea * es, "gemm A");
                GTN_RANGE(q.B, eb * es, "gemm B");
                GTN_RANGE(q.C, ec * es, "gemm C");
            }
            continue;
        }
        // Too few tiles to give every compute unit two workgroups: cut long reductions into chunks
        // (partials to the workspace, summed in a fixed order afterwards) until the launch has about
        // four workgroups per compute unit (the split target).  A chunk keeps at least 256 of the
        // reduction.
        const int es = dtype == GTN_F64 ? 8 : 4;
        std::vector<GemmGroup> orig;
        std::unique_lock<std::mutex> lk(g_ws_mu, std::defer_lock);
        SplitWs *ws = nullptr;
        const int target = target_cfg >= 0 ? target_cfg : split_target();
        if (tiles < 512 && target > 0 && tiles * 2 <= target) {
            int S = 1;
            while (tiles * S < target && S < 16) S *= 2;
            int64_t need = 0;
            int splitting = 0;
            for (int i = 0; i < used; ++i) {
                GemmGroup &g = a.g[i];
                int Sg = S;
                while (Sg > 1 && g.K / Sg < 256) Sg /= 2;
                if (Sg > 1) {
                    const int chunk = (g.K / Sg + 31) / 32 * 32;
                    Sg = (g.K + chunk - 1) / chunk;
                    g.splits = Sg;
                    g.kchunk = chunk;
                    need += (int64_t)Sg * g.batch * g.M * g.N;
                    ++splitting;
                }
            }
            if (splitting) {
                char *base_ws;
                if (fixed_ws) {
                    if (need * es > fixed_bytes) {
                        set_error("%s: the split reduction needs %lld bytes of scratch, above its maximum of %lld bytes",
                                  name, (long long)(need * es), (long long)fixed_bytes);
                        return GTN_ENOMEM;
                    }
                    base_ws = (char *)fixed_ws;
                } else {
                    int dev;
                    GTN_CHECK_HIP(hipGetDevice(&dev));
                    lk.lock();
                    ws = &g_ws[dev];
                    int rc = split_workspace(*ws, (size_t)need * es, s);
                    if (rc) return rc;
                    base_ws = (char *)ws->p;
                }
                orig.assign(a.g, a.g + used);
                int64_t off = 0;
                tiles = 0;
                for (int i = 0; i < used; ++i) {
                    GemmGroup &g = a.g[i];
                    a.tile_start[i] = tiles;
                    if (g.splits > 1) {
                        g.C = base_ws + off * es;
                        g.scm = g.N;
                        g.scn = 1;
                        g.scs = (int64_t)g.M * g.N;
                        g.scb = g.scs * g.splits;
                        g.alpha = 1.0;
                        g.beta = 0.0;
                        off += g.scb * g.batch;
                    }
                    tiles += g.tiles_m * g.tiles_n * g.batch * g.splits;
                }
                a.tile_start[used] = tiles;
            }
        }
        int rc = dtype == GTN_F64 ? launch_gemm<double>(a, precision, tile, s) : launch_gemm<float>(a, precision, tile, s);
        if (rc) return rc;
        if (!orig.empty()) {
            for (int i = 0; i < used; ++i) {
                const GemmGroup &g = a.g[i], &o = orig[i];
                if (g.splits == 1) continue;
                const dim3 grid(grid_for((int64_t)g.M * g.N * g.batch, 256, 4096)), blk(256);
                if (dtype == GTN_F64)
                    hipLaunchKernelGGL(k_split_sum<double>, grid, blk, 0, s, (const double *)g.C, g.splits, g.M, g.N,
                                       g.batch, (double *)o.C, o.scm, o.scn, o.scb, o.alpha, o.beta);
                else
                    hipLaunchKernelGGL(k_split_sum<float>, grid, blk, 0, s, (const float *)g.C, g.splits, g.M, g.N,
                                       g.batch, (float *)o.C, o.scm, o.scn, o.scb, o.alpha, o.beta);
                GTN_CHECK_LAUNCH();
            }
            if (ws) {
                GTN_CHECK_HIP(hipEventRecord(ws->done, s));
                ws->last = s;
                ws->used = true;
            }
        }
    }
    return GTN_OK;
}

// ---- autotuner: launch configuration per call signature -------------------------------------
// Candidates: tile 64 or 128 (128 only where every group is at least 128 x 128) times split-K
// target 0 (no split), 512, 1024 or 2048 workgroups.  The first eager call of a signature (element
// type, precision, and per group M, N, K, batch and operand layouts) times every candidate with HIP
// events (best of kTuneReps launches each) on its own operands and keeps the fastest; calls made while
// the stream is capturing a graph, and calls with beta != 0 (whose output a trial would change), use
// the tuned configuration when there is one and the heuristics otherwise.
struct TuneCfg {
    int tile, target;
    float ms, ms_default;
};
static std::mutex g_tune_mu;
static std::map<std::string, TuneCfg> g_tuned;
static int g_autotune = -1;
constexpr int kTuneReps = 3;

static bool autotune_on() {
    if (g_autotune < 0) {
        const char *e = getenv("GTN_AUTOTUNE");
        g_autotune = e ? (atoi(e) != 0) : 1;
    }
    return g_autotune == 1;
}

static std::string tune_key(const gtn_gemm_desc *g, int n, int dtype, int prec, int dev) {
    std::string k;
    auto put = [&](int64_t v) { k.append((const char *)&v, sizeof(v)); };
    put(dev);
    put(dtype);
    put(prec);
    put(n);
    for (int i = 0; i < n; ++i) {
        put(g[i].M);
        put(g[i].N);
        put(g[i].K);
        put(g[i].batch);
        put((g[i].sak == 1) | ((g[i].sbk == 1) << 1) | ((g[i].scn == 1) << 2));
    }
    return k;
}

static int time_cfg(const gtn_gemm_desc *groups, int n, int dtype, int prec, hipStream_t s, int tile, int target,
                    hipEvent_t e0, hipEvent_t e1, float *best) {
    *best = 1e30f;
    for (int r = 0; r < kTuneReps; ++r) {
        GTN_CHECK_HIP(hipEventRecord(e0, s));
        int rc = gemm_grouped_impl(groups, n, dtype, prec, s, nullptr, 0, "gtn_gemm_grouped", tile, target);
        if (rc) return rc;
        GTN_CHECK_HIP(hipEventRecord(e1, s));
        GTN_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GTN_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < *best) *best = ms;
    }
    return GTN_OK;
}

// the tuned configuration of this signature, if any (no timing)
static int tuned_lookup(const gtn_gemm_desc *groups, int n, int dtype, int prec, int *tile, int *target, bool *found) {
    *tile = 0;
    *target = -1;
    *found = false;
    if (!autotune_on() || dry_run() || n <= 0) return GTN_OK;
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> l(g_tune_mu);
    auto it = g_tuned.find(tune_key(groups, n, dtype, prec, dev));
    if (it != g_tuned.end()) {
        *tile = it->second.tile;
        *target = it->second.target;
        *found = true;
    }
    return GTN_OK;
}

static int tuned_config(const gtn_gemm_desc *groups, int n, int dtype, int prec, hipStream_t s, int *tile, int *target) {
    bool found;
    int rc0 = tuned_lookup(groups, n, dtype, prec, tile, target, &found);
    if (rc0 || found || !autotune_on() || dry_run() || n <= 0) return rc0;
    int dev;
    GTN_CHECK_HIP(hipGetDevice(&dev));
    const std::string key = tune_key(groups, n, dtype, prec, dev);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GTN_CHECK_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone || capture_open()) return GTN_OK;
    bool can128 = true;
    for (int i = 0; i < n; ++i) {
        if (groups[i].beta != 0.0) return GTN_OK;
        if (groups[i].M == 0 || groups[i].N == 0 || groups[i].batch == 0) continue;
        if (groups[i].M < 128 || groups[i].N < 128) can128 = false;
    }
    hipEvent_t e0, e1;
    GTN_CHECK_HIP(hipEventCreate(&e0));
    GTN_CHECK_HIP(hipEventCreate(&e1));
    TuneCfg best{0, -1, 1e30f, 0};
    float ms_default = 0;
    int rc = time_cfg(groups, n, dtype, prec, s, 0, -1, e0, e1, &ms_default);
    const int tiles[2] = {64, 128};
    const int targets[4] = {0, 512, 1024, 2048};
    for (int ti = 0; ti < 2 && !rc; ++ti) {
        if (tiles[ti] == 128 && !can128) continue;
        for (int gi = 0; gi < 4 && !rc; ++gi) {
            float ms;
            rc = time_cfg(groups, n, dtype, prec, s, tiles[ti], targets[gi], e0, e1, &ms);
            if (!rc && ms < best.ms) best = TuneCfg{tiles[ti], targets[gi], ms, 0};
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    best.ms_default = ms_default;
    std::lock_guard<std::mutex> l(g_tune_mu);
    g_tuned[key] = best;
    *tile = best.tile;
    *target = best.target;
    return GTN_OK;
}

extern "C" int gtn_gemm_grouped(const gtn_gemm_desc *groups, int n_groups, int dtype, int precision, void *stream) {
    int tile = 0, target = -1;
    int rc = tuned_config(groups, n_groups, dtype, precision, as_stream(stream), &tile, &target);
    if (rc) return rc;
    return gemm_grouped_impl(groups, n_groups, dtype, precision, stream, nullptr, 0, "gtn_gemm_grouped", tile, target);
}

extern "C" int gtn_autotune_set(int on) {
    g_autotune = on ? 1 : 0;
    return GTN_OK;
}

extern "C" int gtn_autotune_clear(void) {
    std::lock_guard<std::mutex> l(g_tune_mu);
    g_tuned.clear();
    return GTN_OK;
}

extern "C" int gtn_autotune_report(gtn_tune_entry *out, int max_entries) {
    std::lock_guard<std::mutex> l(g_tune_mu);
    int i = 0;
    for (auto &kv : g_tuned) {
        if (!out || i >= max_entries) break;
        const int64_t *k = (const int64_t *)kv.first.data();
        gtn_tune_entry &e = out[i++];
        memset(&e, 0, sizeof(e));
        e.dtype = (int32_t)k[1];
        e.precision = (int32_t)k[2];
        e.n_groups = (int32_t)k[3];
        e.M = k[4];
        e.N = k[5];
        e.K = k[6];
        e.batch = k[7];
        e.tile = kv.second.tile;
        e.split_target = kv.second.target;
        e.ms_best = kv.second.ms;
        e.ms_default = kv.second.ms_default;
    }
    return (int)g_tuned.size();
}

extern "C" int64_t gtn_gemm_noalloc_max_scratch(void) { return kNoAllocScratchBytes; }

// Follows gtn_gemm_grouped's tuned configuration of the signature (never timing one itself), unless
// that configuration's split plan would not fit the scratch: then the heuristics, whose split plans
// always fit (at most ~2 x 1024 workgroups of 64 x 64 FP64 partials = 64 MiB).
extern "C" int gtn_gemm_grouped_noalloc(const gtn_gemm_desc *groups, int n_groups, int dtype, int precision,
                                        void *stream) {
    void *p = nullptr;
    int rc = scratch_ptr(&p);
    if (rc) return rc;
    int tile = 0, target = -1;
    bool found = false;
    rc = tuned_lookup(groups, n_groups, dtype, precision, &tile, &target, &found);
    if (rc) return rc;
    if (found && target > 1024) {
        // a larger split target than the heuristics' may need more scratch than there is
        rc = gemm_grouped_impl(groups, n_groups, dtype, precision, stream, p, kNoAllocScratchBytes,
                               "gtn_gemm_grouped_noalloc", tile, target);
        if (rc != GTN_ENOMEM) return rc;
        tile = 0;
        target = -1;
    }
    return gemm_grouped_impl(groups, n_groups, dtype, precision, stream, p, kNoAllocScratchBytes,
                             "gtn_gemm_grouped_noalloc", tile, target);
}

// int8 16x16x64 lane-layout self-test (the operands are deterministic small integers; one wavefront)
typedef int i32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(64) k_mfma_selftest_i8(int32_t *out) {
    const int l = threadIdx.x;
    if (l >= 64) return;
    int8_t av[16], bv[16];
    for (int t = 0; t < 16; ++t) {
        const int k = 16 * (l / 16) + t;
        av[t] = (int8_t)(((l % 16) + 3 * k) % 7 - 3);
        bv[t] = (int8_t)((2 * (l % 16) + k) % 5 - 2);
    }
    i32x4 a, b;
    memcpy(&a, av, 16);
    memcpy(&b, bv, 16);
    i32x4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) out[4 * l + v] = acc[v];
}

extern "C" int gtn_mfma_selftest_i8(int32_t *out_dev, void *stream) {
    GTN_REQUIRE(out_dev, "gtn_mfma_selftest_i8: null output");
    hipLaunchKernelGGL(k_mfma_selftest_i8, dim3(1), dim3(64), 0, as_stream(stream), out_dev);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

extern "C" int gtn_mfma_selftest(double *out_dev, void *stream) {
    GTN_REQUIRE(out_dev, "gtn_mfma_selftest: null output");
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, as_stream(stream), out_dev);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}
