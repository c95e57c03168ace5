// Truncated SVD / Hermitian eig of a batch of parity blocks (replaces BlockSVD / SortedSVD,
// __init__.py:3931-4029, and BlockEig / SortedEig, :4314-4422).
//
// Large blocks (min(m, n) > kPMax): randomized subspace iteration with block width p, CholeskyQR
// re-orthonormalisation once per iteration (shifted, and repeated on the last pass), then an FP64
// Rayleigh-Ritz step whose p x p symmetric eigenproblem is solved by cyclic Jacobi in LDS.
// Small blocks (min(m, n) <= kPMax): the Gram matrix of the short side (or the Hermitian block
// itself in eig mode) is diagonalised directly by the same Jacobi kernel.
// Every stage processes all blocks of the call together.  Consecutive blocks of equal shape whose
// matrices and workspaces sit at constant strides form a "run": each stage issues one batched launch
// per run (batched GEMM groups, blockIdx.y over the run for the elementwise kernels, one workgroup
// per block for the LDS kernels), so a call over many equal blocks costs the same number of launches
// as a call over one.
#include "gtn_common.h"
#include <string.h>
#include <math.h>
#include <vector>
#include <utility>

namespace gtn {

// kernel launch that is skipped (after its range checks) in host dry-run validation
#define LAUNCH(...)                          \
    do {                                     \
        if (!dry_run()) {                    \
            hipLaunchKernelGGL(__VA_ARGS__); \
            GTN_CHECK_LAUNCH();              \
        }                                    \
    } while (0)
#define R8(p, elems, what) GTN_RANGE((p), (int64_t)(elems) * 8, (what))
// extent (elements) of a rows x cols matrix with leading dimension ld
static inline int64_t ext(int64_t rows, int64_t cols, int64_t ld) { return rows > 0 && cols > 0 ? (rows - 1) * ld + cols : 0; }

constexpr int kPMax = 128;   // largest p x p problem held in LDS (128 KiB of FP64)
constexpr int kMaxBatch = 8; // blocks per launch of the LDS kernels
constexpr int kMaxSweeps = 30;  // Jacobi sweep limit (the rotation log is sized for it)
constexpr int kJacobiTasks = 3;  // 2 x 2 blocks per thread and round: (64 * 65 / 2) / 1024 rounded up

// ---------------------------------------------------------------------------------------------
// counter-based Gaussian fill: a 64-bit mix of (seed, index) -> two uniforms -> Box-Muller
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void k_fill_gaussian(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, int64_t dx) {
    x += blockIdx.y * dx;
    seed = mix64(seed + blockIdx.y);
    int64_t n = rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h1 = mix64(seed ^ mix64((uint64_t)i * 2 + 1));
        uint64_t h2 = mix64(h1 + 0x632BE59BD9B4E019ull);
        double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0, 1]
        double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        int64_t row = i / cols, col = i - row * cols;
        x[row * ld + col] = r * cos(6.283185307179586 * u2);
    }
}

__global__ void k_sum_partials(const double *part, int64_t count, int S, double *out, int64_t dpart, int64_t dout) {
    part += blockIdx.y * dpart;
    out += blockIdx.y * dout;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int b = 0; b < S; ++b) s += part[b * count + i];
        out[i] = s;
    }
}

__global__ void k_copy2d(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                         int64_t dsrc, int64_t ddst) {
    src += blockIdx.y * dsrc;
    dst += blockIdx.y * ddst;
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        dst[r * ldd + c] = src[r * lds + c];
    }
}
// dst (cols x rows) = src (rows x cols)^T
__global__ void k_transpose2d(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols,
                              int64_t dsrc, int64_t ddst) {
    src += blockIdx.y * dsrc;
    dst += blockIdx.y * ddst;
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / rows, r = e - c * rows;
        dst[c * ldd + r] = src[r * lds + c];
    }
}
// G[r][c] = G[c][r] for the entries whose row super-block (r / S) lies below their column's
// (c / S): the part of a symmetric n x n matrix (ld) that a product over its upper block triangle
// left out.  64 x 64 tiles through LDS; grid (n / 64 rounded up, same, blocks), blockIdx.z advancing
// G by dG.
__global__ void __launch_bounds__(256) k_mirror_lower(double *G, int64_t n, int64_t ld, int64_t S, int64_t dG) {
    __shared__ double tile[64][65];
    G += blockIdx.z * dG;
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    if (r0 / S <= (c0 + 63) / S) return;  // the tile is not wholly below the computed block triangle
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {  // read the mirror tile (rows c0.., columns r0..) row by row
        const int64_t rr = c0 + i, cc = r0 + tx;
        tile[i][tx] = (rr < n && cc < n) ? G[rr * ld + cc] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int64_t rr = r0 + i, cc = c0 + tx;
        if (rr < n && cc < n) G[rr * ld + cc] = tile[tx][i];
    }
}
__global__ void k_fill(double *x, int64_t rows, int64_t cols, int64_t ld, double v, int64_t dx) {
    x += blockIdx.y * dx;
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        x[r * ld + c] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// LDS kernels, one workgroup per block
// ---------------------------------------------------------------------------------------------
// Either up to kMaxBatch independent blocks (uniform == 0: pointers and sizes per entry), or any
// number of equal blocks at constant strides (uniform == 1: entry 0 plus the strides d_*).
struct SmallBatch {
    int n;
    int uniform;
    int p[kMaxBatch];
    const double *in[kMaxBatch];
    double *out[kMaxBatch];
    double *aux[kMaxBatch];
    double *lam[kMaxBatch];
    int64_t d_in, d_out, d_aux, d_lam;
    double shift_rel;
    int by_abs;
    int max_sweeps;
};
struct SmallPtrs {
    int p;
    const double *in;
    double *out, *aux, *lam;
};
__device__ __forceinline__ SmallPtrs small_ptrs(const SmallBatch &B) {
    SmallPtrs r;
    const int64_t bi = blockIdx.x;
    if (B.uniform) {
        r.p = B.p[0];
        r.in = B.in[0] + bi * B.d_in;
        r.out = B.out[0] + bi * B.d_out;
        r.aux = B.aux[0] ? B.aux[0] + bi * B.d_aux : nullptr;
        r.lam = B.lam[0] ? B.lam[0] + bi * B.d_lam : nullptr;
    } else {
        r.p = B.p[bi];
        r.in = B.in[bi];
        r.out = B.out[bi];
        r.aux = B.aux[bi];
        r.lam = B.lam[bi];
    }
    return r;
}

__device__ double wg_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// Cholesky W = R^T R (W p x p symmetric positive semi-definite, shifted by shift_rel * trace / p),
// then out = R^{-1} (full p x p, zero below the diagonal).  The matrix is held in LDS padded to
// N = 16 * ceil(p / 16) (row stride N + 1, so that row and column walks of 16 lanes hit distinct
// banks; the padding is zero and factors as dependent columns) and processed in 16 x 16 tiles:
//   - one wavefront factors each diagonal tile in registers (lane c holds column c, broadcasts by
//     lane reads) and inverts it (Dinv_J, kept beside the matrix);
//   - the panel right of it is X = Dinv_J^T A[J, >J] and the trailing update A[I, T] -= X_I^T X_T,
//     one FP64 16x16x4 MFMA tile per wavefront;
//   - R^{-1} fills block columns left to right: X[<J, J] = -(X[<J, <J] R[<J, J]) Dinv_J, the product
//     staged in the (unused) lower block triangle.
// A pivot at or below 1e-14 of the largest diagonal entry (or NaN) marks its column as dependent:
// its row of R and its row and column of R^{-1} are zero, so that a rank-deficient basis gives
// zero columns instead of overflowing ones.
constexpr int kNB = 16;  // tile width of the blocked Cholesky / triangular inverse

typedef double d4 __attribute__((ext_vector_type(4)));

// value of `v` in lane `l` of the wavefront (l uniform)
__device__ __forceinline__ double lane_read(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Factor the 16 x 16 diagonal tile at T (row stride ld) in place (upper triangle R_JJ, zero below)
// and write its inverse to D (16 x 16, row stride 16, zero below).  Called by a whole wavefront;
// lanes 16..63 compute alongside and store nothing.
__device__ __forceinline__ void chol_tile(double *T, int ld, double *D, double tol, int lane) {
    const int c = lane;
    double col[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) col[r] = (c < 16 && r <= c) ? T[r * ld + c] : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double d = lane_read(col[k], k);
        const bool live = d > tol;  // false for NaN as well
        const double rkk = live ? sqrt(d) : 0.0, inv = live ? 1.0 / rkk : 0.0;
        if (c == k)
            col[k] = rkk;
        else if (c > k)
            col[k] *= inv;
#pragma unroll
        for (int i = k + 1; i < 16; ++i) {
            const double aki = lane_read(col[k], i);
            if (c >= i) col[i] -= aki * col[k];
        }
    }
    double inv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) inv[r] = 0.0;
#pragma unroll
    for (int i = 15; i >= 0; --i) {
        double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int q = i + 1; q < 16; ++q) v -= lane_read(col[i], q) * inv[q];
        const double rii = lane_read(col[i], i);
        inv[i] = (i <= c && rii > 0.0) ? v / rii : 0.0;
    }
    if (c < 16) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            T[r * ld + c] = (r <= c) ? col[r] : 0.0;
            D[r * 16 + c] = inv[r];
        }
    }
}

__global__ void __launch_bounds__(1024) k_chol_inv(SmallBatch B) {
    HIP_DYNAMIC_SHARED(double, sm)
    __shared__ double red[16];
    const SmallPtrs P = small_ptrs(B);
    const int p = P.p;
    const int N = (p + kNB - 1) / kNB * kNB, nb = N / kNB, LD = N + 1;
    const double *W = P.in;
    double *Rinv = P.out;
    double *A = sm;            // N x N at stride LD: R (upper), then R^{-1} in place
    double *Dv = sm + N * LD;  // nb inverted diagonal tiles
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    const int li = lane & 15, lq = lane >> 4;  // MFMA operand coordinates of this lane
    double tr = 0, dm = 0;
    for (int e = tid; e < N * N; e += nt) {
        const int i = e / N, j = e - i * N;
        double w = 0.0;
        if (i < p && j < p && j >= i) w = W[i * p + j];
        A[i * LD + j] = w;
        if (i == j && i < p) {
            tr += w;
            dm = fmax(dm, w);
        }
    }
    tr = wg_sum(tr, red);
    for (int o = 32; o > 0; o >>= 1) dm = fmax(dm, __shfl_down(dm, o, 64));
    if (lane == 0) red[wave] = dm;
    __syncthreads();
    dm = 0;
    for (int i = 0; i < nw; ++i) dm = fmax(dm, red[i]);
    __syncthreads();
    const double shift = B.shift_rel * tr / p;
    const double dead_tol = 1e-14 * (dm + shift);
    for (int i = tid; i < p; i += nt) A[i * LD + i] += shift;
    __syncthreads();
    // ---- right-looking blocked Cholesky W = R^T R
    for (int J = 0; J < nb; ++J) {
        const int j0 = J * kNB;
        if (wave == 0) chol_tile(A + j0 * LD + j0, LD, Dv + 256 * J, dead_tol, lane);
        __syncthreads();
        // panel rows of block J: X_T = Dinv_J^T A[J, T]
        for (int T = J + 1 + wave; T < nb; T += nw) {
            d4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = lq + 4 * t;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Dv[256 * J + k * 16 + li], A[(j0 + k) * LD + 16 * T + li], acc,
                                                           0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) A[(j0 + lq + 4 * v) * LD + 16 * T + li] = acc[v];
        }
        __syncthreads();
        // trailing update of the tiles (I, T), J < I <= T: A[I, T] -= X_I^T X_T
        const int m = nb - 1 - J;
        for (int idx = wave; idx < m * (m + 1) / 2; idx += nw) {
            int I = 0, rem = idx;
            while (rem >= m - I) {
                rem -= m - I;
                ++I;
            }
            const int gI = J + 1 + I, gT = gI + rem;
            d4 acc;
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[v] = A[(16 * gI + lq + 4 * v) * LD + 16 * gT + li];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = lq + 4 * t;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-A[(j0 + k) * LD + 16 * gI + li], A[(j0 + k) * LD + 16 * gT + li],
                                                           acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) A[(16 * gI + lq + 4 * v) * LD + 16 * gT + li] = acc[v];
        }
        __syncthreads();
    }
    // ---- R^{-1} in place: the inverted diagonal tiles, then block columns left to right
    for (int e = tid; e < nb * 256; e += nt) {
        const int K = e >> 8, r = (e >> 4) & 15, c = e & 15;
        A[(16 * K + r) * LD + 16 * K + c] = Dv[e];
    }
    __syncthreads();
    for (int J = 1; J < nb; ++J) {
        // S_I = X[I, I..J-1] R[I..J-1, J], staged in block (J, I) of the lower triangle
        for (int I = wave; I < J; I += nw) {
            d4 acc = {0, 0, 0, 0};
            for (int K = I; K < J; ++K) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int k = lq + 4 * t;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(16 * I + li) * LD + 16 * K + k],
                                                               A[(16 * K + k) * LD + 16 * J + li], acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) A[(16 * J + lq + 4 * v) * LD + 16 * I + li] = acc[v];
        }
        __syncthreads();
        // X[I, J] = -S_I Dinv_J
        for (int I = wave; I < J; I += nw) {
            d4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = lq + 4 * t;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-A[(16 * J + li) * LD + 16 * I + k], Dv[256 * J + k * 16 + li],
                                                           acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) A[(16 * I + lq + 4 * v) * LD + 16 * J + li] = acc[v];
        }
        __syncthreads();
    }
    for (int e = tid; e < p * p; e += nt) {
        const int i = e / p, j = e - i * p;
        Rinv[e] = (j >= i) ? A[i * LD + j] : 0.0;
    }
}

// round-robin pairing of round `round` (p even): pair t is (a, b) with a != b, every index once
// (t < p/2 and 0 <= round < p - 1 keep both sums below 2 (p - 1): one conditional subtraction is the
// modulo)
__device__ __forceinline__ void rr_pair(int t, int round, int p, int &i, int &j) {
    int a0 = t - 1 + round, b0 = p - 2 - t + round;
    if (a0 >= p - 1) a0 -= p - 1;
    if (b0 >= p - 1) b0 -= p - 1;
    const int a = (t == 0) ? 0 : a0 + 1;
    const int b = b0 + 1;
    i = a < b ? a : b;
    j = a < b ? b : a;
}

// Cyclic Jacobi on a symmetric p x p matrix (p even) in LDS.  Round-robin pairing gives p/2
// disjoint rotations per round and every pair once per sweep.  Each round computes its rotations,
// then applies both sides at once on the upper block triangle: a thread owns the 2 x 2 block of rows
// {i_a, j_a} x columns {i_b, j_b} for a pair of rotations a <= b and writes it and its mirror, so the
// matrix stays exactly symmetric; blocks whose two rotations are both the identity are left alone.
// Rounds that rotate nothing are not logged; the others append their (c, s) per pair to the log
// (out) and their round-robin index to the index list after the log (out + kMaxSweeps (p-1) p).
// k_jacobi_vectors multiplies the logged rounds out into the eigenvectors afterwards.  A rotation
// is skipped when |a_ij| <= 1e-15 sqrt(|a_ii a_jj|) or |a_ij| <= 1e-17 ||A||; the sweeps stop when a
// whole sweep skips every rotation.  On exit lam holds the eigenvalues sorted by decreasing lambda
// (or |lambda| when by_abs); aux[0] = rounds logged, aux[1 + c] = index of the c-th sorted
// eigenvalue.
__global__ void __launch_bounds__(1024) k_jacobi(SmallBatch B) {
    HIP_DYNAMIC_SHARED(double, sm)
    __shared__ double red[16];
    __shared__ int active, round_mark;
    const SmallPtrs P = small_ptrs(B);
    const int p = P.p;
    const double *Cin = P.in;
    double *info = P.aux;
    double *lg = P.out;
    double *ids = lg + (int64_t)kMaxSweeps * (p - 1) * p;
    double *lam = P.lam;
    const int LD = p + 1;  // row stride: row and column walks of a wavefront hit distinct banks
    double *A = sm;
    double *cs = sm + p * LD;
    int *pr = (int *)(cs + p);
    int *order = pr + p;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int h = p / 2;
    // this thread's blocks (a, b), a <= b, of the upper block triangle (the same every round)
    const int ntask = h / 2 * (h + 1) + (h & 1) * (h + 1) / 2;
    int my_ta[kJacobiTasks], my_tb[kJacobiTasks], my_n = 0;
    for (int e = tid; e < ntask && my_n < kJacobiTasks; e += nt) {
        // rows q and h-1-q of the upper triangle together hold h + 1 blocks
        const int q = e / (h + 1), r = e - q * (h + 1);
        if (r < h - q) {
            my_ta[my_n] = q;
            my_tb[my_n] = q + r;
        } else {
            my_ta[my_n] = h - 1 - q;
            my_tb[my_n] = h - 1 - q + (r - (h - q));
        }
        ++my_n;
    }
    double nrm = 0;
    for (int e = tid; e < p * p; e += nt) {
        const int i = e / p, j = e - i * p;
        double v = Cin[e];
        A[i * LD + j] = v;
        nrm += v * v;
    }
    if (tid == 0) {
        active = 0;
        round_mark = -1;
    }
    nrm = wg_sum(nrm, red);
    const double abs_tol = 1e-17 * sqrt(nrm);
    int rounds = 0, nlog = 0;
    for (int sweep = 0; sweep < B.max_sweeps; ++sweep) {
        for (int round = 0; round < p - 1; ++round, ++rounds) {
            for (int t = tid; t < h; t += nt) {
                int i, j;
                rr_pair(t, round, p, i, j);
                double aii = A[i * LD + i], ajj = A[j * LD + j], aij = A[i * LD + j];
                double c = 1.0, s = 0.0;
                const double fa = fabs(aij);
                if (fa > abs_tol && fa > 1e-15 * sqrt(fabs(aii * ajj))) {
                    double tau = (ajj - aii) / (2.0 * aij);
                    double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                    active = 1;
                    round_mark = rounds;  // every writer stores the same value
                }
                cs[t] = c;
                cs[h + t] = s;
                pr[t] = i;
                pr[h + t] = j;
                lg[((int64_t)nlog * h + t) * 2] = c;
                lg[((int64_t)nlog * h + t) * 2 + 1] = s;
            }
            if (tid == 0) ids[nlog] = (double)round;
            __syncthreads();
            if (round_mark == rounds) {  // uniform: written before the barrier, rewritten only by a later round
#pragma unroll
                for (int x = 0; x < kJacobiTasks; ++x) {
                    if (x >= my_n) break;
                    const int ta = my_ta[x], tb = my_tb[x];
                    const double sa = cs[h + ta], sb = cs[h + tb];
                    if (sa == 0.0 && sb == 0.0) continue;
                    const double ca = cs[ta], cb = cs[tb];
                    const int i = pr[ta], j = pr[h + ta], k = pr[tb], l = pr[h + tb];
                    const double xik = A[i * LD + k], xil = A[i * LD + l], xjk = A[j * LD + k], xjl = A[j * LD + l];
                    const double rik = ca * xik - sa * xjk, rjk = sa * xik + ca * xjk;
                    const double ril = ca * xil - sa * xjl, rjl = sa * xil + ca * xjl;
                    const double yik = cb * rik - sb * ril, yil = sb * rik + cb * ril;
                    const double yjk = cb * rjk - sb * rjl, yjl = sb * rjk + cb * rjl;
                    A[i * LD + k] = yik;
                    A[i * LD + l] = yil;
                    A[j * LD + k] = yjk;
                    A[j * LD + l] = yjl;
                    if (ta != tb) {
                        A[k * LD + i] = yik;
                        A[l * LD + i] = yil;
                        A[k * LD + j] = yjk;
                        A[l * LD + j] = yjl;
                    } else {
                        A[j * LD + i] = yil;  // the diagonal block: keep its two off-diagonal entries equal
                    }
                }
                ++nlog;
            }
            __syncthreads();
        }
        const int go = active;
        __syncthreads();
        if (!go) break;
        if (tid == 0) active = 0;
        __syncthreads();
    }
    // sort keys: lambda (or |lambda|), with NaN ranked below everything, ties broken by index, so
    // that the ranks are a permutation of 0..p-1 whatever the values
    for (int e = tid; e < p; e += nt) {
        double ke = A[e * LD + e];
        if (B.by_abs) ke = fabs(ke);
        if (ke != ke) ke = -INFINITY;
        int rank = 0;
        for (int f = 0; f < p; ++f) {
            double kf = A[f * LD + f];
            if (B.by_abs) kf = fabs(kf);
            if (kf != kf) kf = -INFINITY;
            if (kf > ke || (kf == ke && f < e)) ++rank;
        }
        order[rank] = e;
    }
    __syncthreads();
    for (int i = tid; i < p; i += nt) {
        lam[i] = A[order[i] * LD + order[i]];
        info[1 + i] = (double)order[i];
    }
    if (tid == 0) info[0] = (double)nlog;
}

// Eigenvectors from the rotation log of k_jacobi: V = J_1 J_2 ... J_R applied to the identity.  Rows
// of V evolve independently (v <- v J), so one wavefront owns one row of one block: lane t applies
// pair t of every logged round to the row held in LDS, with only wavefront-level ordering between
// rounds.
// out (p x p): column c = eigenvector of the c-th sorted eigenvalue.  Launch: grid (p, blocks), 64
// threads (p <= 128).
__global__ void __launch_bounds__(64) k_jacobi_vectors(SmallBatch B) {
    __shared__ double row[kPMax];
    SmallPtrs P;
    const int64_t bi = blockIdx.y;
    if (B.uniform) {
        P.p = B.p[0];
        P.in = B.in[0] + bi * B.d_in;
        P.out = B.out[0] + bi * B.d_out;
        P.aux = B.aux[0] + bi * B.d_aux;
    } else {
        P.p = B.p[bi];
        P.in = B.in[bi];
        P.out = B.out[bi];
        P.aux = B.aux[bi];
    }
    const int p = P.p, h = p / 2;
    const int r0 = blockIdx.x;
    if (r0 >= p) return;
    const double *lg = P.in;
    const double *ids = lg + (int64_t)kMaxSweeps * (p - 1) * p;
    const double *info = P.aux;
    const int t = threadIdx.x;
    const bool live = t < h;
    for (int c = t; c < p; c += 64) row[c] = (c == r0) ? 1.0 : 0.0;
    __builtin_amdgcn_wave_barrier();
    const int R = (int)info[0];
    // rounds are applied in chunks of kVecChunk whose (c, s, round index) are fetched one chunk
    // ahead, so that a chunk's loads are in flight while the previous chunk is applied
    constexpr int kVecChunk = 8;
    double cc[kVecChunk], ss[kVecChunk];
    int ii[kVecChunk];
    auto fetch = [&](int base, double (&c)[kVecChunk], double (&s)[kVecChunk], int (&id)[kVecChunk]) {
#pragma unroll
        for (int q = 0; q < kVecChunk; ++q) {
            const int r = base + q;
            const bool ok = live && r < R;
            c[q] = ok ? lg[((int64_t)r * h + t) * 2] : 1.0;
            s[q] = ok ? lg[((int64_t)r * h + t) * 2 + 1] : 0.0;
            id[q] = ok ? (int)ids[r] : 0;
        }
    };
    fetch(0, cc, ss, ii);
    for (int base = 0; base < R; base += kVecChunk) {
        double cn[kVecChunk], sn[kVecChunk];
        int in[kVecChunk];
        fetch(base + kVecChunk, cn, sn, in);
#pragma unroll
        for (int q = 0; q < kVecChunk; ++q) {
            if (live && ss[q] != 0.0) {
                int i, j;
                rr_pair(t, ii[q], p, i, j);
                const double vi = row[i], vj = row[j];
                row[i] = cc[q] * vi - ss[q] * vj;
                row[j] = ss[q] * vi + cc[q] * vj;
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int q = 0; q < kVecChunk; ++q) {
            cc[q] = cn[q];
            ss[q] = sn[q];
            ii[q] = in[q];
        }
    }
    for (int e = t; e < p; e += 64) P.out[(int64_t)r0 * p + e] = row[(int)info[1 + e]];
}

// ---------------------------------------------------------------------------------------------
// finishing kernels
// ---------------------------------------------------------------------------------------------
// sigma_i = sqrt(max(lam_i, 0)) (svd) or lam_i (eig); the rank rule |s_i| / (|s_0| + 1e-14) > 1e-14
// of __init__.py:3939-3941 gives keep_i in {0, 1}; inv_i = keep_i / |s_i|.
__global__ void k_sigma(const double *lam, int k, int eig_mode, double *S, double *inv, double *keep, int64_t dlam,
                        int64_t dS, int64_t dw) {
    lam += blockIdx.y * dlam;
    S += blockIdx.y * dS;
    inv += blockIdx.y * dw;
    keep += blockIdx.y * dw;
    int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= k) return;
    double s0 = eig_mode ? fabs(lam[0]) : sqrt(fmax(lam[0], 0.0));
    double s = eig_mode ? lam[i] : sqrt(fmax(lam[i], 0.0));
    bool kp = fabs(s / (s0 + 1e-14)) > 1e-14;
    // a NaN eigenvalue is passed on, so that the caller sees the breakdown
    S[i] = (lam[i] != lam[i]) ? lam[i] : (kp ? s : 0.0);
    inv[i] = kp ? 1.0 / fabs(s) : 0.0;
    keep[i] = kp ? 1.0 : 0.0;
}
// Singular values from the unnormalised singular vectors: vector i of X (k vectors of `len`
// elements, element j of vector i at X[i * sv + j * se]) is sigma_i times a unit vector, so
// sigma_i = ||x_i|| carries FP64 relative accuracy down to ~1e-16 sigma_0 (sqrt of a Gram eigenvalue
// cannot resolve sigma below ~1e-8 sigma_0).  Then the rank rule of __init__.py:3939-3941 as in
// k_sigma.  One workgroup of 256 threads (four wavefronts) per block; wavefront w reduces vectors
// i = w, w + 4, ...
__global__ void __launch_bounds__(256) k_sigma_norms(const double *X, int k, int64_t len, int64_t sv, int64_t se,
                                                     double *S, double *inv, double *keep, int64_t dX, int64_t dS,
                                                     int64_t dw) {
    X += blockIdx.y * dX;
    S += blockIdx.y * dS;
    inv += blockIdx.y * dw;
    keep += blockIdx.y * dw;
    __shared__ double sh[256];  // k <= 256 (checked on the host)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = w; i < k; i += 4) {
        double acc = 0.0;
        for (int64_t j = lane; j < len; j += 64) {
            double x = X[i * sv + j * se];
            acc = fma(x, x, acc);
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) sh[i] = sqrt(acc);
    }
    __syncthreads();
    const double s0 = sh[0];
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        double s = sh[i];
        bool kp = fabs(s / (s0 + 1e-14)) > 1e-14;
        S[i] = (s != s) ? s : (kp ? s : 0.0);
        inv[i] = kp ? 1.0 / s : 0.0;
        keep[i] = kp ? 1.0 : 0.0;
    }
}
__global__ void k_scale_rows(double *X, int k, int64_t n, int64_t ld, const double *f, int64_t dX, int64_t df) {
    X += blockIdx.y * dX;
    f += blockIdx.y * df;
    int64_t total = (int64_t)k * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / n, c = e - r * n;
        X[r * ld + c] *= f[r];
    }
}
__global__ void k_scale_cols(double *X, int64_t m, int k, int64_t ld, const double *f, int64_t dX, int64_t df) {
    X += blockIdx.y * dX;
    f += blockIdx.y * df;
    int64_t total = m * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / k, c = e - r * k;
        X[r * ld + c] *= f[c];
    }
}
// gauge: column j of U gets its largest-|.| entry (first in row order) positive; the matching row
// of Vt (when given) flips with it.  One workgroup per column.
__global__ void k_gauge(double *U, int64_t m, int64_t ldu, double *Vt, int64_t n, int64_t ldv, int64_t dU, int64_t dV) {
    const int j = blockIdx.x;
    U += blockIdx.y * dU;
    if (Vt) Vt += blockIdx.y * dV;
    __shared__ double bv[256];
    __shared__ int64_t bi[256];
    double best = -1;
    int64_t arg = 0;
    for (int64_t r = threadIdx.x; r < m; r += blockDim.x) {
        double a = fabs(U[r * ldu + j]);
        if (a > best) {
            best = a;
            arg = r;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = arg;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            double b2 = bv[threadIdx.x + o];
            int64_t i2 = bi[threadIdx.x + o];
            if (b2 > bv[threadIdx.x] || (b2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
                bv[threadIdx.x] = b2;
                bi[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    const bool flip = U[bi[0] * ldu + j] < 0;
    __syncthreads();
    if (!flip) return;
    for (int64_t r = threadIdx.x; r < m; r += blockDim.x) U[r * ldu + j] = -U[r * ldu + j];
    if (Vt)
        for (int64_t c = threadIdx.x; c < n; c += blockDim.x) Vt[j * ldv + c] = -Vt[j * ldv + c];
}

// ---------------------------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// Range finder in split BF16.  Reduced-precision iterations keep their operands as BF16 "planes":
// x = hi + lo + O(2^-16 |x|) with hi = bf16(fp32(x)), lo = bf16(fp32(x - hi)), each operand a pair of
// row-major matrices (hi, then lo) whose rows run along the reduction index, zero-padded to 128 rows
// and 128 columns.  The block M is converted once per call into planes of M and of M^T, the basis Y
// into planes of Y^T every iteration, and Z = M^T Y leaves the GEMM as planes of Z^T directly, so
// every product is C = A B^T over two K-contiguous operands (the layout the 16x16x32 BF16 MFMA reads
// with one 16-byte load per lane) and the block's 16 bytes per element of FP64 become 4 (M) per pass.
// ---------------------------------------------------------------------------------------------
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int kRfTile = 128;  // rows per workgroup tile and the padded column count of the B planes
constexpr int kRfBK = 64;     // reduction depth per LDS stage
static inline int64_t rf_pad(int64_t x) { return (x + kRfTile - 1) / kRfTile * kRfTile; }

__device__ __forceinline__ void split_bf16(double x, __bf16 &h, __bf16 &l) {
    h = (__bf16)(float)x;
    l = (__bf16)(float)(x - (double)(float)h);
}

// Planes of X (R x Cc, ld ldx; zero beyond) over the padded Rp x Cp grid: direct planes D (Rp x Cp,
// hi at D, lo at D + Rp Cp) and / or transposed planes T (Cp x Rp, hi at T, lo at T + Rp Cp).
// grid (Cp / 64, Rp / 64, blocks); blockIdx.z advances X by dX doubles and D, T by dP bf16 elements.
__global__ void __launch_bounds__(256) k_planes(const double *X, int64_t ldx, int64_t R, int64_t Cc, int64_t Rp,
                                                int64_t Cp, __bf16 *D, __bf16 *T, int64_t dX, int64_t dP) {
    __shared__ double tile[64][65];
    X += blockIdx.z * dX;
    const int64_t plane = Rp * Cp;
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tc = threadIdx.x & 63, tr = threadIdx.x >> 6;
#pragma unroll 4
    for (int i = tr; i < 64; i += 4) {
        const int64_t r = r0 + i, c = c0 + tc;
        const double x = (r < R && c < Cc) ? X[r * ldx + c] : 0.0;
        tile[i][tc] = x;
        if (D) {
            __bf16 h, l;
            split_bf16(x, h, l);
            __bf16 *d = D + blockIdx.z * dP + r * Cp + c;
            d[0] = h;
            d[plane] = l;
        }
    }
    if (!T) return;
    __syncthreads();
#pragma unroll 4
    for (int i = tr; i < 64; i += 4) {
        // T[c0 + i][r0 + tc] = X[r0 + tc][c0 + i]
        __bf16 h, l;
        split_bf16(tile[tc][i], h, l);
        __bf16 *t = T + blockIdx.z * dP + (c0 + i) * Rp + r0 + tc;
        t[0] = h;
        t[plane] = l;
    }
}

struct RfArgs {
    const __bf16 *A, *B;     // hi planes; lo planes follow at +plane_a / +plane_b
    void *C;
    int64_t K;               // reduction length (a multiple of kRfBK; both operands zero beyond the data)
    int64_t plane_a, plane_b;
    int64_t rows, rows_pad;  // rows of C with data / rows of the A planes (a multiple of kRfTile)
    int64_t cols, ldc;       // columns of C with data (<= kRfTile); FP64 output leading dimension
    int64_t dA, dB, dC;      // per-block strides (bf16 elements for planes, doubles for FP64 C)
    int32_t nb;              // blocks
};

// C = A B^T for every block, A (rows_pad x K) and B (kRfTile x K) in split-BF16 planes: the products
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  OUT 0: C FP64
// (rows x cols, ld ldc); OUT 1: C^T as planes (kRfTile x rows_pad, ld rows_pad), padding included
// (zero where A or B is zero padding).  One workgroup (4 waves, 2 x 2, 64 x 64 each) per 128-row
// tile; tiles of one block are enumerated contiguously per XCD so that they share its L2 (B).
template <int OUT>
__global__ void __launch_bounds__(256) k_rf_nt(RfArgs a) {
    const int64_t tiles = a.rows_pad / kRfTile, total = tiles * a.nb;
    const int64_t per = (total + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= total) return;
    const int64_t blk = t / tiles, m0 = (t - blk * tiles) * kRfTile;
    const __bf16 *A = a.A + blk * a.dA + m0 * a.K;
    const __bf16 *B = a.B + blk * a.dB;
    constexpr int PAD = 8, LD = kRfBK + PAD;
    __shared__ __attribute__((aligned(16))) __bf16 As[2][kRfTile][LD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][kRfTile][LD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0, 0, 0, 0};
    // 16-byte chunk s of this thread: plane s / 4, row (tid + 256 (s % 4)) / 8, k chunk tid % 8
    u32x4 ra[8], rb[8];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int c = threadIdx.x + 256 * (s & 3);
            const int row = c >> 3, kc = c & 7;
            const int64_t off = (int64_t)row * a.K + k0 + 8 * kc;
            ra[s] = *(const u32x4 *)(A + (s >> 2) * a.plane_a + off);
            rb[s] = *(const u32x4 *)(B + (s >> 2) * a.plane_b + off);
        }
    };
    load(0);
    for (int64_t k0 = 0; k0 < a.K; k0 += kRfBK) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int c = threadIdx.x + 256 * (s & 3);
            const int row = c >> 3, kc = c & 7;
            *(u32x4 *)&As[s >> 2][row][8 * kc] = ra[s];
            *(u32x4 *)&Bs[s >> 2][row][8 * kc] = rb[s];
        }
        __syncthreads();
        if (k0 + kRfBK < a.K) load(k0 + kRfBK);
#pragma unroll
        for (int kk = 0; kk < kRfBK; kk += 32) {
            bf8v ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ah[i] = *(const bf8v *)&As[0][wm + 16 * i + lane % 16][kk + 8 * (lane / 16)];
                al[i] = *(const bf8v *)&As[1][wm + 16 * i + lane % 16][kk + 8 * (lane / 16)];
                bh[i] = *(const bf8v *)&Bs[0][wn + 16 * i + lane % 16][kk + 8 * (lane / 16)];
                bl[i] = *(const bf8v *)&Bs[1][wn + 16 * i + lane % 16][kk + 8 * (lane / 16)];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        }
    }
    // result register v of lane l holds C[4 (l / 16) + v][l % 16] of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t r = m0 + wm + 16 * i + 4 * (lane / 16);
            const int64_t c = wn + 16 * j + lane % 16;
            if (OUT == 0) {
                double *C = (double *)a.C + blk * a.dC;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (r + v < a.rows && c < a.cols) C[(r + v) * a.ldc + c] = (double)acc[i][j][v];
            } else {
                __bf16 *C = (__bf16 *)a.C + blk * a.dC + c * a.rows_pad + r;
                __bf16 h[4], l[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) split_bf16((double)acc[i][j][v], h[v], l[v]);
                u32x2 hv, lv;
                memcpy(&hv, h, 8);
                memcpy(&lv, l, 8);
                *(u32x2 *)C = hv;
                *(u32x2 *)(C + (int64_t)kRfTile * a.rows_pad) = lv;
            }
        }
}

// ---------------------------------------------------------------------------------------------
// FP64 products on the int8 matrix cores (Ozaki scheme).  An operand row r (length K) is scaled by
// 2^-e_r, e_r = ilogb(max_k |x_rk|) + 1, and cut into kOzSlices signed 7-bit integer slices:
// x_rk = 2^e_r (sum_s q_s 2^(-7 s) + rho), |q_s| <= 127, |rho| < 2^(-7 kOzSlices) (exact: scaling by a
// power of two, then repeated multiplication by 128 and truncation, all exact in FP64).  A product
// C = A B^T over two sliced operands is sum over slice pairs (s, t) of 2^(-7 (s + t)) (Q_s P_t^T): the
// integer products are exact in int32 (v_mfma_i32_16x16x64_i8); pairs with s + t <= kOzSlices + 1
// are kept (the dropped ones weigh <= 2^(-7 (kOzSlices + 2)) relative), their sums per diagonal
// d = s + t stay exact in int32 for K <= kOzMaxK, and the diagonals are combined in FP64, smallest
// first, then scaled by 2^(e_r + e'_c).  With eight slices (56 bits) the error of C_ij relative to
// sum_k |A_ik| |B_jk| matches a native FP64 product's (1.5e-16 against 1.1e-16 on Gaussian 2048-deep
// products; seven slices give 1.6e-14).
// Slice planes: kOzSlices int8 matrices (rows_pad x K_pad, row-major, K contiguous), zero padded, and
// one int32 exponent per row.
// ---------------------------------------------------------------------------------------------
constexpr int kOzSlices = 8;
constexpr int kOzTile = 64;      // workgroup tile (rows and columns); also the row / K padding unit
constexpr int64_t kOzMaxK = 16384;  // exact int32 diagonal sums: K * kOzSlices * 127^2 < 2^31
static inline int64_t oz_pad(int64_t x) { return (x + kOzTile - 1) / kOzTile * kOzTile; }
typedef int i32x4 __attribute__((ext_vector_type(4)));

// one wavefront per row: row r of X (element (r, k) at X[r sr + k sk], k < K) -> slices and exponent;
// rows r >= R (up to the grid's padded count) and columns k >= K get zero slices.  grid (Rpad / 4,
// blocks), 256 lanes; blockIdx.y advances X by dX doubles, the planes by dP bytes, the exponents by dE.
__global__ void __launch_bounds__(256) k_oz_slice(const double *X, int64_t sr, int64_t sk, int64_t R, int64_t K,
                                                  int64_t Rpad, int64_t Kpad, int8_t *planes, int32_t *expo, int64_t dX,
                                                  int64_t dP, int64_t dE) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= Rpad) return;
    X += blockIdx.y * dX;
    planes += blockIdx.y * dP;
    expo += blockIdx.y * dE;
    const bool live = r < R;
    double mx = 0.0;
    if (live)
        for (int64_t k = lane; k < K; k += 64) mx = fmax(mx, fabs(X[r * sr + k * sk]));
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int e = mx > 0.0 ? ilogb(mx) + 1 : 0;
    if (lane == 0) expo[r] = e;
    const int64_t plane = Rpad * Kpad;
    for (int64_t k0 = 16 * lane; k0 < Kpad; k0 += 1024) {
        int8_t q[kOzSlices][16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int64_t k = k0 + t;
            double x = (live && k < K) ? ldexp(X[r * sr + k * sk], -e) : 0.0;
#pragma unroll
            for (int s = 0; s < kOzSlices; ++s) {
                x *= 128.0;
                const double qi = trunc(x);
                q[s][t] = (int8_t)qi;
                x -= qi;
            }
        }
#pragma unroll
        for (int s = 0; s < kOzSlices; ++s) {
            i32x4 v;
            memcpy(&v, q[s], 16);
            *(i32x4 *)(planes + s * plane + r * Kpad + k0) = v;
        }
    }
}

struct OzArgs {
    const int8_t *A, *B;      // slice planes: A (rows_pad x K), B (cols_pad x K)
    const int32_t *eA, *eB;   // row exponents of A and of B
    double *C;                // C (rows x cols) at C[i ldc + j] (trans = 0) or C[j ldc + i] (trans = 1)
    int64_t K, rows, cols, rows_pad, cols_pad, ldc;
    int64_t dA, dB, dEA, dEB, dC;  // per-block strides (bytes for planes, ints for exponents, doubles for C)
    int32_t nb, trans;
};

// C = A B^T, 64 x 64 tile per workgroup (4 waves, 2 x 2, 32 x 32 each); tiles of a block numbered
// contiguously per XCD.  Per 64-deep K step: both operands' slice tiles in LDS (2 x 8 x 4 KiB), then
// for every wave's 2 x 2 16x16 tiles and every kept slice pair one int8 MFMA into that pair's
// diagonal accumulator.
__global__ void __launch_bounds__(256) k_oz_nt(OzArgs a) {
    const int64_t tr = a.rows_pad / kOzTile, tc = a.cols_pad / kOzTile, tiles = tr * tc, total = tiles * a.nb;
    const int64_t per = (total + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= total) return;
    const int64_t blk = t / tiles, rem = t - blk * tiles, m0 = (rem / tc) * kOzTile, n0 = (rem % tc) * kOzTile;
    const int8_t *A = a.A + blk * a.dA + m0 * a.K;
    const int8_t *B = a.B + blk * a.dB + n0 * a.K;
    const int64_t pa = a.rows_pad * a.K, pb = a.cols_pad * a.K;
    // LDS rows of 64 bytes; the 16-byte chunk q of row r sits at chunk q ^ (r & 3), so the 16 rows a
    // fragment read touches at one chunk index spread over four bank groups
    __shared__ __attribute__((aligned(16))) int8_t As[kOzSlices][kOzTile][kOzTile];
    __shared__ __attribute__((aligned(16))) int8_t Bs[kOzSlices][kOzTile][kOzTile];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    i32x4 acc[kOzSlices][2][2];
#pragma unroll
    for (int d = 0; d < kOzSlices; ++d)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[d][i][j] = i32x4{0, 0, 0, 0};
    // 16-byte chunk c of a slice tile: row c / 4, bytes 16 (c % 4); 256 chunks per slice tile
    const int crow = threadIdx.x >> 2, ccol = 16 * (threadIdx.x & 3), cswz = 16 * ((threadIdx.x & 3) ^ (crow & 3));
    for (int64_t k0 = 0; k0 < a.K; k0 += kOzTile) {
        i32x4 ra[kOzSlices], rb[kOzSlices];
#pragma unroll
        for (int s = 0; s < kOzSlices; ++s) {
            ra[s] = *(const i32x4 *)(A + s * pa + crow * a.K + k0 + ccol);
            rb[s] = *(const i32x4 *)(B + s * pb + crow * a.K + k0 + ccol);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kOzSlices; ++s) {
            *(i32x4 *)&As[s][crow][cswz] = ra[s];
            *(i32x4 *)&Bs[s][crow][cswz] = rb[s];
        }
        __syncthreads();
        i32x4 fa[kOzSlices][2], fb[kOzSlices][2];
#pragma unroll
        for (int s = 0; s < kOzSlices; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ra_ = wm + 16 * i + lane % 16, rb_ = wn + 16 * i + lane % 16;
                fa[s][i] = *(const i32x4 *)&As[s][ra_][16 * ((lane / 16) ^ (ra_ & 3))];
                fb[s][i] = *(const i32x4 *)&Bs[s][rb_][16 * ((lane / 16) ^ (rb_ & 3))];
            }
#pragma unroll
        for (int s = 0; s < kOzSlices; ++s)
#pragma unroll
            for (int u = 0; u < kOzSlices - s; ++u)  // pair (s + 1, u + 1), diagonal index s + u
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[s + u][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[s][i], fb[u][j], acc[s + u][i][j], 0, 0, 0);
    }
    // result register v of lane l: C[4 (l / 16) + v][l % 16] of each 16 x 16 tile
    const int32_t *eA = a.eA + blk * a.dEA, *eB = a.eB + blk * a.dEB;
    double *C = a.C + blk * a.dC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t c = n0 + wn + 16 * j + lane % 16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = m0 + wm + 16 * i + 4 * (lane / 16) + v;
                if (r >= a.rows || c >= a.cols) continue;
                double sum = 0.0;
#pragma unroll
                for (int d = kOzSlices - 1; d >= 0; --d) sum += ldexp((double)acc[d][i][j][v], -7 * (d + 2));
                const double val = ldexp(sum, eA[r] + eB[c]);
                if (a.trans)
                    C[c * a.ldc + r] = val;
                else
                    C[r * a.ldc + c] = val;
            }
        }
}

struct Blk {
    gtn_svd_block b;
    int64_t p;
    bool small, left;
    double *ws;
    double *Y, *Z, *T, *Bm, *W, *G, *R, *V, *Vs, *lam, *inv, *keep, *X, *part, *Lg;
    // split-BF16 planes of the range finder (planes_of(opts) and a large block): M (mp x np), M^T
    // (np x mp), Y^T (kRfTile x mp), Z^T (kRfTile x np), each hi then lo
    __bf16 *Mp, *MTp, *YTp, *ZTp;
    int64_t mp, np;
    // int8 slices of the FP64-by-int8 range finder (oz_of(opts) and a large block): M (omp x onp),
    // M^T (onp x omp), Y^T (opp x omp), Z^T (opp x onp), kOzSlices planes each, with row exponents
    int8_t *Ms, *MTs, *YTs, *ZTs;
    int32_t *eM, *eMT, *eYT, *eZT;
    int64_t omp, onp, opp;
    double *Gm;  // m x m Gram matrix M M^T of the Gram-iteration range finder (gram_of(opts))
    int64_t bytes;
};

// range finder iterating on the FP64 Gram matrix M M^T (SVD mode, m <= n)
static bool gram_of(const gtn_svd_opts &o, const gtn_svd_block &b) {
    return o.mode == 0 && o.precision == GTN_PREC_F64_GRAM && b.m <= b.n;
}

// range-finder products on int8 slices (SVD mode, blocks up to kOzMaxK on either side)
static bool oz_of(const gtn_svd_opts &o, const gtn_svd_block &b) {
    return o.mode == 0 && o.precision == GTN_PREC_F64_INT8 && oz_pad(b.m) <= kOzMaxK && oz_pad(b.n) <= kOzMaxK;
}

// reduced range-finder iterations on split-BF16 planes (SVD mode, BF16X3 arithmetic)
static bool planes_of(const gtn_svd_opts &o) {
    return o.mode == 0 && o.precision == GTN_PREC_BF16X3 && o.iters > o.f64_tail;
}

static int pick_split(int64_t K) {
    int S = 1;
    while (S < 32 && K % (2 * S) == 0 && K / (2 * S) >= 64) S *= 2;
    return S;
}

static void plan_block(const gtn_svd_block &b, const gtn_svd_opts &o, Blk &P, double *ws) {
    P.b = b;
    int64_t mn = b.m < b.n ? b.m : b.n;
    P.small = mn <= kPMax;
    P.left = b.m <= b.n;
    int64_t p;
    if (P.small) {
        p = (o.mode == 1) ? b.m : mn;
    } else {
        p = b.k + (o.oversample > 0 ? o.oversample : 0);
        if (p > kPMax) p = kPMax;
        if (p < b.k) p = b.k;
    }
    if (p & 1) ++p;  // Jacobi pairs need an even size; the pad row/column stays zero
    if (p < 2) p = 2;
    P.p = p;
    int64_t off = 0;
    auto take = [&](int64_t elems) {
        int64_t o2 = off;
        off += (elems + 31) / 32 * 32;
        return o2;
    };
    const int64_t big = b.m > b.n ? b.m : b.n;
    const int64_t kk = b.k > 0 ? b.k : 1;
    int64_t oY = take(big * p), oZ = take(big * p), oT = take(big * p), oB = take(p * big);
    int64_t oW = take(p * p), oG = take(p * p), oR = take(p * p), oV = take(p * p), oVs = take(p * p);
    int64_t ol = take(p), oi = take(p), ok = take(p), oX = take(big * kk);
    int64_t opart = take((int64_t)32 * p * p);  // up to 32 split-K partial Gram matrices
    int64_t olg = take((int64_t)kMaxSweeps * (p - 1) * (p + 1));  // Jacobi rotation log: (c, s) per pair, round ids
    P.mp = rf_pad(b.m);
    P.np = rf_pad(b.n);
    const bool planes = !P.small && planes_of(o);
    // bf16 planes, in doubles: 2 (hi, lo) x 2 (M, M^T) x mp np / 4, 2 x kRfTile x (mp or np) / 4
    int64_t oMp = planes ? take(P.mp * P.np) : 0, oYT = planes ? take(kRfTile * P.mp / 2) : 0,
            oZT = planes ? take(kRfTile * P.np / 2) : 0;
    const bool oz = !P.small && oz_of(o, b);
    P.omp = oz_pad(b.m);
    P.onp = oz_pad(b.n);
    P.opp = oz_pad(p);
    auto dbl = [](int64_t bytes) { return (bytes + 7) / 8; };
    int64_t oMs = oz ? take(dbl(2 * kOzSlices * P.omp * P.onp)) : 0, oYs = oz ? take(dbl(kOzSlices * P.opp * P.omp)) : 0,
            oZs = oz ? take(dbl(kOzSlices * P.opp * P.onp)) : 0,
            oE = oz ? take(dbl(4 * (P.omp + P.onp + 2 * P.opp))) : 0;
    const bool gram = !P.small && gram_of(o, b);
    int64_t oGm = gram ? take(b.m * b.m) : 0;
    P.bytes = (off * 8 + 255) / 256 * 256;
    P.ws = ws;
    if (ws) {
        P.Y = ws + oY;
        P.Z = ws + oZ;
        P.T = ws + oT;
        P.Bm = ws + oB;
        P.W = ws + oW;
        P.G = ws + oG;
        P.R = ws + oR;
        P.V = ws + oV;
        P.Vs = ws + oVs;
        P.lam = ws + ol;
        P.inv = ws + oi;
        P.keep = ws + ok;
        P.X = ws + oX;
        P.part = ws + opart;
        P.Lg = ws + olg;
        P.Mp = planes ? (__bf16 *)(ws + oMp) : nullptr;
        P.MTp = planes ? P.Mp + 2 * P.mp * P.np : nullptr;
        P.YTp = planes ? (__bf16 *)(ws + oYT) : nullptr;
        P.ZTp = planes ? (__bf16 *)(ws + oZT) : nullptr;
        P.Ms = oz ? (int8_t *)(ws + oMs) : nullptr;
        P.MTs = oz ? P.Ms + kOzSlices * P.omp * P.onp : nullptr;
        P.YTs = oz ? (int8_t *)(ws + oYs) : nullptr;
        P.ZTs = oz ? (int8_t *)(ws + oZs) : nullptr;
        P.eM = oz ? (int32_t *)(ws + oE) : nullptr;
        P.eMT = oz ? P.eM + P.omp : nullptr;
        P.eYT = oz ? P.eMT + P.onp : nullptr;
        P.eZT = oz ? P.eYT + P.opp : nullptr;
        P.Gm = gram ? ws + oGm : nullptr;
    }
}

static gtn_gemm_desc gd(const double *A, int64_t sam, int64_t sak, const double *B, int64_t sbk, int64_t sbn, double *C,
                        int64_t scm, int64_t M, int64_t N, int64_t K, int64_t batch = 1, int64_t sab = 0, int64_t sbb = 0,
                        int64_t scb = 0) {
    gtn_gemm_desc d;
    memset(&d, 0, sizeof(d));
    d.A = A;
    d.B = B;
    d.C = C;
    d.M = M;
    d.N = N;
    d.K = K;
    d.batch = batch;
    d.sam = sam;
    d.sak = sak;
    d.sbk = sbk;
    d.sbn = sbn;
    d.scm = scm;
    d.scn = 1;
    d.sab = sab;
    d.sbb = sbb;
    d.scb = scb;
    d.alpha = 1.0;
    return d;
}

static int run(std::vector<gtn_gemm_desc> &g, int prec, hipStream_t s) {
    if (g.empty()) return GTN_OK;
    return gtn_gemm_grouped(g.data(), (int)g.size(), GTN_F64, prec, (void *)s);
}

// A run: consecutive blocks of equal geometry whose input / output matrices and workspaces advance
// by constant strides (in doubles), so that every stage can process the run in one launch.
struct Run {
    Blk *f;          // first block
    int n;           // blocks in the run (the batch of every launch)
    int first;       // index of the first block in the call (seeds)
    int64_t dM, dU, dS, dV, dws;
    std::vector<Blk *> mem;  // every block of the run, f first
};

static bool delta_of(const void *a, const void *b, int64_t &d) {
    if ((a == nullptr) != (b == nullptr)) return false;
    if (!a) {
        d = 0;
        return true;
    }
    int64_t diff = (int64_t)((const char *)b - (const char *)a);
    if (diff % 8) return false;
    d = diff / 8;
    return true;
}

static std::vector<Run> make_runs(std::vector<Blk *> &bs, std::vector<int> &index) {
    std::vector<Run> runs;
    for (size_t i = 0; i < bs.size(); ++i) {
        Blk *b = bs[i];
        if (!runs.empty()) {
            Run &r = runs.back();
            Blk *a = r.f;
            const gtn_svd_block &x = a->b, &y = b->b;
            bool same = x.m == y.m && x.n == y.n && x.k == y.k && x.lda == y.lda && x.ldu == y.ldu &&
                        x.ldv == y.ldv && a->p == b->p && a->small == b->small && a->left == b->left &&
                        a->bytes == b->bytes && r.n < 65535;
            int64_t dM, dU, dS, dV, dws;
            // the previous block of the run is f + (n-1) strides; compare with the last one
            Blk *last = bs[i - 1];
            if (same && delta_of(last->b.M, y.M, dM) && delta_of(last->b.U, y.U, dU) && delta_of(last->b.S, y.S, dS) &&
                delta_of(last->b.Vt, y.Vt, dV) && delta_of(last->ws, b->ws, dws)) {
                if (r.n == 1) {
                    r.dM = dM, r.dU = dU, r.dS = dS, r.dV = dV, r.dws = dws;
                    ++r.n;
                    r.mem.push_back(b);
                    continue;
                }
                if (dM == r.dM && dU == r.dU && dS == r.dS && dV == r.dV && dws == r.dws) {
                    ++r.n;
                    r.mem.push_back(b);
                    continue;
                }
            }
        }
        Run r;
        r.f = b;
        r.n = 1;
        r.first = index[i];
        r.dM = r.dU = r.dS = r.dV = r.dws = 0;
        r.mem.push_back(b);
        runs.push_back(r);
    }
    return runs;
}

// range checks (dry run) of a strided family of n ranges
#define R8N(p, elems, delta, n, what)                                                     \
    do {                                                                                    \
        R8((p), (elems), (what));                                                           \
        if ((n) > 1) R8((p) + (int64_t)((n) - 1) * (delta), (elems), (what));               \
    } while (0)

static dim3 grid2(int64_t elems, int n) { return dim3(grid_for(elems, 256, 4096), n); }

// gemm over a run: C_b = A_b B_b with the per-block strides of the run
static gtn_gemm_desc gdr(const double *A, int64_t sam, int64_t sak, const double *B, int64_t sbk, int64_t sbn, double *C,
                         int64_t scm, int64_t M, int64_t N, int64_t K, int n, int64_t dA, int64_t dB, int64_t dC) {
    return gd(A, sam, sak, B, sbk, sbn, C, scm, M, N, K, n, dA, dB, dC);
}

// split of a Gram reduction of length K: none for long runs (enough workgroups already), at most 8
// chunks for short runs (one grouped launch), up to 32 for a single block
static int split_of(const Run &r, int64_t K) {
    if (r.n >= 64) return 1;
    int S = pick_split(K);
    if (r.n > 1 && S > 8) S = 8;
    return S;
}

// descriptors of W_b = A_b A_b^T for every block b of a run, where A (p x K) has element (i, kk) at
// A[i * sam + kk * sak] and consecutive K-chunks start chunk_stride * (K / S) apart; with S > 1 the
// chunks' products go to the partials buffer `part` (S consecutive p x p matrices)
static void gram_descs(std::vector<gtn_gemm_desc> &g, const Run &r, const double *A, int64_t sam, int64_t sak, int64_t K,
                       int64_t chunk_stride, double *out, double *part, int64_t p, int S) {
    const int64_t ch = K / S;
    if (S == 1) {
        g.push_back(gdr(A, sam, sak, A, sak, sam, out, p, p, p, K, r.n, r.dws, r.dws, r.dws));
    } else if (r.n == 1) {
        g.push_back(gd(A, sam, sak, A, sak, sam, part, p, p, p, ch, S, ch * chunk_stride, ch * chunk_stride, p * p));
    } else {
        for (int c = 0; c < S; ++c)
            g.push_back(gdr(A + c * ch * chunk_stride, sam, sak, A + c * ch * chunk_stride, sak, sam, part + (int64_t)c * p * p,
                            p, p, p, ch, r.n, r.dws, r.dws, r.dws));
    }
}

static int sum_partials(std::vector<Run> &runs, int64_t (*K_of)(Blk &), double *(*out)(Blk &), hipStream_t s) {
    for (Run &r : runs) {
        Blk &f = *r.f;
        const int64_t p = f.p;
        const int S = split_of(r, K_of(f));
        if (S == 1) continue;
        R8N(f.part, (int64_t)S * p * p, r.dws, r.n, "gram partials");
        R8N(out(f), p * p, r.dws, r.n, "gram");
        LAUNCH(k_sum_partials, grid2(p * p, r.n), dim3(256), 0, s, (const double *)f.part, p * p, S, out(f), r.dws, r.dws);
    }
    return GTN_OK;
}

static int64_t rows_of(Blk &b) { return b.b.m; }
static int64_t cols_of(Blk &b) { return b.b.n; }

// Gram matrix W = Y^T Y (Y rows x p, ld p) of every block
static int gram_Y(std::vector<Run> &runs, hipStream_t s) {
    std::vector<gtn_gemm_desc> g;
    for (Run &r : runs) {
        Blk &f = *r.f;
        const int64_t rows = f.b.m, p = f.p;
        R8N(f.Y, rows * p, r.dws, r.n, "gram input");
        gram_descs(g, r, f.Y, 1, p, rows, p, f.W, f.part, p, split_of(r, rows));
    }
    int rc = run(g, GTN_PREC_F64, s);
    if (rc) return rc;
    return sum_partials(runs, rows_of, [](Blk &b) { return b.W; }, s);
}

// launch an LDS kernel over the runs: one uniform launch per multi-block run; single blocks are
// packed kMaxBatch to a launch
template <typename F>
static int small_launch(std::vector<Run> &runs, F fill, void (*kern)(SmallBatch), size_t shm_per_p2, size_t shm_extra,
                        SmallBatch proto, hipStream_t s) {
    std::vector<Run *> singles;
    auto go = [&](SmallBatch &sb, int64_t pmax, int grid) -> int {
        const int64_t pn = (pmax + kNB - 1) / kNB * kNB;  // kernels may pad p to whole 16-wide tiles
        size_t shm = shm_per_p2 * pn * pn + shm_extra * pn;
        GTN_REQUIRE(shm + 256 <= 160 * 1024, "gtn_svd_blocks: LDS request of %zu bytes exceeds 160 KiB", shm);
        if (dry_run()) return GTN_OK;
        // LDS beyond 64 KiB must be requested explicitly (gfx950 has 160 KiB per CU)
        GTN_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        LAUNCH(kern, dim3(grid), dim3(1024), shm, s, sb);
        return GTN_OK;
    };
    for (Run &r : runs) {
        if (r.n == 1) {
            singles.push_back(&r);
            continue;
        }
        SmallBatch sb = proto;
        sb.n = r.n;
        sb.uniform = 1;
        fill(*r.f, sb, 0);
        sb.d_in = sb.d_out = sb.d_aux = sb.d_lam = r.dws;
        const int64_t p2 = (int64_t)sb.p[0] * sb.p[0];
        R8N(sb.in[0], p2, r.dws, r.n, "small kernel input");
        R8N(sb.out[0], p2, r.dws, r.n, "small kernel output");
        if (sb.aux[0]) R8N(sb.aux[0], p2, r.dws, r.n, "small kernel aux");
        if (sb.lam[0]) R8N(sb.lam[0], sb.p[0], r.dws, r.n, "small kernel lambda");
        int rc = go(sb, sb.p[0], r.n);
        if (rc) return rc;
    }
    for (size_t base = 0; base < singles.size(); base += kMaxBatch) {
        SmallBatch sb = proto;
        sb.n = 0;
        sb.uniform = 0;
        int64_t pmax = 0;
        for (size_t i = base; i < singles.size() && i < base + kMaxBatch; ++i) {
            fill(*singles[i]->f, sb, sb.n);
            if (singles[i]->f->p > pmax) pmax = singles[i]->f->p;
            ++sb.n;
        }
        for (int i = 0; i < sb.n; ++i) {
            int64_t p2 = (int64_t)sb.p[i] * sb.p[i];
            R8(sb.in[i], p2, "small kernel input");
            R8(sb.out[i], p2, "small kernel output");
            if (sb.aux[i]) R8(sb.aux[i], p2, "small kernel aux");
            if (sb.lam[i]) R8(sb.lam[i], sb.p[i], "small kernel lambda");
        }
        int rc = go(sb, pmax, sb.n);
        if (rc) return rc;
    }
    return GTN_OK;
}

// orthonormalise Y (rows x p, ld p) of every block in place (CholeskyQR, `passes` times)
static int orthonormalise(std::vector<Run> &runs, int passes, hipStream_t s) {
    for (int pass = 0; pass < passes; ++pass) {
        int rc = gram_Y(runs, s);
        if (rc) return rc;
        SmallBatch proto;
        memset(&proto, 0, sizeof(proto));
        proto.shift_rel = pass == 0 ? 1e-13 : 0.0;
        rc = small_launch(
            runs,
            [](Blk &b, SmallBatch &sb, int i) {
                sb.p[i] = (int)b.p;
                sb.in[i] = b.W;
                sb.out[i] = b.R;
            },
            k_chol_inv, sizeof(double), sizeof(double) * (kNB + 1), proto, s);
        if (rc) return rc;
        std::vector<gtn_gemm_desc> g;
        for (Run &r : runs) {
            Blk &f = *r.f;
            int64_t rows = f.b.m, p = f.p;
            g.push_back(gdr(f.Y, p, 1, f.R, p, 1, f.T, p, rows, p, p, r.n, r.dws, r.dws, r.dws));
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        // the product is the new basis: Y and its scratch trade places in every block (same offsets
        // in every block's workspace, so the runs' strides hold)
        for (Run &r : runs)
            for (Blk *b : r.mem) std::swap(b->Y, b->T);
    }
    return GTN_OK;
}

static int jacobi(std::vector<Run> &runs, int by_abs, hipStream_t s) {
    SmallBatch proto;
    memset(&proto, 0, sizeof(proto));
    proto.by_abs = by_abs;
    proto.max_sweeps = kMaxSweeps;
    // eigenvalues, order and the rotation log (out = log, aux = info)
    int rc = small_launch(
        runs,
        [](Blk &b, SmallBatch &sb, int i) {
            sb.p[i] = (int)b.p;
            sb.in[i] = b.G;
            sb.out[i] = b.Lg;
            sb.aux[i] = b.V;
            sb.lam[i] = b.lam;
        },
        k_jacobi, sizeof(double), 2 * sizeof(double) + 3 * sizeof(int), proto, s);
    if (rc) return rc;
    // eigenvectors: one launch per run (uniform) or per single block
    for (Run &r : runs) {
        Blk &f = *r.f;
        const int64_t p = f.p;
        R8N(f.Lg, (int64_t)kMaxSweeps * (p - 1) * (p + 1), r.dws, r.n, "rotation log");
        R8N(f.V, p + 1, r.dws, r.n, "jacobi info");
        R8N(f.Vs, p * p, r.dws, r.n, "eigenvectors");
        SmallBatch sb;
        memset(&sb, 0, sizeof(sb));
        sb.n = r.n;
        sb.uniform = 1;
        sb.p[0] = (int)p;
        sb.in[0] = f.Lg;
        sb.out[0] = f.Vs;
        sb.aux[0] = f.V;
        sb.d_in = sb.d_out = sb.d_aux = r.dws;
        LAUNCH(k_jacobi_vectors, dim3((unsigned)p, r.n), dim3(64), 0, s, sb);
    }
    return GTN_OK;
}

// split-BF16 planes of every block's M and M^T (once per call)
static int rf_planes_of_blocks(std::vector<Run> &rl, hipStream_t s) {
    for (Run &r : rl) {
        Blk &f = *r.f;
        GTN_REQUIRE(f.Mp && f.mp / 64 <= 65535 && f.np / 64 <= 65535, "gtn_svd_blocks: no range-finder planes");
        R8N(f.b.M, ext(f.b.m, f.b.n, f.b.lda), r.dM, r.n, "block");
        R8N((const double *)f.Mp, f.mp * f.np, r.dws, r.n, "range-finder planes of M");
        LAUNCH(k_planes, dim3((unsigned)(f.np / 64), (unsigned)(f.mp / 64), r.n), dim3(256), 0, s, f.b.M, f.b.lda, f.b.m,
               f.b.n, f.mp, f.np, f.Mp, f.MTp, r.dM, 4 * r.dws);
    }
    return GTN_OK;
}

static dim3 rf_grid(int64_t rows_pad, int nb) {
    const int64_t total = rows_pad / kRfTile * nb;
    return dim3((unsigned)((total + 7) / 8 * 8));
}

// Y <- M Z (FP64) from the planes of M and of Z^T
static int rf_y_from_zt(std::vector<Run> &rl, hipStream_t s) {
    for (Run &r : rl) {
        Blk &f = *r.f;
        RfArgs a;
        a.A = f.Mp, a.B = f.ZTp, a.C = f.Y;
        a.K = f.np, a.plane_a = f.mp * f.np, a.plane_b = kRfTile * f.np;
        a.rows = f.b.m, a.rows_pad = f.mp, a.cols = f.p, a.ldc = f.p;
        a.dA = a.dB = 4 * r.dws, a.dC = r.dws, a.nb = r.n;
        R8N(f.Y, f.b.m * f.p, r.dws, r.n, "range-finder basis");
        LAUNCH(k_rf_nt<0>, rf_grid(f.mp, r.n), dim3(256), 0, s, a);
    }
    return GTN_OK;
}

// one reduced iteration on planes: Y^T planes <- Y; Z^T planes <- (M^T Y)^T; Y <- M Z (FP64 out)
static int rf_iteration_planes(std::vector<Run> &rl, hipStream_t s) {
    for (Run &r : rl) {
        Blk &f = *r.f;
        GTN_REQUIRE(f.p <= kRfTile, "gtn_svd_blocks: range-finder width above the plane tile");
        R8N(f.Y, f.b.m * f.p, r.dws, r.n, "range-finder basis");
        R8N((const double *)f.YTp, kRfTile * f.mp / 2, r.dws, r.n, "range-finder planes of Y");
        LAUNCH(k_planes, dim3(kRfTile / 64, (unsigned)(f.mp / 64), r.n), dim3(256), 0, s, (const double *)f.Y, f.p,
               f.b.m, f.p, f.mp, (int64_t)kRfTile, (__bf16 *)nullptr, f.YTp, r.dws, 4 * r.dws);
    }
    for (Run &r : rl) {
        Blk &f = *r.f;
        RfArgs a;
        a.A = f.MTp, a.B = f.YTp, a.C = f.ZTp;
        a.K = f.mp, a.plane_a = f.np * f.mp, a.plane_b = kRfTile * f.mp;
        a.rows = f.b.n, a.rows_pad = f.np, a.cols = f.p, a.ldc = 0;
        a.dA = a.dB = a.dC = 4 * r.dws, a.nb = r.n;
        R8N((const double *)f.ZTp, kRfTile * f.np / 2, r.dws, r.n, "range-finder planes of Z");
        LAUNCH(k_rf_nt<1>, rf_grid(f.np, r.n), dim3(256), 0, s, a);
    }
    return rf_y_from_zt(rl, s);
}

// int8 slices of every block's M (rows of M) and M^T (columns of M), once per call
static int oz_slices_of_blocks(std::vector<Run> &rl, hipStream_t s) {
    for (Run &r : rl) {
        Blk &f = *r.f;
        GTN_REQUIRE(f.Ms, "gtn_svd_blocks: no range-finder slices");
        R8N(f.b.M, ext(f.b.m, f.b.n, f.b.lda), r.dM, r.n, "block");
        R8N((const double *)f.Ms, (2 * kOzSlices * f.omp * f.onp + 7) / 8, r.dws, r.n, "slices of M");
        R8N((const double *)f.eM, (4 * (f.omp + f.onp + 2 * f.opp) + 7) / 8, r.dws, r.n, "slice exponents");
        LAUNCH(k_oz_slice, dim3((unsigned)(f.omp / 4), r.n), dim3(256), 0, s, f.b.M, f.b.lda, (int64_t)1, f.b.m, f.b.n,
               f.omp, f.onp, f.Ms, f.eM, r.dM, 8 * r.dws, 2 * r.dws);
        LAUNCH(k_oz_slice, dim3((unsigned)(f.onp / 4), r.n), dim3(256), 0, s, f.b.M, (int64_t)1, f.b.lda, f.b.n, f.b.m,
               f.onp, f.omp, f.MTs, f.eMT, r.dM, 8 * r.dws, 2 * r.dws);
    }
    return GTN_OK;
}

static dim3 oz_grid(int64_t rows_pad, int64_t cols_pad, int nb) {
    const int64_t total = rows_pad / kOzTile * (cols_pad / kOzTile) * nb;
    return dim3((unsigned)((total + 7) / 8 * 8));
}

// out (rows x p FP64, ld p) = (A planes, rows) x (slices of the p columns of X (K x p FP64, ld p))
static int oz_product(std::vector<Run> &rl, bool transposed_m, double *(Blk::*X), double *(Blk::*out), hipStream_t s) {
    for (Run &r : rl) {
        Blk &f = *r.f;
        const int64_t K = transposed_m ? f.b.m : f.b.n, Kp = transposed_m ? f.omp : f.onp;
        const int64_t rows = transposed_m ? f.b.n : f.b.m, rows_pad = transposed_m ? f.onp : f.omp;
        int8_t *Bs = transposed_m ? f.YTs : f.ZTs;
        int32_t *eB = transposed_m ? f.eYT : f.eZT;
        R8N(f.*X, K * f.p, r.dws, r.n, "range-finder operand");
        R8N(f.*out, rows * f.p, r.dws, r.n, "range-finder product");
        LAUNCH(k_oz_slice, dim3((unsigned)(f.opp / 4), r.n), dim3(256), 0, s, (const double *)(f.*X), (int64_t)1, f.p,
               f.p, K, f.opp, Kp, Bs, eB, r.dws, 8 * r.dws, 2 * r.dws);
        OzArgs a;
        a.A = transposed_m ? f.MTs : f.Ms, a.B = Bs;
        a.eA = transposed_m ? f.eMT : f.eM, a.eB = eB;
        a.C = f.*out;
        a.K = Kp, a.rows = rows, a.cols = f.p, a.rows_pad = rows_pad, a.cols_pad = f.opp, a.ldc = f.p;
        a.dA = a.dB = 8 * r.dws, a.dEA = a.dEB = 2 * r.dws, a.dC = r.dws;
        a.nb = r.n, a.trans = 0;
        LAUNCH(k_oz_nt, oz_grid(rows_pad, f.opp, r.n), dim3(256), 0, s, a);
    }
    return GTN_OK;
}

// the Gram matrices G = M M^T of every block (m x m, FP64): products over the upper block triangle
// in super-rows of S rows (grouped launches of at most 12 super-rows), then the lower part mirrored.
// S = 128 (one 128-row tile) computes (1 + 128 / m) / 2 of the full product: 53 % at m = 2048.
static int64_t gram_super_rows(int64_t m) {
    int64_t S = 128;
    while ((m + S - 1) / S > 16) S += 128;  // at most 16 super-rows per block
    return S;
}
static int gram_of_blocks(std::vector<Run> &rl, hipStream_t s) {
    std::vector<gtn_gemm_desc> g;
    for (Run &r : rl) {
        Blk &f = *r.f;
        GTN_REQUIRE(f.Gm, "gtn_svd_blocks: no Gram matrix");
        const int64_t m = f.b.m, S = gram_super_rows(m);
        R8N(f.Gm, m * m, r.dws, r.n, "Gram matrix");
        R8N(f.b.M, ext(f.b.m, f.b.n, f.b.lda), r.dM, r.n, "block");
        if (!g.empty() && g.size() + (size_t)((m + S - 1) / S) > 12) {  // a launch takes at most 12 groups
            int rc = run(g, GTN_PREC_F64, s);
            if (rc) return rc;
            g.clear();
        }
        for (int64_t r0 = 0; r0 < m; r0 += S) {
            const int64_t rows = m - r0 < S ? m - r0 : S;
            // G[r0.., r0..] = M[r0.., :] M[r0.., :]^T
            g.push_back(gdr(f.b.M + r0 * f.b.lda, f.b.lda, 1, f.b.M + r0 * f.b.lda, 1, f.b.lda, f.Gm + r0 * m + r0, m, rows,
                            m - r0, f.b.n, r.n, r.dM, r.dM, r.dws));
        }
    }
    int rc = run(g, GTN_PREC_F64, s);
    if (rc) return rc;
    for (Run &r : rl) {
        Blk &f = *r.f;
        const int64_t m = f.b.m, t = (m + 63) / 64;
        GTN_REQUIRE(t <= 65535, "gtn_svd_blocks: Gram matrix too large");
        LAUNCH(k_mirror_lower, dim3((unsigned)t, (unsigned)t, r.n), dim3(256), 0, s, f.Gm, m, m, gram_super_rows(m), r.dws);
    }
    return GTN_OK;
}

static int run_blocks(std::vector<Blk> &blocks, const gtn_svd_opts &o, hipStream_t s) {
    const int eig = o.mode == 1;
    std::vector<Blk *> small, large, all;
    std::vector<int> ismall, ilarge, iall;
    for (size_t i = 0; i < blocks.size(); ++i) {
        Blk &b = blocks[i];
        if (b.b.k <= 0 || b.b.m <= 0 || b.b.n <= 0) continue;
        all.push_back(&b);
        iall.push_back((int)i);
        if (b.small) {
            small.push_back(&b);
            ismall.push_back((int)i);
        } else {
            large.push_back(&b);
            ilarge.push_back((int)i);
        }
    }
    if (all.empty()) return GTN_OK;
    std::vector<Run> rs = make_runs(small, ismall), rl = make_runs(large, ilarge), ra = make_runs(all, iall);
    int rc;
    // ---- small blocks: G = Gram of the short side (svd) or the block itself (eig)
    if (!rs.empty()) {
        std::vector<gtn_gemm_desc> g;
        for (Run &r : rs) {
            Blk &f = *r.f;
            const gtn_svd_block &B = f.b;
            R8N(f.G, f.p * f.p, r.dws, r.n, "small gram");
            LAUNCH(k_fill, grid2(f.p * f.p, r.n), dim3(256), 0, s, f.G, f.p, f.p, f.p, 0.0, r.dws);
            if (eig) {
                GTN_REQUIRE(B.m <= f.p && B.n <= f.p, "gtn_svd_blocks: eig block larger than its work size");
                R8N(B.M, ext(B.m, B.n, B.lda), r.dM, r.n, "small eig input");
                LAUNCH(k_copy2d, grid2(B.m * B.n, r.n), dim3(256), 0, s, B.M, B.lda, f.G, f.p, B.m, B.n, r.dM, r.dws);
            } else if (f.left) {
                g.push_back(gdr(B.M, B.lda, 1, B.M, 1, B.lda, f.G, f.p, B.m, B.m, B.n, r.n, r.dM, r.dM, r.dws));  // M M^T
            } else {
                g.push_back(gdr(B.M, 1, B.lda, B.M, B.lda, 1, f.G, f.p, B.n, B.n, B.m, r.n, r.dM, r.dM, r.dws));  // M^T M
            }
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
    }
    // ---- large blocks: randomized subspace iteration
    if (!rl.empty()) {
        std::vector<gtn_gemm_desc> g;
        const bool planes = planes_of(o);
        if (planes) {
            rc = rf_planes_of_blocks(rl, s);
            if (rc) return rc;
        }
        bool oz = !eig && o.precision == GTN_PREC_F64_INT8;
        for (Run &r : rl) oz = oz && r.f->Ms != nullptr;
        if (oz) {
            rc = oz_slices_of_blocks(rl, s);
            if (rc) return rc;
        }
        bool gram = !eig && o.precision == GTN_PREC_F64_GRAM;
        for (Run &r : rl) gram = gram && r.f->Gm != nullptr;
        if (gram) {
            rc = gram_of_blocks(rl, s);
            if (rc) return rc;
        }
        for (Run &r : rl) {
            Blk &f = *r.f;
            R8N(f.Z, f.b.n * f.p, r.dws, r.n, "gaussian");
            R8N(f.b.M, ext(f.b.m, f.b.n, f.b.lda), r.dM, r.n, "block");
            LAUNCH(k_fill_gaussian, grid2(f.b.n * f.p, r.n), dim3(256), 0, s, f.Z, f.b.n, f.p, f.p,
                   o.seed * 0x9E3779B97F4A7C15ull + (uint64_t)r.first, r.dws);
            if (planes) {  // Z^T planes of the Gaussian block
                R8N((const double *)f.ZTp, kRfTile * f.np / 2, r.dws, r.n, "range-finder planes of Z");
                LAUNCH(k_planes, dim3(kRfTile / 64, (unsigned)(f.np / 64), r.n), dim3(256), 0, s, (const double *)f.Z,
                       f.p, f.b.n, f.p, f.np, (int64_t)kRfTile, (__bf16 *)nullptr, f.ZTp, r.dws, 4 * r.dws);
            } else if (!oz) {
                g.push_back(gdr(f.b.M, f.b.lda, 1, f.Z, f.p, 1, f.Y, f.p, f.b.m, f.p, f.b.n, r.n, r.dM, r.dws, r.dws));  // Y = M Z
            }
        }
        // FP64-by-int8 and the Gram iteration fall back to the alternating FP64 products where a
        // block has no slices (beyond kOzMaxK) or no Gram matrix (m > n); their first product is FP64
        const int gprec = (o.precision == GTN_PREC_F64_INT8 || o.precision == GTN_PREC_F64_GRAM) ? GTN_PREC_F64 : o.precision;
        rc = planes ? rf_y_from_zt(rl, s) : (oz ? oz_product(rl, false, &Blk::Z, &Blk::Y, s) : run(g, gprec, s));
        if (rc) return rc;
        rc = orthonormalise(rl, o.iters > 0 ? 1 : 2, s);
        if (rc) return rc;
        for (int it = 0; it < o.iters; ++it) {
            const int prec = it >= o.iters - o.f64_tail ? GTN_PREC_F64 : (oz ? o.precision : gprec);
            if (gram) {  // Y = G Y (into Z, which then becomes the basis)
                g.clear();
                for (Run &r : rl) {
                    Blk &f = *r.f;
                    R8N(f.Z, f.b.m * f.p, r.dws, r.n, "Gram iterate");
                    g.push_back(gdr(f.Gm, f.b.m, 1, f.Y, f.p, 1, f.Z, f.p, f.b.m, f.p, f.b.m, r.n, r.dws, r.dws, r.dws));
                }
                rc = run(g, GTN_PREC_F64, s);
                if (rc) return rc;
                for (Run &r : rl)
                    for (Blk *b : r.mem) std::swap(b->Y, b->Z);
                const int every = o.orth_every > 1 ? o.orth_every : 1;
                rc = (it + 1 == o.iters) ? orthonormalise(rl, 2, s)
                                         : ((it + 1) % every == 0 ? orthonormalise(rl, 1, s) : GTN_OK);
                if (rc) return rc;
                continue;
            }
            if (oz && prec != GTN_PREC_F64) {  // Z = M^T Y, Y = M Z on int8 slices
                rc = oz_product(rl, true, &Blk::Y, &Blk::Z, s);
                if (rc) return rc;
                rc = oz_product(rl, false, &Blk::Z, &Blk::Y, s);
                if (rc) return rc;
                const int every = o.orth_every > 1 ? o.orth_every : 1;
                rc = (it + 1 == o.iters) ? orthonormalise(rl, 2, s)
                                         : ((it + 1) % every == 0 ? orthonormalise(rl, 1, s) : GTN_OK);
                if (rc) return rc;
                continue;
            }
            if (planes && prec != GTN_PREC_F64) {
                rc = rf_iteration_planes(rl, s);
                if (rc) return rc;
                const int every = o.orth_every > 1 ? o.orth_every : 1;
                rc = (it + 1 == o.iters) ? orthonormalise(rl, 2, s)
                                         : ((it + 1) % every == 0 ? orthonormalise(rl, 1, s) : GTN_OK);
                if (rc) return rc;
                continue;
            }
            g.clear();
            for (Run &r : rl) {
                Blk &f = *r.f;
                if (eig)  // Z = M Y
                    g.push_back(gdr(f.b.M, f.b.lda, 1, f.Y, f.p, 1, f.Z, f.p, f.b.m, f.p, f.b.n, r.n, r.dM, r.dws, r.dws));
                else  // Z = M^T Y
                    g.push_back(gdr(f.b.M, 1, f.b.lda, f.Y, f.p, 1, f.Z, f.p, f.b.n, f.p, f.b.m, r.n, r.dM, r.dws, r.dws));
            }
            rc = run(g, prec, s);
            if (rc) return rc;
            if (!eig) {
                g.clear();
                for (Run &r : rl) {  // Y = M Z
                    Blk &f = *r.f;
                    g.push_back(gdr(f.b.M, f.b.lda, 1, f.Z, f.p, 1, f.Y, f.p, f.b.m, f.p, f.b.n, r.n, r.dM, r.dws, r.dws));
                }
                rc = run(g, prec, s);
                if (rc) return rc;
            } else {
                for (Run &r : rl) {
                    Blk &f = *r.f;
                    R8N(f.Z, f.b.m * f.p, r.dws, r.n, "eig iterate");
                    R8N(f.Y, f.b.m * f.p, r.dws, r.n, "eig basis");
                    LAUNCH(k_copy2d, grid2(f.b.m * f.p, r.n), dim3(256), 0, s, (const double *)f.Z, f.p, f.Y, f.p,
                           f.b.m, f.p, r.dws, r.dws);
                }
            }
            const int every = o.orth_every > 1 ? o.orth_every : 1;
            if (it + 1 == o.iters) {
                rc = orthonormalise(rl, 2, s);
            } else if ((it + 1) % every == 0) {
                rc = orthonormalise(rl, 1, s);
            }
            if (rc) return rc;
        }
        // Rayleigh-Ritz in FP64: Bm = Y^T M (p x n); G = Bm Bm^T (svd) or Bm Y (eig)
        g.clear();
        for (Run &r : rl) {
            Blk &f = *r.f;
            g.push_back(gdr(f.Y, 1, f.p, f.b.M, f.b.lda, 1, f.Bm, f.b.n, f.p, f.b.n, f.b.m, r.n, r.dws, r.dM, r.dws));
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        if (eig) {
            g.clear();
            for (Run &r : rl) {
                Blk &f = *r.f;
                g.push_back(gdr(f.Bm, f.b.n, 1, f.Y, f.p, 1, f.G, f.p, f.p, f.p, f.b.n, r.n, r.dws, r.dws, r.dws));
            }
            rc = run(g, GTN_PREC_F64, s);
            if (rc) return rc;
        } else {
            // G = Bm Bm^T (Bm p x n, ld n)
            g.clear();
            for (Run &r : rl) {
                Blk &f = *r.f;
                R8N(f.Bm, f.p * f.b.n, r.dws, r.n, "ritz input");
                gram_descs(g, r, f.Bm, f.b.n, 1, f.b.n, 1, f.G, f.part, f.p, split_of(r, f.b.n));
            }
            rc = run(g, GTN_PREC_F64, s);
            if (rc) return rc;
            rc = sum_partials(rl, cols_of, [](Blk &b) { return b.G; }, s);
            if (rc) return rc;
        }
    }
    // ---- p x p eigenproblems of every block at once
    rc = jacobi(ra, eig, s);
    if (rc) return rc;
    for (Run &r : ra) {
        Blk &f = *r.f;
        GTN_REQUIRE(f.b.k <= 256 && f.b.k <= f.p, "gtn_svd_blocks: k too large");
        if (!eig) continue;  // svd: the values come from the vectors' norms below
        R8N(f.lam, f.b.k, r.dws, r.n, "lambda");
        R8N(f.b.S, f.b.k, r.dS, r.n, "S");
        R8N(f.inv, f.b.k, r.dws, r.n, "inv");
        R8N(f.keep, f.b.k, r.dws, r.n, "keep");
        LAUNCH(k_sigma, dim3(1, r.n), dim3(256), 0, s, (const double *)f.lam, f.b.k, eig, f.b.S, f.inv, f.keep, r.dws, r.dS,
               r.dws);
    }
    // ---- singular vectors (svd: unnormalised, sigma_i times the unit vector, on the side computed
    // by a product with the block)
    std::vector<gtn_gemm_desc> g;
    for (Run &r : ra) {
        Blk &f = *r.f;
        const gtn_svd_block &B = f.b;
        const int64_t k = B.k, p = f.p;
        if (f.small) {
            if (eig || f.left) {
                GTN_REQUIRE(B.m <= p, "gtn_svd_blocks: small block rows exceed the work size");
                R8N(f.Vs, ext(B.m, k, p), r.dws, r.n, "eigenvectors");
                R8N(B.U, ext(B.m, k, B.ldu), r.dU, r.n, "U");
                LAUNCH(k_copy2d, grid2(B.m * k, r.n), dim3(256), 0, s, (const double *)f.Vs, p, B.U, B.ldu, B.m, k, r.dws,
                       r.dU);
                if (!eig)  // Vt = U^T M
                    g.push_back(gdr(f.Vs, 1, p, B.M, B.lda, 1, B.Vt, B.ldv, k, B.n, B.m, r.n, r.dws, r.dM, r.dV));
            } else {
                GTN_REQUIRE(B.n <= p, "gtn_svd_blocks: small block columns exceed the work size");
                R8N(f.Vs, ext(B.n, k, p), r.dws, r.n, "eigenvectors");
                R8N(B.Vt, ext(k, B.n, B.ldv), r.dV, r.n, "Vt");
                LAUNCH(k_transpose2d, grid2(B.n * k, r.n), dim3(256), 0, s, (const double *)f.Vs, p, B.Vt, B.ldv, B.n, k,
                       r.dws, r.dV);  // Vt = Vs[:, :k]^T
                g.push_back(gdr(B.M, B.lda, 1, f.Vs, p, 1, B.U, B.ldu, B.m, k, B.n, r.n, r.dM, r.dws, r.dU));  // U = M V
            }
        } else {
            g.push_back(gdr(f.Y, p, 1, f.Vs, p, 1, B.U, B.ldu, B.m, k, p, r.n, r.dws, r.dws, r.dU));  // U = Y Vs
            if (!eig)  // Vt = Vs^T Bm
                g.push_back(gdr(f.Vs, 1, p, f.Bm, B.n, 1, B.Vt, B.ldv, k, B.n, p, r.n, r.dws, r.dws, r.dV));
        }
    }
    rc = run(g, GTN_PREC_F64, s);
    if (rc) return rc;
    if (!eig) {
        // sigma_i = norm of the unnormalised vector i: row i of Vt, or column i of U = M V
        for (Run &r : ra) {
            Blk &f = *r.f;
            const gtn_svd_block &B = f.b;
            const int64_t k = B.k;
            R8N(f.b.S, k, r.dS, r.n, "S");
            R8N(f.inv, k, r.dws, r.n, "inv");
            R8N(f.keep, k, r.dws, r.n, "keep");
            if (f.small && !f.left) {
                R8N(B.U, ext(B.m, k, B.ldu), r.dU, r.n, "U");
                LAUNCH(k_sigma_norms, dim3(1, r.n), dim3(256), 0, s, (const double *)B.U, (int)k, B.m, (int64_t)1, B.ldu,
                       B.S, f.inv, f.keep, r.dU, r.dS, r.dws);
            } else {
                R8N(B.Vt, ext(k, B.n, B.ldv), r.dV, r.n, "Vt");
                LAUNCH(k_sigma_norms, dim3(1, r.n), dim3(256), 0, s, (const double *)B.Vt, (int)k, B.n, B.ldv, (int64_t)1,
                       B.S, f.inv, f.keep, r.dV, r.dS, r.dws);
            }
        }
    }
    for (Run &r : ra) {
        Blk &f = *r.f;
        const gtn_svd_block &B = f.b;
        const int64_t k = B.k;
        if (!eig) {
            bool u_from_m = f.small && !f.left;
            R8N(B.U, ext(B.m, k, B.ldu), r.dU, r.n, "U");
            R8N(B.Vt, ext(k, B.n, B.ldv), r.dV, r.n, "Vt");
            R8N(f.inv, k, r.dws, r.n, "inv");
            R8N(f.keep, k, r.dws, r.n, "keep");
            if (u_from_m) {
                LAUNCH(k_scale_cols, grid2(B.m * k, r.n), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu, (const double *)f.inv,
                       r.dU, r.dws);
                LAUNCH(k_scale_rows, grid2(k * B.n, r.n), dim3(256), 0, s, B.Vt, (int)k, B.n, B.ldv,
                       (const double *)f.keep, r.dV, r.dws);
            } else {
                LAUNCH(k_scale_rows, grid2(k * B.n, r.n), dim3(256), 0, s, B.Vt, (int)k, B.n, B.ldv, (const double *)f.inv,
                       r.dV, r.dws);
                LAUNCH(k_scale_cols, grid2(B.m * k, r.n), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu,
                       (const double *)f.keep, r.dU, r.dws);
            }
            LAUNCH(k_gauge, dim3(k, r.n), dim3(256), 0, s, B.U, B.m, B.ldu, B.Vt, B.n, B.ldv, r.dU, r.dV);
        } else {
            R8N(B.U, ext(B.m, k, B.ldu), r.dU, r.n, "U");
            R8N(f.keep, k, r.dws, r.n, "keep");
            LAUNCH(k_scale_cols, grid2(B.m * k, r.n), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu, (const double *)f.keep,
                   r.dU, r.dws);
            LAUNCH(k_gauge, dim3(k, r.n), dim3(256), 0, s, B.U, B.m, B.ldu, (double *)nullptr, (int64_t)0, (int64_t)0, r.dU,
                   (int64_t)0);
            if (B.Vt) {
                R8N(B.Vt, ext(k, B.m, B.ldv), r.dV, r.n, "Vt");
                LAUNCH(k_transpose2d, grid2(B.m * k, r.n), dim3(256), 0, s, (const double *)B.U, B.ldu, B.Vt, B.ldv, B.m, k,
                       r.dU, r.dV);
            }
        }
    }
    return GTN_OK;
}

}  // namespace gtn

using namespace gtn;

extern "C" int64_t gtn_svd_workspace_bytes(const gtn_svd_block *blocks, int n_blocks, const gtn_svd_opts *opts) {
    int64_t total = 0;
    for (int i = 0; i < n_blocks; ++i) {
        Blk P;
        plan_block(blocks[i], *opts, P, nullptr);
        total += P.bytes;
    }
    return total;
}

extern "C" int gtn_svd_blocks(const gtn_svd_block *blocks, int n_blocks, const gtn_svd_opts *opts, void *workspace,
                              int64_t workspace_bytes, void *stream) {
    GTN_REQUIRE(opts && (n_blocks == 0 || blocks), "gtn_svd_blocks: null argument");
    GTN_REQUIRE(opts->mode == 0 || opts->mode == 1, "gtn_svd_blocks: mode must be 0 (svd) or 1 (eig)");
    GTN_REQUIRE(opts->iters >= 0 && opts->iters <= 1000, "gtn_svd_blocks: iters out of range");
    GTN_REQUIRE(opts->f64_tail >= 0, "gtn_svd_blocks: negative f64_tail");
    int64_t need = gtn_svd_workspace_bytes(blocks, n_blocks, opts);
    if (need > workspace_bytes || (need > 0 && !workspace)) {
        set_error("gtn_svd_blocks: workspace too small (%lld < %lld bytes)", (long long)workspace_bytes, (long long)need);
        return GTN_ENOMEM;
    }
    std::vector<Blk> bs(n_blocks);
    char *base = (char *)workspace;
    for (int i = 0; i < n_blocks; ++i) {
        const gtn_svd_block &b = blocks[i];
        GTN_REQUIRE(b.m >= 0 && b.n >= 0 && b.k >= 0, "gtn_svd_blocks: negative size");
        int64_t mn = b.m < b.n ? b.m : b.n;
        GTN_REQUIRE(b.k <= mn || b.k == 0, "gtn_svd_blocks: k=%d exceeds min(m,n)=%lld", b.k, (long long)mn);
        if (b.k > 0) {
            GTN_REQUIRE(b.M && b.U && b.S, "gtn_svd_blocks: null matrix pointer");
            GTN_REQUIRE(opts->mode == 1 || b.Vt, "gtn_svd_blocks: null Vt");
            GTN_REQUIRE(opts->mode != 1 || b.m == b.n, "gtn_svd_blocks: eig needs a square block");
            GTN_REQUIRE(b.lda >= b.n && b.ldu >= b.k && (b.Vt == nullptr || b.ldv >= b.n), "gtn_svd_blocks: bad leading dims");
        }
        plan_block(b, *opts, bs[i], (double *)base);
        base += bs[i].bytes;
    }
    return run_blocks(bs, *opts, as_stream(stream));
}
