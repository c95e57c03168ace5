// Truncated SVD / Hermitian eig of a batch of parity blocks (replaces BlockSVD / SortedSVD,
// __init__.py:3931-4029, and BlockEig / SortedEig, :4314-4422).
//
// Large blocks (min(m, n) > kPMax): randomized subspace iteration with block width p, CholeskyQR
// re-orthonormalisation once per iteration (shifted, and repeated on the last pass), then an FP64
// Rayleigh-Ritz step whose p x p symmetric eigenproblem is solved by cyclic Jacobi in LDS.
// Small blocks (min(m, n) <= kPMax): the Gram matrix of the short side (or the Hermitian block
// itself in eig mode) is diagonalised directly by the same Jacobi kernel.
// Every stage processes all blocks of the call together: GEMMs are grouped (one group per block)
// and the LDS kernels run one workgroup per block.
#include "gtn_common.h"
#include <string.h>
#include <math.h>
#include <vector>

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

// ---------------------------------------------------------------------------------------------
// counter-based Gaussian fill: a 64-bit mix of (seed, index) -> two uniforms -> Box-Muller
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void k_fill_gaussian(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed) {
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

__global__ void k_sum_partials(const double *part, int64_t count, int S, double *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int b = 0; b < S; ++b) s += part[b * count + i];
        out[i] = s;
    }
}

__global__ void k_copy2d(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols) {
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        dst[r * ldd + c] = src[r * lds + c];
    }
}
// dst (cols x rows) = src (rows x cols)^T
__global__ void k_transpose2d(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows, int64_t cols) {
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = e / rows, r = e - c * rows;
        dst[c * ldd + r] = src[r * lds + c];
    }
}
__global__ void k_fill(double *x, int64_t rows, int64_t cols, int64_t ld, double v) {
    int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        x[r * ld + c] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// LDS kernels, one workgroup per block
// ---------------------------------------------------------------------------------------------
struct SmallBatch {
    int n;
    int p[kMaxBatch];
    const double *in[kMaxBatch];
    double *out[kMaxBatch];
    double *aux[kMaxBatch];
    double *lam[kMaxBatch];
    double shift_rel;
    int by_abs;
    int max_sweeps;
};

__device__ double wg_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// Cholesky W = R^T R (W p x p SPD, shifted by shift_rel * trace / p), then out = R^{-1} (full
// p x p, zero below the diagonal).  R and R^{-1} are kept in LDS in packed upper-triangular form
// (2 x p(p+1)/2 doubles, 132 KiB at p = 128).  Pivots that fall to <= 0 are replaced by a tiny
// positive value so that rank-deficient inputs still give finite columns (they are removed later
// by the rank rule).
__device__ __forceinline__ int up_idx(int i, int j, int p) { return i * p - (i * (i - 1)) / 2 + (j - i); }

__global__ void __launch_bounds__(1024) k_chol_inv(SmallBatch B) {
    HIP_DYNAMIC_SHARED(double, sm)
    __shared__ double red[16];
    const int bi = blockIdx.x;
    const int p = B.p[bi];
    const double *W = B.in[bi];
    double *Rinv = B.out[bi];
    const int np = p * (p + 1) / 2;
    double *A = sm;       // packed upper R
    double *X = sm + np;  // packed upper R^{-1}
    const int tid = threadIdx.x, nt = blockDim.x;
    double tr = 0;
    for (int e = tid; e < p * p; e += nt) {
        int i = e / p, j = e - i * p;
        if (j >= i) A[up_idx(i, j, p)] = W[e];
        if (i == j) tr += W[e];
    }
    tr = wg_sum(tr, red);
    const double shift = B.shift_rel * tr / p;
    const double floor_v = 1e-300 + 1e-30 * (tr / p);
    for (int i = tid; i < p; i += nt) A[up_idx(i, i, p)] += shift;
    __syncthreads();
    for (int k = 0; k < p; ++k) {
        if (tid == 0) {
            double d = A[up_idx(k, k, p)];
            A[up_idx(k, k, p)] = sqrt(d > floor_v ? d : floor_v);
        }
        __syncthreads();
        const double rkk = A[up_idx(k, k, p)];
        for (int j = k + 1 + tid; j < p; j += nt) A[up_idx(k, j, p)] /= rkk;
        __syncthreads();
        const int rem = p - k - 1;
        for (int e = tid; e < rem * rem; e += nt) {
            int i = k + 1 + e / rem, j = k + 1 + e % rem;
            if (j >= i) A[up_idx(i, j, p)] -= A[up_idx(k, i, p)] * A[up_idx(k, j, p)];
        }
        __syncthreads();
    }
    // X = R^{-1} by back substitution, one column per thread
    for (int c = tid; c < p; c += nt) {
        for (int i = c; i >= 0; --i) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int j = i + 1; j <= c; ++j) s -= A[up_idx(i, j, p)] * X[up_idx(j, c, p)];
            X[up_idx(i, c, p)] = s / A[up_idx(i, i, p)];
        }
    }
    __syncthreads();
    for (int e = tid; e < p * p; e += nt) {
        int i = e / p, j = e - i * p;
        Rinv[e] = (j >= i) ? X[up_idx(i, j, p)] : 0.0;
    }
}

// Cyclic Jacobi on a symmetric p x p matrix (p even) in LDS.  Round-robin pairing gives p/2
// disjoint rotations per round and every pair once per sweep.  Eigenvectors accumulate in
// global memory (aux, p x p, column e = eigenvector e); on exit lam and the columns of `out` are
// sorted by decreasing lambda (or |lambda| when by_abs).
__global__ void __launch_bounds__(1024) k_jacobi(SmallBatch B) {
    HIP_DYNAMIC_SHARED(double, sm)
    __shared__ double red[16];
    const int bi = blockIdx.x;
    const int p = B.p[bi];
    const double *Cin = B.in[bi];
    double *V = B.aux[bi];
    double *Vout = B.out[bi];
    double *lam = B.lam[bi];
    double *A = sm;
    double *cs = sm + p * p;
    int *pr = (int *)(cs + p);
    int *order = pr + p;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int h = p / 2;
    double nrm = 0;
    for (int i = tid; i < p * p; i += nt) {
        double v = Cin[i];
        A[i] = v;
        nrm += v * v;
        V[i] = (i / p == i % p) ? 1.0 : 0.0;
    }
    nrm = wg_sum(nrm, red);
    const double tol = 1e-32 * nrm;
    for (int sweep = 0; sweep < B.max_sweeps; ++sweep) {
        double off = 0;
        for (int i = tid; i < p * p; i += nt)
            if (i / p != i % p) off += A[i] * A[i];
        off = wg_sum(off, red);
        if (off <= tol) break;
        for (int round = 0; round < p - 1; ++round) {
            for (int t = tid; t < h; t += nt) {
                int a = (t == 0) ? 0 : ((t - 1 + round) % (p - 1)) + 1;
                int b = ((p - 2 - t + round) % (p - 1)) + 1;
                int i = a < b ? a : b, j = a < b ? b : a;
                double aii = A[i * p + i], ajj = A[j * p + j], aij = A[i * p + j];
                double c = 1.0, s = 0.0;
                if (fabs(aij) > 1e-300) {
                    double tau = (ajj - aii) / (2.0 * aij);
                    double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                }
                cs[t] = c;
                cs[h + t] = s;
                pr[t] = i;
                pr[h + t] = j;
            }
            __syncthreads();
            for (int e = tid; e < h * p; e += nt) {  // rows: A <- J^T A
                int t = e / p, col = e - t * p;
                int i = pr[t], j = pr[h + t];
                double c = cs[t], s = cs[h + t];
                double ai = A[i * p + col], aj = A[j * p + col];
                A[i * p + col] = c * ai - s * aj;
                A[j * p + col] = s * ai + c * aj;
            }
            __syncthreads();
            for (int e = tid; e < h * p; e += nt) {  // columns: A <- A J, V <- V J
                int t = e / p, row = e - t * p;
                int i = pr[t], j = pr[h + t];
                double c = cs[t], s = cs[h + t];
                double ai = A[row * p + i], aj = A[row * p + j];
                A[row * p + i] = c * ai - s * aj;
                A[row * p + j] = s * ai + c * aj;
                double vi = V[row * p + i], vj = V[row * p + j];
                V[row * p + i] = c * vi - s * vj;
                V[row * p + j] = s * vi + c * vj;
            }
            __syncthreads();
            for (int t = tid; t < h; t += nt) {
                A[pr[t] * p + pr[h + t]] = 0.0;
                A[pr[h + t] * p + pr[t]] = 0.0;
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < p; e += nt) {
        double ke = A[e * p + e];
        if (B.by_abs) ke = fabs(ke);
        int rank = 0;
        for (int f = 0; f < p; ++f) {
            double kf = A[f * p + f];
            if (B.by_abs) kf = fabs(kf);
            if (kf > ke || (kf == ke && f < e)) ++rank;
        }
        order[rank] = e;
    }
    __syncthreads();
    for (int i = tid; i < p; i += nt) lam[i] = A[order[i] * p + order[i]];
    for (int i = tid; i < p * p; i += nt) {
        int r = i / p, c = i - r * p;
        Vout[i] = V[r * p + order[c]];
    }
}

// ---------------------------------------------------------------------------------------------
// finishing kernels
// ---------------------------------------------------------------------------------------------
// sigma_i = sqrt(max(lam_i, 0)) (svd) or lam_i (eig); the rank rule |s_i| / (|s_0| + 1e-14) > 1e-14
// of __init__.py:3939-3941 gives keep_i in {0, 1}; inv_i = keep_i / |s_i|.
__global__ void k_sigma(const double *lam, int k, int eig_mode, double *S, double *inv, double *keep) {
    int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= k) return;
    double s0 = eig_mode ? fabs(lam[0]) : sqrt(fmax(lam[0], 0.0));
    double s = eig_mode ? lam[i] : sqrt(fmax(lam[i], 0.0));
    bool kp = fabs(s / (s0 + 1e-14)) > 1e-14;
    S[i] = kp ? s : 0.0;
    inv[i] = kp ? 1.0 / fabs(s) : 0.0;
    keep[i] = kp ? 1.0 : 0.0;
}
__global__ void k_scale_rows(double *X, int k, int64_t n, int64_t ld, const double *f) {
    int64_t total = (int64_t)k * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / n, c = e - r * n;
        X[r * ld + c] *= f[r];
    }
}
__global__ void k_scale_cols(double *X, int64_t m, int k, int64_t ld, const double *f) {
    int64_t total = m * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / k, c = e - r * k;
        X[r * ld + c] *= f[c];
    }
}
// gauge: column j of U gets its largest-|.| entry (first in row order) positive; the matching row
// of Vt (when given) flips with it.  One workgroup per column.
__global__ void k_gauge(double *U, int64_t m, int64_t ldu, double *Vt, int64_t n, int64_t ldv) {
    const int j = blockIdx.x;
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
struct Blk {
    gtn_svd_block b;
    int64_t p;
    bool small, left;
    double *ws;
    double *Y, *Z, *T, *Bm, *W, *G, *R, *V, *Vs, *lam, *inv, *keep, *X, *part;
    int64_t bytes;
};

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

// launch an LDS kernel over the blocks, kMaxBatch at a time
template <typename F>
static int small_launch(std::vector<Blk *> &bs, F fill, void (*kern)(SmallBatch), size_t shm_per_p2, size_t shm_extra,
                        SmallBatch proto, hipStream_t s) {
    for (size_t base = 0; base < bs.size(); base += kMaxBatch) {
        SmallBatch sb = proto;
        sb.n = 0;
        int64_t pmax = 0;
        for (size_t i = base; i < bs.size() && i < base + kMaxBatch; ++i) {
            fill(*bs[i], sb, sb.n);
            if (bs[i]->p > pmax) pmax = bs[i]->p;
            ++sb.n;
        }
        size_t shm = shm_per_p2 * pmax * pmax + shm_extra * pmax;
        GTN_REQUIRE(shm + 256 <= 160 * 1024, "gtn_svd_blocks: LDS request of %zu bytes exceeds 160 KiB", shm);
        for (int i = 0; i < sb.n; ++i) {
            int64_t p2 = (int64_t)sb.p[i] * sb.p[i];
            R8(sb.in[i], p2, "small kernel input");
            R8(sb.out[i], p2, "small kernel output");
            if (sb.aux[i]) R8(sb.aux[i], p2, "small kernel aux");
            if (sb.lam[i]) R8(sb.lam[i], sb.p[i], "small kernel lambda");
        }
        if (dry_run()) continue;
        // LDS beyond 64 KiB must be requested explicitly (gfx950 has 160 KiB per CU)
        GTN_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        LAUNCH(kern, dim3(sb.n), dim3(1024), shm, s, sb);
    }
    return GTN_OK;
}

// orthonormalise X (rows x p, ld p) of every block in place (CholeskyQR, `passes` times)
static int orthonormalise(std::vector<Blk *> &bs, double *(*pick)(Blk &), bool rows_m, int passes, hipStream_t s) {
    for (int pass = 0; pass < passes; ++pass) {
        std::vector<gtn_gemm_desc> g;
        for (Blk *b : bs) {
            double *X = pick(*b);
            int64_t rows = rows_m ? b->b.m : b->b.n, p = b->p;
            int S = pick_split(rows);
            int64_t ch = rows / S;
            g.push_back(gd(X, 1, p, X, p, 1, b->part, p, p, p, ch, S, ch * p, ch * p, p * p));
        }
        int rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        for (Blk *b : bs) {
            int64_t rows = rows_m ? b->b.m : b->b.n;
            R8(b->part, (int64_t)pick_split(rows) * b->p * b->p, "gram partials");
            R8(b->W, b->p * b->p, "gram");
            LAUNCH(k_sum_partials, dim3(grid_for(b->p * b->p, 256)), dim3(256), 0, s, (const double *)b->part,
                               b->p * b->p, pick_split(rows), b->W);
        }
        SmallBatch proto;
        memset(&proto, 0, sizeof(proto));
        proto.shift_rel = pass == 0 ? 1e-13 : 0.0;
        rc = small_launch(
            bs,
            [](Blk &b, SmallBatch &sb, int i) {
                sb.p[i] = (int)b.p;
                sb.in[i] = b.W;
                sb.out[i] = b.R;
            },
            k_chol_inv, sizeof(double), sizeof(double), proto, s);
        if (rc) return rc;
        g.clear();
        for (Blk *b : bs) {
            int64_t rows = rows_m ? b->b.m : b->b.n, p = b->p;
            g.push_back(gd(pick(*b), p, 1, b->R, p, 1, b->T, p, rows, p, p));
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        for (Blk *b : bs) {
            int64_t rows = rows_m ? b->b.m : b->b.n, p = b->p;
            R8(b->T, rows * p, "orth scratch");
            R8(pick(*b), rows * p, "orth basis");
            LAUNCH(k_copy2d, dim3(grid_for(rows * p, 256)), dim3(256), 0, s, (const double *)b->T, p, pick(*b), p,
                               rows, p);
        }
    }
    return GTN_OK;
}

static int jacobi(std::vector<Blk *> &bs, int by_abs, hipStream_t s) {
    SmallBatch proto;
    memset(&proto, 0, sizeof(proto));
    proto.by_abs = by_abs;
    proto.max_sweeps = 30;
    return small_launch(
        bs,
        [](Blk &b, SmallBatch &sb, int i) {
            sb.p[i] = (int)b.p;
            sb.in[i] = b.G;
            sb.out[i] = b.Vs;
            sb.aux[i] = b.V;
            sb.lam[i] = b.lam;
        },
        k_jacobi, sizeof(double), sizeof(double) + 3 * sizeof(int), proto, s);
}

static int run_blocks(std::vector<Blk> &blocks, const gtn_svd_opts &o, hipStream_t s) {
    const int eig = o.mode == 1;
    std::vector<Blk *> small, large, all;
    for (Blk &b : blocks) {
        if (b.b.k <= 0 || b.b.m <= 0 || b.b.n <= 0) continue;
        all.push_back(&b);
        (b.small ? small : large).push_back(&b);
    }
    if (all.empty()) return GTN_OK;
    int rc;
    // ---- small blocks: G = Gram of the short side (svd) or the block itself (eig)
    if (!small.empty()) {
        std::vector<gtn_gemm_desc> g;
        for (Blk *b : small) {
            const gtn_svd_block &B = b->b;
            R8(b->G, b->p * b->p, "small gram");
            LAUNCH(k_fill, dim3(grid_for(b->p * b->p, 256)), dim3(256), 0, s, b->G, b->p, b->p, b->p, 0.0);
            if (eig) {
                GTN_REQUIRE(B.m <= b->p && B.n <= b->p, "gtn_svd_blocks: eig block larger than its work size");
                R8(B.M, ext(B.m, B.n, B.lda), "small eig input");
                LAUNCH(k_copy2d, dim3(grid_for(B.m * B.n, 256)), dim3(256), 0, s, B.M, B.lda, b->G, b->p, B.m, B.n);
            } else if (b->left) {
                g.push_back(gd(B.M, B.lda, 1, B.M, 1, B.lda, b->G, b->p, B.m, B.m, B.n));  // M M^T
            } else {
                g.push_back(gd(B.M, 1, B.lda, B.M, B.lda, 1, b->G, b->p, B.n, B.n, B.m));  // M^T M
            }
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
    }
    // ---- large blocks: randomized subspace iteration
    if (!large.empty()) {
        std::vector<gtn_gemm_desc> g;
        int bi = 0;
        for (Blk *b : large) {
            R8(b->Z, b->b.n * b->p, "gaussian");
            LAUNCH(k_fill_gaussian, dim3(grid_for(b->b.n * b->p, 256)), dim3(256), 0, s, b->Z, b->b.n, b->p, b->p,
                               o.seed * 0x9E3779B97F4A7C15ull + (uint64_t)(bi++));
            g.push_back(gd(b->b.M, b->b.lda, 1, b->Z, b->p, 1, b->Y, b->p, b->b.m, b->p, b->b.n));  // Y = M Z
        }
        rc = run(g, o.precision, s);
        if (rc) return rc;
        auto pickY = [](Blk &b) { return b.Y; };
        rc = orthonormalise(large, pickY, true, o.iters > 0 ? 1 : 2, s);
        if (rc) return rc;
        for (int it = 0; it < o.iters; ++it) {
            g.clear();
            for (Blk *b : large) {
                if (eig)  // Z = M Y
                    g.push_back(gd(b->b.M, b->b.lda, 1, b->Y, b->p, 1, b->Z, b->p, b->b.m, b->p, b->b.n));
                else  // Z = M^T Y
                    g.push_back(gd(b->b.M, 1, b->b.lda, b->Y, b->p, 1, b->Z, b->p, b->b.n, b->p, b->b.m));
            }
            rc = run(g, o.precision, s);
            if (rc) return rc;
            if (!eig) {
                g.clear();
                for (Blk *b : large)  // Y = M Z
                    g.push_back(gd(b->b.M, b->b.lda, 1, b->Z, b->p, 1, b->Y, b->p, b->b.m, b->p, b->b.n));
                rc = run(g, o.precision, s);
                if (rc) return rc;
            } else {
                for (Blk *b : large) {
                    R8(b->Z, b->b.m * b->p, "eig iterate");
                    R8(b->Y, b->b.m * b->p, "eig basis");
                    LAUNCH(k_copy2d, dim3(grid_for(b->b.m * b->p, 256)), dim3(256), 0, s, (const double *)b->Z,
                                       b->p, b->Y, b->p, b->b.m, b->p);
                }
            }
            rc = orthonormalise(large, pickY, true, it + 1 == o.iters ? 2 : 1, s);
            if (rc) return rc;
        }
        // Rayleigh-Ritz in FP64: Bm = Y^T M (p x n); G = Bm Bm^T (svd) or Bm Y (eig)
        g.clear();
        for (Blk *b : large)
            g.push_back(gd(b->Y, 1, b->p, b->b.M, b->b.lda, 1, b->Bm, b->b.n, b->p, b->b.n, b->b.m));
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        g.clear();
        for (Blk *b : large) {
            if (eig) {
                g.push_back(gd(b->Bm, b->b.n, 1, b->Y, b->p, 1, b->G, b->p, b->p, b->p, b->b.n));
            } else {
                int S = pick_split(b->b.n);
                int64_t ch = b->b.n / S;
                g.push_back(gd(b->Bm, b->b.n, 1, b->Bm, 1, b->b.n, b->part, b->p, b->p, b->p, ch, S, ch, ch, b->p * b->p));
            }
        }
        rc = run(g, GTN_PREC_F64, s);
        if (rc) return rc;
        if (!eig) {
            for (Blk *b : large) {
                R8(b->part, (int64_t)pick_split(b->b.n) * b->p * b->p, "ritz partials");
                R8(b->G, b->p * b->p, "ritz gram");
                LAUNCH(k_sum_partials, dim3(grid_for(b->p * b->p, 256)), dim3(256), 0, s, (const double *)b->part,
                                   b->p * b->p, pick_split(b->b.n), b->G);
            }
        }
    }
    // ---- p x p eigenproblems of every block at once
    rc = jacobi(all, eig, s);
    if (rc) return rc;
    for (Blk *b : all) {
        GTN_REQUIRE(b->b.k <= 256 && b->b.k <= b->p, "gtn_svd_blocks: k too large");
        R8(b->lam, b->b.k, "lambda");
        R8(b->b.S, b->b.k, "S");
        R8(b->inv, b->b.k, "inv");
        R8(b->keep, b->b.k, "keep");
        LAUNCH(k_sigma, dim3(1), dim3(256), 0, s, (const double *)b->lam, b->b.k, eig, b->b.S, b->inv, b->keep);
    }
    // ---- singular vectors
    std::vector<gtn_gemm_desc> g;
    for (Blk *b : all) {
        const gtn_svd_block &B = b->b;
        const int64_t k = B.k, p = b->p;
        if (b->small) {
            if (eig || b->left) {
                GTN_REQUIRE(B.m <= p, "gtn_svd_blocks: small block rows exceed the work size");
                R8(b->Vs, ext(B.m, k, p), "eigenvectors");
                R8(B.U, ext(B.m, k, B.ldu), "U");
                LAUNCH(k_copy2d, dim3(grid_for(B.m * k, 256)), dim3(256), 0, s, (const double *)b->Vs, p, B.U, B.ldu,
                                   B.m, k);
                if (!eig) g.push_back(gd(b->Vs, 1, p, B.M, B.lda, 1, B.Vt, B.ldv, k, B.n, B.m));  // Vt = U^T M
            } else {
                GTN_REQUIRE(B.n <= p, "gtn_svd_blocks: small block columns exceed the work size");
                R8(b->Vs, ext(B.n, k, p), "eigenvectors");
                R8(B.Vt, ext(k, B.n, B.ldv), "Vt");
                LAUNCH(k_transpose2d, dim3(grid_for(B.n * k, 256)), dim3(256), 0, s, (const double *)b->Vs, p, B.Vt,
                                   B.ldv, B.n, k);  // Vt = Vs[:, :k]^T
                g.push_back(gd(B.M, B.lda, 1, b->Vs, p, 1, B.U, B.ldu, B.m, k, B.n));  // U = M V
            }
        } else {
            g.push_back(gd(b->Y, p, 1, b->Vs, p, 1, B.U, B.ldu, B.m, k, p));  // U = Y Vs
            if (!eig) g.push_back(gd(b->Vs, 1, p, b->Bm, B.n, 1, B.Vt, B.ldv, k, B.n, p));  // Vt = Vs^T Bm
        }
    }
    rc = run(g, GTN_PREC_F64, s);
    if (rc) return rc;
    for (Blk *b : all) {
        const gtn_svd_block &B = b->b;
        const int64_t k = B.k;
        if (!eig) {
            bool u_from_m = b->small && !b->left;
            R8(B.U, ext(B.m, k, B.ldu), "U");
            R8(B.Vt, ext(k, B.n, B.ldv), "Vt");
            R8(b->inv, k, "inv");
            R8(b->keep, k, "keep");
            if (u_from_m) {
                LAUNCH(k_scale_cols, dim3(grid_for(B.m * k, 256)), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu,
                                   (const double *)b->inv);
                LAUNCH(k_scale_rows, dim3(grid_for(k * B.n, 256)), dim3(256), 0, s, B.Vt, (int)k, B.n, B.ldv,
                                   (const double *)b->keep);
            } else {
                LAUNCH(k_scale_rows, dim3(grid_for(k * B.n, 256)), dim3(256), 0, s, B.Vt, (int)k, B.n, B.ldv,
                                   (const double *)b->inv);
                LAUNCH(k_scale_cols, dim3(grid_for(B.m * k, 256)), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu,
                                   (const double *)b->keep);
            }
            LAUNCH(k_gauge, dim3(k), dim3(256), 0, s, B.U, B.m, B.ldu, B.Vt, B.n, B.ldv);
        } else {
            R8(B.U, ext(B.m, k, B.ldu), "U");
            R8(b->keep, k, "keep");
            LAUNCH(k_scale_cols, dim3(grid_for(B.m * k, 256)), dim3(256), 0, s, B.U, B.m, (int)k, B.ldu,
                               (const double *)b->keep);
            LAUNCH(k_gauge, dim3(k), dim3(256), 0, s, B.U, B.m, B.ldu, (double *)nullptr, (int64_t)0, (int64_t)0);
            if (B.Vt) {
                R8(B.Vt, ext(k, B.m, B.ldv), "Vt");
                LAUNCH(k_transpose2d, dim3(grid_for(B.m * k, 256)), dim3(256), 0, s, (const double *)B.U, B.ldu,
                                   B.Vt, B.ldv, B.m, k);
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
