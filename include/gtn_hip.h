/*
 * gtn_hip.h -- C ABI of the MI355X (gfx950) Grassmann-tensor kernels.
 *
 * The reference (grassmanntn) is pure Python/numpy: it has no FFI of its own.  Each entry point here
 * replaces one numeric seam of that Python code; the comment on each names the reference function
 * (file:line, paths relative to the grassmanntn tree) whose arithmetic it performs.  The Python
 * package grassmanntn_amd binds these with ctypes (see grassmanntn_amd/_lib.py).
 *
 * Conventions
 *   - every pointer argument named *_dev is device memory (HBM); descriptors are host memory;
 *   - `stream` is a hipStream_t passed as void*; NULL means the default stream;
 *   - every entry point returns 0 on success and a nonzero GTN_E* code on failure, after recording
 *     a message that gtn_last_error() returns (thread-local);
 *   - all index arithmetic is exact integer arithmetic; sign factors are +-1 computed on the device
 *     from popcounts (parity p(i) = popcount(i) & 1, sigma(i) = (-1)^(p(p-1)/2) with p = popcount(i),
 *     encoder eps(i) = i ^ (popcount(i >> 1) & 1)).
 */
#ifndef GTN_HIP_H
#define GTN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTN_OK 0
#define GTN_EINVAL 1   /* bad argument or descriptor */
#define GTN_EDEVICE 2  /* HIP runtime error */
#define GTN_ENOMEM 3   /* workspace too small */
#define GTN_ENOTSUP 4  /* unsupported combination */

/* element types */
#define GTN_F64 0
#define GTN_F32 1
#define GTN_C128 2
#define GTN_I8 3

/* arithmetic precisions of the contraction GEMMs */
#define GTN_PREC_F64 0    /* FP64 MFMA, FP64 accumulation */
#define GTN_PREC_F32 1    /* operands rounded to FP32, FP32 MFMA, FP32 accumulation */
#define GTN_PREC_BF16X3 2 /* operands split hi+lo in BF16, 3 BF16 MFMA products, FP32 accumulation */
#define GTN_PREC_BF16X6 3 /* operands split in 3 BF16 parts, 6 products, FP32 accumulation */
#define GTN_PREC_BF16 4   /* operands rounded to BF16, one product, FP32 accumulation */
#define GTN_PREC_BF16_SPLIT_A 5 /* A split hi+lo in BF16, B rounded to BF16, 2 products, FP32 accumulation
                                 * (BF16 weights times activations kept to ~16 bits) */
#define GTN_PREC_F64_INT8 6 /* range finder of gtn_svd_blocks only: FP64 products on the int8 matrix cores
                             * (each operand row scaled by a power of two and cut into eight 7-bit integer
                             * slices, exact int32 slice products, FP64 recombination: the error of a
                             * native FP64 product) */
#define GTN_PREC_F64_GRAM 7 /* range finder of gtn_svd_blocks only: FP64 subspace iteration on the Gram
                             * matrix G = M M^T (formed once per call from its upper block triangle), one
                             * product G Y per iteration instead of M^T Y and M Z; blocks with m > n
                             * use the alternating FP64 products */

int gtn_version(void);

/* Host-only validation ("dry run"): while enabled, no kernel is launched and no device memory is
 * touched; every GEMM operand range and every buffer range a kernel would access is checked
 * against the registered regions [base, base + bytes), and the first violation makes the entry
 * point fail with GTN_EINVAL.  Used by CPU tests to check launch geometry with placeholder
 * addresses.  n_regions = 0 disables the mode. */
int gtn_set_dry_run(int n_regions, const uint64_t *bases, const int64_t *bytes);
const char *gtn_last_error(void);
/* Checks that the current device is a gfx950; fills its architecture name and compute-unit count. */
int gtn_device_info(char *name, int name_len, int *n_cu);

/* ------------------------------------------------------------------------------------------------
 * Device memory and ordering.  The host layer keeps every tensor in HBM allocated here (a caching
 * allocator: blocks freed with gtn_free are reused by later gtn_malloc calls of the same size class,
 * in stream order on the one stream the host layer uses).
 * gtn_memcpy kind: 1 host->device, 2 device->host (both synchronous), 3 device->device (ordered on
 * `stream`).  gtn_fill writes `value` to n elements (complex: only 0 is accepted); gtn_copy2d copies
 * a rows x cols block between leading dimensions lds and ldd.
 * ---------------------------------------------------------------------------------------------- */
int gtn_malloc(int64_t bytes, void **out_dev);
int gtn_free(void *dev);
/* Bytes the allocator has handed out and not taken back (live), bytes it holds for reuse (cached:
 * the general cache and the free blocks of graph pools), and the device's free / total memory
 * (hipMemGetInfo).  No counterpart in the reference, which keeps its arrays on the host; callers use
 * it to size batches. */
int gtn_mem_info(int64_t *live_bytes, int64_t *cached_bytes, int64_t *device_free, int64_t *device_total);
int gtn_memcpy(void *dst, const void *src, int64_t bytes, int kind, void *stream);
int gtn_fill(void *dev, int64_t n, int dtype, double value, void *stream);
int gtn_copy2d(void *dst_dev, int64_t ldd, const void *src_dev, int64_t lds, int64_t rows, int64_t cols, int dtype,
               void *stream);
int gtn_synchronize(void);
int gtn_event_create(void **ev);
int gtn_event_record(void *ev, void *stream);
int gtn_event_elapsed_ms(void *start, void *stop, float *ms);
int gtn_event_destroy(void *ev);

/* ------------------------------------------------------------------------------------------------
 * Streams and HIP graphs (no counterpart in the reference, whose coarse-graining loop,
 * example.py:162-196, is plain Python): the host layer orders its work on a stream created here, and
 * a coarse-graining step issued between gtn_graph_begin and gtn_graph_end is recorded once and
 * replayed with gtn_graph_launch.
 * gtn_stream_create: a non-blocking stream at the default priority.
 * gtn_graph_begin(stream, pool): work issued on `stream` is recorded (relaxed capture mode), not run,
 *   until gtn_graph_end; meanwhile gtn_malloc serves allocations from memory pool `pool` (> 0).  A
 *   pool's blocks, freed during or after the capture, are reused only by allocations from the same
 *   pool, so every address a graph holds stays its own; inside one capture a freed block may be
 *   reused in stream order, which every replay keeps.  One capture at a time per process.
 * gtn_graph_end(stream, &exec): closes the capture and instantiates it (exec NULL: abandon it).
 * gtn_graph_launch(exec, stream): one replay, ordered on `stream`.
 * gtn_pool_release(pool): after the pool's graphs are destroyed, its blocks go back to the general
 *   cache (blocks still referenced return there when freed).
 * While a graph exists or a capture is open the library frees none of its own buffers (gather
 * tables, split-K workspaces).
 * ---------------------------------------------------------------------------------------------- */
int gtn_stream_create(void **stream);
int gtn_stream_destroy(void *stream);
int gtn_stream_synchronize(void *stream);
int gtn_graph_begin(void *stream, int pool);
int gtn_graph_end(void *stream, void **graph_exec);
int gtn_graph_launch(void *graph_exec, void *stream);
int gtn_graph_destroy(void *graph_exec);
int gtn_pool_release(int pool);

/* ------------------------------------------------------------------------------------------------
 * (A1) index tables.  Replaces param.gparity / param.sgn / param.encoder (param.py:61-83; the
 * literal tables of param.py:4-6 are their values on [0, 65536)).
 * popc[i] = popcount(i), sigma[i] = +-1, enc[i] = eps(i), for 0 <= i < n.  Any pointer may be NULL.
 * ---------------------------------------------------------------------------------------------- */
int gtn_index_tables(int64_t n, int32_t *popc_dev, int8_t *sigma_dev, int32_t *enc_dev, void *stream);

/* ------------------------------------------------------------------------------------------------
 * Signed gather ("permute").  One descriptor covers the reference's dense.switch_encoder
 * (__init__.py:1070-1102), switch_format (:1011-1068), switch_parity (:1104-1154), the
 * single-operand einsum with its S3 sign tensor (:2045-2126, :2295), the grouping signs and
 * reshapes of join_legs / split_legs (:2947-3170), the parity-block extraction of BlockSVD
 * (:3990-3996) and the operand preparation of every contraction (S1 / S2 folded in).
 *
 * Output element with raw multi-index (o_0..o_{n_out-1}) (C order, last axis fastest):
 *   virtual axes:   V_v = sum_a o_a * out_vmul[a][t] over the (a, t) with out_virt[a][t] == v
 *                   c_v = virt_enc[v] ? eps(V_v) : V_v                   (canonical fused index)
 *                   sign *= virt_sigma[v] ? sigma(c_v) : 1
 *                   c_v is split, C order, into virt_ncomp[v] consecutive components k_j
 *   components:     bits ^= (popcount(k_j) & 1) << comp_bit[j]          (if comp_bit[j] >= 0)
 *                   sign *= comp_sigma[j] ? sigma(k_j) : 1
 *                   F[comp_in[j][t]] += k_j * comp_mult[j][t]           (t = 0, 1; -1 = unused)
 *   input axes:     e_b = in_enc[b] ? eps(F_b) : F_b ;  sign *= in_sigma[b] ? sigma(F_b) : 1
 *                   offset += in_split[b] ? (e_b & 1) * in_stride_lo[b] + (e_b >> 1) * in_stride[b]
 *                                         : e_b * in_stride[b]
 *   table:          2 bits per entry, entry `bits`: 0 -> +1, 1 -> -1, 2 -> the element is 0
 *   dst[o] = alpha * sign * src[offset]   (src is not read for a zero entry)
 * ---------------------------------------------------------------------------------------------- */
#define GTN_PERM_MAX_AX 10
#define GTN_PERM_MAX_COMP 16
#define GTN_PERM_MAX_BITS 10
typedef struct gtn_perm_desc {
    int32_t dtype; /* GTN_F64 / GTN_F32 / GTN_C128 (src and dst share it) */
    int32_t n_out, n_virt, n_comp, n_in, n_bits;
    int64_t out_dim[GTN_PERM_MAX_AX];
    int32_t out_virt[GTN_PERM_MAX_AX][2];
    int64_t out_vmul[GTN_PERM_MAX_AX][2];
    int64_t virt_dim[GTN_PERM_MAX_AX];
    int32_t virt_enc[GTN_PERM_MAX_AX];
    int32_t virt_sigma[GTN_PERM_MAX_AX];
    int32_t virt_ncomp[GTN_PERM_MAX_AX];
    int64_t comp_dim[GTN_PERM_MAX_COMP];
    int32_t comp_bit[GTN_PERM_MAX_COMP];
    int32_t comp_sigma[GTN_PERM_MAX_COMP];
    int32_t comp_in[GTN_PERM_MAX_COMP][2];
    int64_t comp_mult[GTN_PERM_MAX_COMP][2];
    int64_t in_dim[GTN_PERM_MAX_AX];
    int32_t in_enc[GTN_PERM_MAX_AX];
    int32_t in_sigma[GTN_PERM_MAX_AX];
    int32_t in_split[GTN_PERM_MAX_AX];
    int64_t in_stride[GTN_PERM_MAX_AX];
    int64_t in_stride_lo[GTN_PERM_MAX_AX];
    uint32_t table[2 * (1 << GTN_PERM_MAX_BITS) / 32];
    double alpha_re, alpha_im;
    int64_t src_elems; /* number of elements addressable from src (checked on the host) */
} gtn_perm_desc;
int gtn_permute(const gtn_perm_desc *desc, const void *src_dev, void *dst_dev, void *stream);
/* Two gathers composed in one pass (no counterpart in the reference, whose contractions materialise
 * every intermediate): dst = outer(inner(src)), i.e. dst[i] = sign_outer(i) * sign_inner(j) *
 * src[off_inner(j)] with j = off_outer(i), the outer gather reading the contiguous tensor the inner
 * one would write (outer->src_elems must equal the inner gather's output size); both F64 or both
 * F32; alpha multiplies. */
int gtn_permute2(const gtn_perm_desc *outer, const gtn_perm_desc *inner, const void *src_dev, void *dst_dev,
                 void *stream);

/* ------------------------------------------------------------------------------------------------
 * (A3) sign tensor, as the reference materialises S1 / S3 (__init__.py:1944-1999, :2088-2126):
 * out[idx] = table[bits(idx)] where bit j of bits(idx) is popcount(idx_j) & 1 (canonical indices);
 * table: 1 bit per entry, set => -1.  Output int8 (+1 / -1), C order over dims[0..ndim).
 * ---------------------------------------------------------------------------------------------- */
int gtn_sign_tensor(int ndim, const int64_t *dims, const uint32_t *table, int8_t *out_dev, void *stream);

/* ------------------------------------------------------------------------------------------------
 * (A4/A5) grouped, batched GEMM:  C_g[b] = alpha_g * A_g[b] * B_g[b] + beta_g * C_g[b]
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C(m,n) = C[m*scm + n*scn]; batch strides
 * sab/sbb/scb.  This is the parity-block-grouped contraction that replaces oe.contract
 * (__init__.py:2295) and einsum_block's per-sector np.einsum (:2781): every nonzero parity sector of
 * a contraction is one group.  Element type GTN_F64 or GTN_F32; `precision` is a GTN_PREC_* value.
 * For each operand exactly one of its two strides must be 1.
 * ---------------------------------------------------------------------------------------------- */
typedef struct gtn_gemm_desc {
    const void *A;
    const void *B;
    void *C;
    int64_t M, N, K, batch;
    int64_t sam, sak, sbk, sbn, scm, scn;
    int64_t sab, sbb, scb;
    double alpha, beta;
} gtn_gemm_desc;
int gtn_gemm_grouped(const gtn_gemm_desc *groups, int n_groups, int dtype, int precision, void *stream);

/* Autotuner of gtn_gemm_grouped's launch configuration (no counterpart in the reference).  With it on
 * (the default; GTN_AUTOTUNE=0 in the environment or gtn_autotune_set(0) turns it off) the first
 * eager call of each signature -- element type, precision, and per group M, N, K, batch and operand
 * layouts -- runs every candidate configuration (tile 64 or 128, the latter only where every group
 * is at least 128 x 128; split-K target off, 512, 1024 or 2048 workgroups) on its own operands, times
 * each with HIP events (best of 3), and keeps the fastest for every later call of that signature in
 * the process.  Calls issued while their stream is capturing a graph, and calls with beta != 0, are
 * never timed: they use the tuned configuration if there is one, else the heuristics (128 x 128 tiles
 * when every group is at least 128 x 128 and the launch has >= 256 of them; split-K towards 1024
 * workgroups when it has < 512 tiles).  Results are identical up to the summation order of split-K
 * partials.  gtn_autotune_report copies up to max_entries tuned signatures (first group's sizes) and
 * returns how many there are. */
typedef struct gtn_tune_entry {
    int32_t dtype, precision, n_groups, tile, split_target, pad_;
    int64_t M, N, K, batch;
    float ms_best, ms_default;
} gtn_tune_entry;
int gtn_autotune_set(int on);
int gtn_autotune_clear(void);
int gtn_autotune_report(gtn_tune_entry *out, int max_entries);

/* Allocation-free variant of gtn_gemm_grouped.  It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels, split plan and results (bit for bit)
 * as gtn_gemm_grouped (it follows the autotuner's configuration of the signature, below, when that
 * one's split plan fits the scratch, else the default plan); the split-reduction partials that
 * gtn_gemm_grouped keeps in a workspace it
 * allocates go instead to a fixed per-device scratch array of gtn_gemm_noalloc_max_scratch() bytes
 * (64 MiB) that the library holds from its first use on a device until the process ends.  Nothing
 * is allocated or freed here.  A call whose split plan needs more scratch than that fails with
 * GTN_ENOMEM and a message giving both sizes.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap in time on one GPU (different streams, different host threads) race; calls
 * ordered on one stream do not.  Nothing in the scratch needs to start at zero (every partial is
 * written before it is read). */
int gtn_gemm_grouped_noalloc(const gtn_gemm_desc *groups, int n_groups, int dtype, int precision, void *stream);
int64_t gtn_gemm_noalloc_max_scratch(void);

/* Device self-test of the matrix-core operand/result lane layouts the GEMM kernels assume (no
 * counterpart in the reference: a diagnostic for users of this library).  One wavefront runs each
 * MFMA form (FP64 16x16x4, FP32 16x16x4, BF16 16x16x32) on inputs whose product is
 * D[i][j] = 16 i + j + 1 and stores every lane's result registers: out_dev[256 f + 4 lane + v]
 * (768 doubles).  The layout the kernels use: out_dev[256 f + 4 lane + v] = D[row][lane % 16] with
 * row = lane / 16 + 4 v for FP64 and row = 4 (lane / 16) + v for FP32 and BF16. */
int gtn_mfma_selftest(double *out_dev, void *stream);

/* Device self-test of the int8 matrix-core form the FP64-by-int8 products use (no counterpart in the
 * reference): one wavefront runs v_mfma_i32_16x16x64_i8 on A[i][k] = (i + 3k) % 7 - 3 and
 * B[k][j] = (2j + k) % 5 - 2 (i, j < 16, k < 64), lane l feeding A[l % 16][16 (l / 16) + t] and
 * B[16 (l / 16) + t][l % 16] as bytes t = 0..15, and stores its four result registers:
 * out_dev[4 lane + v] (256 int32), expected to hold D[4 (lane / 16) + v][lane % 16]. */
int gtn_mfma_selftest_i8(int32_t *out_dev, void *stream);

/* ------------------------------------------------------------------------------------------------
 * Reductions and elementwise helpers (FP64 accumulation).
 * gtn_sumsq: out_dev[0] = sum |x|^2 over n elements (Frobenius norm^2; dense.norm __init__.py:888-891,
 *   gauge2d.py:1746).
 * gtn_scale_by_norm: x /= sqrt(nrm2_dev[0]) (gauge2d.py:1747-1748).
 * gtn_reduce_last: dst[r] = sum_{t<inner} src[r*inner + t] (traced axes of a single-operand einsum).
 * gtn_lognorm_update: (log-norm accumulation of example.py:192) lognorm[0] = 2*lognorm[0] + 0.5*log(nrm2[0]).
 * gtn_power_diag: matrix-format power of a diagonal with the reference's |x| > rcond rule
 *   (power_ds, __init__.py:6059-6069): x = (|x| > rcond) ? x^p : 0, elementwise on n elements.
 * gtn_elementwise: x = (|x| > rcond) ? f(x) : 0 elementwise on n elements, f by op: GTN_EW_POWER
 *   x^p (power_ds), GTN_EW_INV 1/x (inv_nz_ds, __init__.py:5992-6004), GTN_EW_EXP exp(x) (exp_ds,
 *   __init__.py:6093-6106); the matrix-format canonical coefficients are what the caller passes.
 * gtn_axpby: out = alpha * x + beta * y elementwise (dense arithmetic, __init__.py:933-987); y may be
 *   NULL (then out = alpha * x); complex tensors take complex alpha/beta.
 * gtn_odd_norm: out_dev[0] = sum |x|^2 over entries whose fermionic parity is odd (is_grassmann_even,
 *   __init__.py:3900-3925); parity of each axis from its index (PP: LSB, canonical: popcount).
 * ---------------------------------------------------------------------------------------------- */
int gtn_sumsq(const void *x_dev, int64_t n, int dtype, double *out_dev, void *stream);
int gtn_scale_by_norm(void *x_dev, int64_t n, int dtype, const double *nrm2_dev, void *stream);
int gtn_reduce_last(const void *src_dev, void *dst_dev, int64_t rows, int64_t inner, int dtype, void *stream);
int gtn_lognorm_update(double *lognorm_dev, const double *nrm2_dev, void *stream);
int gtn_power_diag(void *x_dev, int64_t n, int64_t stride, int dtype, double p, double rcond, void *stream);
#define GTN_EW_POWER 0
#define GTN_EW_INV 1
#define GTN_EW_EXP 2
int gtn_elementwise(void *x_dev, int64_t n, int dtype, int op, double p, double rcond, void *stream);
int gtn_axpby(const void *x_dev, const void *y_dev, void *out_dev, int64_t n, int dtype, double alpha_re,
              double alpha_im, double beta_re, double beta_im, void *stream);
int gtn_odd_norm(const void *x_dev, int ndim, const int64_t *dims, const int32_t *fermionic,
                 const int32_t *pp, int dtype, double *out_dev, void *stream);
/* Per-point normalisation of a batch of coarse-grained tensors with the driver's log-norm rule
 * (gauge2d.py:1746-1748 and example.py:191-193 for every point at once): for b < nbatch (<= 1024),
 * x_b = x[b n .. b n + n): nrm2[b] = ||x_b||^2 (FP64 accumulation), x_b / ||x_b|| stored to dst_b
 * (dst_dev NULL: in place into x_b), lognorm[b] = 2 lognorm[b] + log ||x_b|| (when lognorm is not
 * NULL), hist[b hist_len + k] = log ||x_b|| with k = step[0] (when hist and step are given and
 * k < hist_len), then step[0] += 1. */
int gtn_normalize_batch(void *x_dev, void *dst_dev, int64_t nbatch, int64_t n, int dtype, double *nrm2_dev,
                        double *lognorm_dev, double *hist_dev, int hist_len, double *step_dev, void *stream);
/* The diagonal singular-value matrices of a batch of parity-blocked decompositions
 * (get_full_matrix of BlockSVD, __init__.py:4017-4028): dst (nbatch, 2d, 2d) with
 * dst[b][r][r] = src[b][r % 2][r / 2] (src (nbatch, 2, d): the even then the odd block's values) and
 * zero off the diagonal. */
int gtn_diag_interleave(void *dst_dev, int64_t nbatch, int64_t d, const void *src_dev, int dtype, void *stream);

/* ------------------------------------------------------------------------------------------------
 * (A10/A11) truncated SVD / Hermitian eig of a batch of parity blocks (BlockSVD / SortedSVD,
 * __init__.py:3931-4029; BlockEig / SortedEig :4314-4422).  For each block b (an m_b x n_b FP64
 * matrix, row-major with leading dimension lda_b) the top k_b singular triplets are returned:
 *   U_b (m_b x k_b, ld ldu_b), S_b (k_b), Vt_b (k_b x n_b, ld ldv_b),
 * ordered by decreasing singular value, with singular values at or below 1e-14 relative to the
 * largest set to zero together with their vectors (the rank rule of __init__.py:3939-3941).
 * Algorithm: randomized subspace iteration with block width p_b = k_b + oversample, `iters`
 * iterations, CholeskyQR2 re-orthonormalisation, then an FP64 Rayleigh-Ritz step (B = Q^T M,
 * eigen-decomposition of B B^T by cyclic Jacobi) -- the range finder may run in a reduced
 * `precision`, optionally with its last f64_tail iterations in FP64 (the iteration is a fixed-point
 * map: the FP64 iterations damp the rounding error of the reduced ones outside the top-p
 * subspace); the Rayleigh-Ritz step is always FP64.
 * Gauge: each column of U is signed so that its entry of largest magnitude is positive.
 * eig mode (mode=1): the block must be Hermitian (checked on the host path); eigenvalues are the
 * signed Rayleigh quotients lambda_k = u_k^T M u_k, ordered by |lambda|, and Vt = U^T.
 * ---------------------------------------------------------------------------------------------- */
typedef struct gtn_svd_block {
    const double *M;
    double *U, *S, *Vt;
    int64_t m, n, lda, ldu, ldv;
    int32_t k;
} gtn_svd_block;
typedef struct gtn_svd_opts {
    int32_t oversample; /* extra columns in the range finder */
    int32_t iters;      /* subspace iterations */
    int32_t precision;  /* GTN_PREC_* for the range finder GEMMs */
    int32_t mode;       /* 0: SVD, 1: Hermitian eig */
    uint64_t seed;
    int32_t orth_every; /* re-orthonormalise the basis every this many iterations (<= 1: every one) */
    int32_t f64_tail;   /* the last f64_tail iterations run in FP64 whatever `precision` is (0: none) */
} gtn_svd_opts;
int64_t gtn_svd_workspace_bytes(const gtn_svd_block *blocks, int n_blocks, const gtn_svd_opts *opts);
int gtn_svd_blocks(const gtn_svd_block *blocks, int n_blocks, const gtn_svd_opts *opts, void *workspace_dev,
                   int64_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------------------------------
 * Learned surrogate of the coupling sweep (no counterpart in the reference): an ensemble of E MLPs
 * whose matrix products run on gtn_gemm_grouped (GTN_PREC_BF16_SPLIT_A: BF16 weights, FP32
 * accumulation) and whose other stages are these kernels.  Activations of a layer are E row-major
 * (n x c) FP32 matrices one after the other; bias is (E x c).
 * gtn_mlp_bias_act: y += bias (per column); pre (optional) = y; y = act(y) (GTN_ACT_GELU:
 *   x Phi(x) = 0.5 x (1 + erf(x / sqrt 2)); GTN_ACT_NONE: identity).
 * gtn_mlp_act_backward: dy *= GELU'(pre) over count values.
 * gtn_mlp_colsum: out[e][j] = sum_r x[e][r][j] (bias gradients; FP64 accumulation in a fixed order:
 *   partial sums of row chunks in part_dev, gtn_mlp_colsum_workspace(E, c) bytes, then their sum).
 * gtn_mlp_mse: with target (n x c) shared by the members and weight (c): dout = 2 w_j (out - t) / n,
 *   loss[e] = sum w_j (out - t)^2 / n (loss may be NULL).
 * gtn_adamw: AdamW (decoupled weight decay) on count parameters at step t = step[0] + 1, learning
 *   rate lr_min + (lr_max - lr_min) (1 + cos(pi (t - 1) / (total_steps - 1))) / 2; then step[0] += 1.
 * gtn_ensemble_stats: mean and population standard deviation over the E members (sd may be NULL).
 * ---------------------------------------------------------------------------------------------- */
#define GTN_ACT_NONE 0
#define GTN_ACT_GELU 1
typedef struct gtn_adamw_opts {
    double lr_max, lr_min, beta1, beta2, eps, weight_decay;
    int64_t total_steps;
} gtn_adamw_opts;
int gtn_mlp_bias_act(float *y_dev, int64_t E, int64_t n, int64_t c, const float *bias_dev, int act, float *pre_dev,
                     void *stream);
int gtn_mlp_act_backward(float *dy_dev, const float *pre_dev, int64_t count, int act, void *stream);
int gtn_mlp_colsum(const float *x_dev, int64_t E, int64_t n, int64_t c, float *out_dev, double *part_dev, void *stream);
int64_t gtn_mlp_colsum_workspace(int64_t E, int64_t c); /* bytes of part_dev */
int gtn_mlp_mse(const float *out_dev, const float *target_dev, const float *weight_dev, int64_t E, int64_t n, int64_t c,
                float *dout_dev, double *loss_dev, void *stream);
int gtn_adamw(float *w_dev, const float *g_dev, float *m_dev, float *v_dev, int64_t count, const gtn_adamw_opts *opts,
              double *step_dev, void *stream);
int gtn_ensemble_stats(const float *out_dev, int64_t E, int64_t count, float *mean_dev, float *sd_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
