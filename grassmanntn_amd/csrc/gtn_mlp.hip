// Elementwise stages of an ensemble of MLPs (the learned surrogate of the coupling sweep; no
// counterpart in the reference).  The matrix products run on the grouped GEMM of gtn_gemm.hip in
// its BF16 modes; the kernels here are the rest of a training or inference step: bias + GELU with
// the pre-activation kept for the backward pass, the GELU derivative, bias gradients, a weighted
// mean-squared error, AdamW with a cosine learning-rate schedule, and the ensemble mean / spread.
//
// Layout: an ensemble of E members holds, per layer, activations as E row-major (n x c) FP32
// matrices, one after the other.  GELU is x * Phi(x) = 0.5 x (1 + erf(x / sqrt 2)).
#include "gtn_common.h"
#include <math.h>

namespace gtn {

constexpr double kInvSqrt2 = 0.70710678118654752440;
constexpr double kInvSqrt2Pi = 0.39894228040143267794;

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * (float)kInvSqrt2)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * (float)kInvSqrt2)) + x * (float)kInvSqrt2Pi * expf(-0.5f * x * x);
}

// y[e][r][j] += bias[e][j]; pre = y (optional); y = act(y).  grid (blocks, E)
__global__ void k_bias_act(float *y, int64_t n, int64_t c, const float *bias, int act, float *pre) {
    const int64_t e = blockIdx.y, total = n * c;
    float *ye = y + e * total;
    float *pe = pre ? pre + e * total : nullptr;
    const float *be = bias + e * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        float v = ye[i] + be[i % c];
        if (pe) pe[i] = v;
        ye[i] = act == 1 ? gelu(v) : v;
    }
}

__global__ void k_act_backward(float *dy, const float *pre, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dy[i] *= gelu_grad(pre[i]);
}

// out[e][j] = sum_r x[e][r][j] in two passes with a fixed summation order (FP64 accumulation).
// Pass 1, grid (ceil(c / 64), E, kColChunks): rows [q n / R, (q + 1) n / R) of 64 columns; lane ->
// column, the four wavefronts take every fourth row; part[(e R + q) c + j].
constexpr int kColChunks = 32;
__global__ void __launch_bounds__(256) k_colsum_partial(const float *x, int64_t n, int64_t c, double *part) {
    __shared__ double sh[4][64];
    const int64_t e = blockIdx.y, q = blockIdx.z, j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int w = threadIdx.x >> 6;
    const int64_t r0 = q * n / kColChunks, r1 = (q + 1) * n / kColChunks;
    double acc = 0.0;
    if (j < c) {
        const float *xe = x + e * n * c + j;
#pragma unroll 4
        for (int64_t r = r0 + w; r < r1; r += 4) acc += (double)xe[r * c];
    }
    sh[w][threadIdx.x & 63] = acc;
    __syncthreads();
    if (w == 0 && j < c)
        part[(e * kColChunks + q) * c + j] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}
__global__ void k_colsum_final(const double *part, int64_t E, int64_t c, float *out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * c) return;
    const int64_t e = t / c, j = t - e * c;
    double acc = 0.0;
    for (int q = 0; q < kColChunks; ++q) acc += part[(e * kColChunks + q) * c + j];
    out[t] = (float)acc;
}

// dout = 2 w_j (out - target) / n; loss[e] = sum_{r,j} w_j (out - target)^2 / n.  One workgroup per
// member; the target (n x c) is shared by the members.
__global__ void k_mse(const float *out, const float *target, const float *w, int64_t n, int64_t c, float *dout,
                      double *loss) {
    __shared__ double sh[16];
    const int64_t e = blockIdx.x, total = n * c;
    const float *oe = out + e * total;
    float *de = dout + e * total;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < total; i += blockDim.x) {
        const int64_t j = i % c;
        const double d = (double)oe[i] - (double)target[i];
        de[i] = (float)(2.0 * w[j] * d / (double)n);
        acc += w[j] * d * d;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
        if (loss) loss[e] = t / (double)n;
    }
}

// AdamW (decoupled weight decay) at step t = step[0] + 1 with a cosine learning rate from lr_max
// (t = 1) to lr_min (t = total)
__global__ void k_adamw(float *w, const float *g, float *m, float *v, int64_t count, gtn_adamw_opts o, const double *step) {
    const double t = step[0] + 1.0;
    const double frac = o.total_steps > 1 ? fmin((t - 1.0) / (double)(o.total_steps - 1), 1.0) : 0.0;
    const double lr = o.lr_min + 0.5 * (o.lr_max - o.lr_min) * (1.0 + cos(3.14159265358979323846 * frac));
    const double c1 = 1.0 / (1.0 - pow(o.beta1, t)), c2 = 1.0 / (1.0 - pow(o.beta2, t));
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const double gi = g[i];
        const double mi = o.beta1 * m[i] + (1.0 - o.beta1) * gi;
        const double vi = o.beta2 * v[i] + (1.0 - o.beta2) * gi * gi;
        m[i] = (float)mi;
        v[i] = (float)vi;
        const double wi = (double)w[i] * (1.0 - lr * o.weight_decay);
        w[i] = (float)(wi - lr * (mi * c1) / (sqrt(vi * c2) + o.eps));
    }
}

__global__ void k_step_inc(double *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] += 1.0;
}

// mean / standard deviation (population, divisor E) over the members of count values each
__global__ void k_ens_stats(const float *out, int64_t E, int64_t count, float *mean, float *sd) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0, q = 0.0;
        for (int64_t e = 0; e < E; ++e) s += (double)out[e * count + i];
        const double mu = s / (double)E;
        for (int64_t e = 0; e < E; ++e) {
            const double d = (double)out[e * count + i] - mu;
            q += d * d;
        }
        mean[i] = (float)mu;
        if (sd) sd[i] = (float)sqrt(q / (double)E);
    }
}

}  // namespace gtn

using namespace gtn;

extern "C" {

int gtn_mlp_bias_act(float *y, int64_t E, int64_t n, int64_t c, const float *bias, int act, float *pre, void *stream) {
    GTN_REQUIRE(E >= 1 && n >= 0 && c >= 1 && y && bias, "gtn_mlp_bias_act: bad argument");
    GTN_REQUIRE(act == GTN_ACT_NONE || act == GTN_ACT_GELU, "gtn_mlp_bias_act: unknown activation %d", act);
    GTN_REQUIRE(E <= 65535, "gtn_mlp_bias_act: too many members");
    if (n == 0) return GTN_OK;
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(n * c, 256, 4096), (unsigned)E), dim3(256), 0, as_stream(stream), y, n, c,
                       bias, act, pre);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_mlp_act_backward(float *dy, const float *pre, int64_t count, int act, void *stream) {
    GTN_REQUIRE(act == GTN_ACT_GELU, "gtn_mlp_act_backward: only GELU has a derivative pass");
    GTN_REQUIRE(count >= 0 && (count == 0 || (dy && pre)), "gtn_mlp_act_backward: bad argument");
    if (count == 0) return GTN_OK;
    hipLaunchKernelGGL(k_act_backward, dim3(grid_for(count, 256, 1 << 16)), dim3(256), 0, as_stream(stream), dy, pre, count);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int64_t gtn_mlp_colsum_workspace(int64_t E, int64_t c) { return E * c * kColChunks * (int64_t)sizeof(double); }

int gtn_mlp_colsum(const float *x, int64_t E, int64_t n, int64_t c, float *out, double *part, void *stream) {
    GTN_REQUIRE(E >= 1 && E <= 65535 && n >= 0 && c >= 1 && x && out && part, "gtn_mlp_colsum: bad argument");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_colsum_partial, dim3((unsigned)((c + 63) / 64), (unsigned)E, kColChunks), dim3(256), 0, s, x, n,
                       c, part);
    GTN_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_colsum_final, dim3(grid_for(E * c, 256)), dim3(256), 0, s, (const double *)part, E, c, out);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_mlp_mse(const float *out, const float *target, const float *weight, int64_t E, int64_t n, int64_t c, float *dout,
                double *loss, void *stream) {
    GTN_REQUIRE(E >= 1 && n >= 1 && c >= 1 && out && target && weight && dout, "gtn_mlp_mse: bad argument");
    hipLaunchKernelGGL(k_mse, dim3((unsigned)E), dim3(256), 0, as_stream(stream), out, target, weight, n, c, dout, loss);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_adamw(float *w, const float *g, float *m, float *v, int64_t count, const gtn_adamw_opts *opts, double *step,
              void *stream) {
    GTN_REQUIRE(count >= 0 && opts && step && (count == 0 || (w && g && m && v)), "gtn_adamw: bad argument");
    GTN_REQUIRE(opts->beta1 >= 0 && opts->beta1 < 1 && opts->beta2 >= 0 && opts->beta2 < 1 && opts->eps > 0,
                "gtn_adamw: bad moments");
    hipStream_t s = as_stream(stream);
    if (count) {
        hipLaunchKernelGGL(k_adamw, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s, w, g, m, v, count, *opts,
                           (const double *)step);
        GTN_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

int gtn_ensemble_stats(const float *out, int64_t E, int64_t count, float *mean, float *sd, void *stream) {
    GTN_REQUIRE(E >= 1 && count >= 0 && (count == 0 || (out && mean)), "gtn_ensemble_stats: bad argument");
    if (count == 0) return GTN_OK;
    hipLaunchKernelGGL(k_ens_stats, dim3(grid_for(count, 256)), dim3(256), 0, as_stream(stream), out, E, count, mean, sd);
    GTN_CHECK_LAUNCH();
    return GTN_OK;
}

}  // extern "C"
