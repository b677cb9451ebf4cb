// sbx_mlp.h -- the skin model: a per-cell MLP 12 -> 64 -> 64 -> 2 with GELU, FP32.
//
// It has no counterpart in the reference (SpatialGrid rebuilds its grid every step).  From 12
// features of a coarse cell's recent motion it predicts (out0) the log of the Verlet skin width the
// cell's bodies need to last the target rebuild horizon and (out1) the log of the number of steps
// that skin is predicted to last.  The prediction only sizes the neighbour lists: the exact
// displacement guard decides every rebuild, so pair sets never depend on it.
//
// Parameter layout (one float array of MLP_NPARAM):
//   W1[64][12] | b1[64] | W2[64][64] | b2[64] | W3[2][64] | b3[2] | xMean[12] | xStd[12]
// xMean/xStd standardise the inputs (x - mean) / std before the first layer, clamped to
// [-MLP_XCLAMP, MLP_XCLAMP] so an input far outside the training data cannot dominate; they are
// data statistics, not trained.
#pragma once
#include <hip/hip_runtime.h>

namespace sbx {

constexpr int MLP_IN = 12, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NTRAIN = MLP_B3 + MLP_OUT;       // trained parameters
constexpr int MLP_XMEAN = MLP_NTRAIN;
constexpr int MLP_XSTD = MLP_XMEAN + MLP_IN;
constexpr int MLP_NPARAM = MLP_XSTD + MLP_IN;
constexpr float MLP_XCLAMP = 6.0f;

__device__ __forceinline__ float mlp_standardise(float x, float mean, float sd) {
    return fminf(MLP_XCLAMP, fmaxf(-MLP_XCLAMP, (x - mean) / sd));
}

// exact GELU: x * Phi(x) = 0.5 x (1 + erf(x / sqrt 2)), and its derivative Phi(x) + x phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// Forward pass of one sample; w may point to LDS or global memory.  Layer 2 is accumulated
// column by column (h1[k] is consumed as soon as it is produced), so the only per-thread array is
// the 64 layer-2 accumulators, indexed by unrolled constants: it stays in registers.
__device__ __forceinline__ void mlp_forward(const float *__restrict__ w, const float x_raw[MLP_IN],
                                            float y[MLP_OUT]) {
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) x[k] = mlp_standardise(x_raw[k], w[MLP_XMEAN + k], w[MLP_XSTD + k]);
    // layer 1 in registers, then one layer-2 unit at a time (each unit sums over k in ascending order,
    // so the result does not depend on this loop order; 64 live accumulators would cap occupancy at 1)
    float h1[MLP_H];
#pragma unroll
    for (int k = 0; k < MLP_H; k++) {
        float a = w[MLP_B1 + k];
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) a += w[MLP_W1 + k * MLP_IN + i] * x[i];
        h1[k] = gelu(a);
    }
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) y[o] = w[MLP_B3 + o];
#pragma unroll 1
    for (int j = 0; j < MLP_H; j++) {
        float a = w[MLP_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; k++) a += w[MLP_W2 + j * MLP_H + k] * h1[k];
        float h = gelu(a);
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) y[o] += w[MLP_W3 + o * MLP_H + j] * h;
    }
}

}  // namespace sbx
