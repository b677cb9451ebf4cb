// sbx_mlp.hip -- training and inference kernels of the skin model (sbx_mlp.h).
//
// Training is mean-squared error on the two outputs over a batch of B samples:
//   L = 1/B * sum_b sum_o 0.5 * m_o * (y_bo - t_bo)^2        (m_o: per-output weight)
//   k_mlp_fwd_bwd: one thread per sample -- forward, then back-propagation down to the
//                  pre-activation gradients dA1[b][64], dA2[b][64] and dY[b][2]; it also stores
//                  the layer inputs Xs[b][12] (standardised), H1[b][64], H2[b][64] and the GELU
//                  slopes G1, G2;
//   k_mlp_wgrad:   one block per (weight row, sample partition), one thread per column: partial
//                  dW[j][k] = sum_b dA[b][j] * In[b][k] in ascending b, plus the bias gradient;
//                  k_mlp_grad_reduce sums the partitions in order (deterministic, no atomics);
//   k_mlp_adam:    Adam update of the trained parameters.
// All FP32.  The forward pass used at run time is mlp_forward in sbx_mlp.h.
#include "sbx_internal.h"
#include "sbx_mlp.h"

using namespace sbx;

namespace {

constexpr int T = 256;

__global__ void __launch_bounds__(T) k_mlp_forward(const float *__restrict__ w, const float *__restrict__ x, int B,
                                                   float *__restrict__ y) {
    __shared__ float ws[MLP_NPARAM];
    for (int k = threadIdx.x; k < MLP_NPARAM; k += T) ws[k] = w[k];
    __syncthreads();
    int b = blockIdx.x * T + threadIdx.x;
    if (b >= B) return;
    float xr[MLP_IN], out[MLP_OUT];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) xr[k] = x[(size_t)b * MLP_IN + k];
    mlp_forward(ws, xr, out);
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) y[(size_t)b * MLP_OUT + o] = out[o];
}

struct MlpScratch {
    float *xs, *h1, *h2, *g1, *g2, *da1, *da2, *dy, *loss;
};

// Per-sample forward + back-propagation.  The sums that would need a per-thread array indexed by
// a loop variable are accumulated column by column into 64 register accumulators instead; the
// per-unit values the weight gradients need go to the [B][64] scratch arrays.
__global__ void __launch_bounds__(T) k_mlp_fwd_bwd(const float *__restrict__ w, const float *__restrict__ x,
                                                   const float *__restrict__ target, int B, float m0, float m1,
                                                   MlpScratch s) {
    __shared__ float ws[MLP_NPARAM];
    for (int k = threadIdx.x; k < MLP_NPARAM; k += T) ws[k] = w[k];
    __syncthreads();
    int b = blockIdx.x * T + threadIdx.x;
    if (b >= B) return;
    const size_t rowH = (size_t)b * MLP_H;
    float xs[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) {
        xs[k] = mlp_standardise(x[(size_t)b * MLP_IN + k], ws[MLP_XMEAN + k], ws[MLP_XSTD + k]);
        s.xs[(size_t)b * MLP_IN + k] = xs[k];
    }
    float acc[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) acc[j] = ws[MLP_B2 + j];
    for (int k = 0; k < MLP_H; k++) {
        float a = ws[MLP_B1 + k];
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) a += ws[MLP_W1 + k * MLP_IN + i] * xs[i];
        float h = gelu(a);
        s.h1[rowH + k] = h;
        s.g1[rowH + k] = gelu_grad(a);
#pragma unroll
        for (int j = 0; j < MLP_H; j++) acc[j] += ws[MLP_W2 + j * MLP_H + k] * h;
    }
    float y[MLP_OUT];
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) y[o] = ws[MLP_B3 + o];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        float h = gelu(acc[j]);
        s.h2[rowH + j] = h;
        s.g2[rowH + j] = gelu_grad(acc[j]);
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) y[o] += ws[MLP_W3 + o * MLP_H + j] * h;
    }
    float dy[MLP_OUT];
    float loss = 0.0f;
    const float invB = 1.0f / (float)B;
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
        float m = o == 0 ? m0 : m1;
        float e = y[o] - target[(size_t)b * MLP_OUT + o];
        loss += 0.5f * m * e * e;
        dy[o] = m * e * invB;
        s.dy[(size_t)b * MLP_OUT + o] = dy[o];
    }
    s.loss[b] = loss;
    // dA2[k] = (W3^T dy)[k] * gelu'(a2[k]);  dH1 = W2^T dA2 accumulated column by column
#pragma unroll
    for (int i = 0; i < MLP_H; i++) acc[i] = 0.0f;
    for (int k = 0; k < MLP_H; k++) {
        float d = 0.0f;
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) d += ws[MLP_W3 + o * MLP_H + k] * dy[o];
        float da2 = d * s.g2[rowH + k];
        s.da2[rowH + k] = da2;
#pragma unroll
        for (int i = 0; i < MLP_H; i++) acc[i] += ws[MLP_W2 + k * MLP_H + i] * da2;
    }
#pragma unroll
    for (int i = 0; i < MLP_H; i++) s.da1[rowH + i] = acc[i] * s.g1[rowH + i];
}

// Block (r, p) handles weight row r over sample partition p: rows [0,64) W1/b1, [64,128) W2/b2,
// [128,130) W3/b3.  Thread k < nIn: partial dW[r][k] = sum_{b in p} dA[b][r] * In[b][k]; thread
// nIn: partial db[r].  Partitions are summed in order by k_mlp_grad_reduce, so the gradient is
// deterministic (no atomics).
constexpr int MLP_PARTS = 32;
__global__ void __launch_bounds__(128) k_mlp_wgrad(MlpScratch s, int B, float *__restrict__ partial) {
    int r = blockIdx.x, part = blockIdx.y, k = threadIdx.x;
    const float *dA, *In;
    int nIn, ldA, row, wOff, bOff;
    if (r < MLP_H) {
        dA = s.da1; ldA = MLP_H; In = s.xs; nIn = MLP_IN; row = r; wOff = MLP_W1; bOff = MLP_B1;
    } else if (r < 2 * MLP_H) {
        dA = s.da2; ldA = MLP_H; In = s.h1; nIn = MLP_H; row = r - MLP_H; wOff = MLP_W2; bOff = MLP_B2;
    } else {
        dA = s.dy; ldA = MLP_OUT; In = s.h2; nIn = MLP_H; row = r - 2 * MLP_H; wOff = MLP_W3; bOff = MLP_B3;
    }
    if (k > nIn) return;
    int per = (B + MLP_PARTS - 1) / MLP_PARTS;
    int b0 = part * per, b1 = b0 + per < B ? b0 + per : B;
    float acc = 0.0f;
    float *out = partial + (size_t)part * MLP_NTRAIN;
    if (k < nIn) {
        for (int b = b0; b < b1; b++) acc += dA[(size_t)b * ldA + row] * In[(size_t)b * nIn + k];
        out[wOff + row * nIn + k] = acc;
    } else {
        for (int b = b0; b < b1; b++) acc += dA[(size_t)b * ldA + row];
        out[bOff + row] = acc;
    }
}

__global__ void __launch_bounds__(256) k_mlp_grad_reduce(const float *__restrict__ partial,
                                                         float *__restrict__ grad) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= MLP_NTRAIN) return;
    float acc = 0.0f;
    for (int p = 0; p < MLP_PARTS; p++) acc += partial[(size_t)p * MLP_NTRAIN + k];
    grad[k] = acc;
}

__global__ void __launch_bounds__(T) k_mlp_loss_sum(const float *__restrict__ loss, int B, float *__restrict__ out) {
    __shared__ float part[T];
    float acc = 0.0f;
    for (int b = threadIdx.x; b < B; b += T) acc += loss[b];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int d = T / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0] / (float)B;
}

__global__ void __launch_bounds__(T) k_mlp_adam(float *__restrict__ w, const float *__restrict__ grad,
                                                float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                float b2, float eps, float c1, float c2, int n) {
    int k = blockIdx.x * T + threadIdx.x;
    if (k >= n) return;
    float g = grad[k];
    float mk = b1 * m[k] + (1.0f - b1) * g;
    float vk = b2 * v[k] + (1.0f - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

}  // namespace

extern "C" {

int sbx_mlp_nparam(void) { return MLP_NPARAM; }

int sbx_mlp_forward(sbx_ctx *c, const float *weights, const float *x, int nSamples, float *y) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSamples >= 0, "mlp_forward: nSamples < 0");
    if (nSamples == 0) return SBX_OK;
    SBX_REQUIRE(c, weights && x && y, "mlp_forward: null array");
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(nSamples, T)), dim3(T), 0, c->stream, weights, x, nSamples, y);
    return check_launch(c, "k_mlp_forward");
}

int sbx_mlp_loss_grad(sbx_ctx *c, const float *weights, const float *x, const float *target, int nSamples,
                      const float *outWeight, float *scratch, float *grad, float *loss) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSamples > 0, "mlp_loss_grad: need nSamples > 0");
    SBX_REQUIRE(c, weights && x && target && scratch && grad && loss && outWeight, "mlp_loss_grad: null argument");
    size_t B = (size_t)nSamples;
    MlpScratch s;
    s.xs = scratch;
    s.h1 = s.xs + B * MLP_IN;
    s.h2 = s.h1 + B * MLP_H;
    s.g1 = s.h2 + B * MLP_H;
    s.g2 = s.g1 + B * MLP_H;
    s.da1 = s.g2 + B * MLP_H;
    s.da2 = s.da1 + B * MLP_H;
    s.dy = s.da2 + B * MLP_H;
    s.loss = s.dy + B * MLP_OUT;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(blocks_for(nSamples, T)), dim3(T), 0, c->stream, weights, x, target,
                       nSamples, outWeight[0], outWeight[1], s);
    float *partial = s.loss + B;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(2 * MLP_H + MLP_OUT, MLP_PARTS), dim3(128), 0, c->stream, s, nSamples,
                       partial);
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3(blocks_for(MLP_NTRAIN, 256)), dim3(256), 0, c->stream, partial, grad);
    hipLaunchKernelGGL(k_mlp_loss_sum, dim3(1), dim3(T), 0, c->stream, s.loss, nSamples, loss);
    return check_launch(c, "mlp_loss_grad");
}

size_t sbx_mlp_scratch_floats(int nSamples) {
    return (size_t)nSamples * (MLP_IN + 6 * MLP_H + MLP_OUT + 1) + (size_t)MLP_PARTS * MLP_NTRAIN;
}

int sbx_mlp_adam(sbx_ctx *c, float *weights, const float *grad, float *m, float *v, int t, float lr, float beta1,
                 float beta2, float eps) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, weights && grad && m && v && t >= 1, "mlp_adam: bad argument");
    float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(MLP_NTRAIN, T)), dim3(T), 0, c->stream, weights, grad, m, v, lr,
                       beta1, beta2, eps, c1, c2, MLP_NTRAIN);
    return check_launch(c, "k_mlp_adam");
}

}  // extern "C"
