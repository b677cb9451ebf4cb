// sbx_skin.hip -- training-data kernels of the skin model (features and labels from a trajectory
// of the per-step-rebuild run).  No counterpart in the reference.
//
// A sample is one coarse cell at a snapshot time t0, with its members binned at t0 - H:
//   features  skin_cell_features over the members' motion from t0 - H to t0 (what the learned
//             mode sees at a rebuild: the motion since the previous rebuild);
//   labels    out0 = log of the root mean square over members of each member's largest
//             displacement within H steps after t0 (the cell's displacement scale over the
//             horizon), out1 = log of the first step after t0 at which a member is 3 such scales
//             from where it started (censored at the tracking window).
#include "sbx_internal.h"
#include "sbx_skin.h"

using namespace sbx;

namespace {

constexpr int T = 256;

GridGeom min_image_geom(int periodic, const float *boxLen) {
    GridGeom g{};
    g.periodic = periodic;
    if (periodic) {
        g.Lx = boxLen[0]; g.Ly = boxLen[1]; g.Lz = boxLen[2];
        g.iLx = 1.0f / g.Lx; g.iLy = 1.0f / g.Ly; g.iLz = 1.0f / g.Lz;
    }
    return g;
}

__global__ void __launch_bounds__(T) k_features(const int *__restrict__ off, const int *__restrict__ cont, int nCells,
                                                int n, const float *__restrict__ cNow, const float *__restrict__ cPrev,
                                                const float *__restrict__ uNow, const float *__restrict__ uPrev,
                                                const int *__restrict__ links, float H, float cellSize, GridGeom g,
                                                float *__restrict__ out) {
    int c = blockIdx.x * T + threadIdx.x;
    if (c >= nCells) return;
    float f[SKIN_FEATURES];
    skin_cell_features(c, off, cont, n, cNow, cPrev, uNow, uPrev, links, H, cellSize, g, f);
#pragma unroll
    for (int k = 0; k < SKIN_FEATURES; k++) out[(size_t)c * SKIN_FEATURES + k] = f[k];
}

__global__ void __launch_bounds__(T) k_track(const float *__restrict__ coord, const float *__restrict__ x0, int n,
                                             GridGeom g, float *__restrict__ maxDisp, const float *__restrict__ thr,
                                             int *__restrict__ breach, int step) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    float d = sqrtf(disp_sq(coord, x0, i, n, g));
    if (maxDisp) maxDisp[i] = fmaxf(maxDisp[i], d);
    if (thr && breach[i] < 0 && d > thr[i]) breach[i] = step;
}

// mode 0: per-cell max of a per-body float; mode 1: per-cell min of a per-body int (>= 0 only);
// mode 2: per-cell root mean square of a per-body float; `fill` when no member qualifies
__global__ void __launch_bounds__(T) k_cell_reduce(const int *__restrict__ off, const int *__restrict__ cont,
                                                   int nCells, const float *__restrict__ fv,
                                                   const int *__restrict__ iv, int mode, float fill,
                                                   float *__restrict__ out) {
    int c = blockIdx.x * T + threadIdx.x;
    if (c >= nCells) return;
    float acc = mode == 1 ? fill : 0.0f;
    int any = 0;
    for (int p = off[c]; p < off[c + 1]; p++) {
        int i = cont[p];
        if (mode == 0) {
            acc = fmaxf(acc, fv[i]);
            any++;
        } else if (mode == 2) {
            acc += fv[i] * fv[i];
            any++;
        } else if (iv[i] >= 0) {
            acc = any ? fminf(acc, (float)iv[i]) : (float)iv[i];
            any++;
        }
    }
    if (mode == 2 && any) acc = sqrtf(acc / (float)any);
    out[c] = any ? acc : fill;
}

__global__ void __launch_bounds__(T) k_cell_scatter(const int *__restrict__ off, const int *__restrict__ cont,
                                                    int nCells, const float *__restrict__ cellVal,
                                                    float *__restrict__ bodyVal) {
    int c = blockIdx.x * T + threadIdx.x;
    if (c >= nCells) return;
    float v = cellVal[c];
    for (int p = off[c]; p < off[c + 1]; p++) bodyVal[cont[p]] = v;
}

}  // namespace

extern "C" {

int sbx_skin_features(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells, int n,
                      const float *coordNow, const float *coordPrev, const float *uNow, const float *uPrev,
                      const int *activeLinkCount, float H, float cellSize, int periodic, const float *boxLen,
                      float *features) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nCells >= 0 && n >= 0 && H > 0.0f && cellSize > 0.0f, "skin_features: bad sizes, H or cellSize");
    SBX_REQUIRE(c, !periodic || boxLen, "skin_features: periodic needs boxLen");
    if (nCells == 0) return SBX_OK;
    SBX_REQUIRE(c, coarseOffsets && coarseContents && coordNow && coordPrev && uNow && uPrev && features,
                "skin_features: null array");
    hipLaunchKernelGGL(k_features, dim3(blocks_for(nCells, T)), dim3(T), 0, c->stream, coarseOffsets, coarseContents,
                       nCells, n, coordNow, coordPrev, uNow, uPrev, activeLinkCount, H, cellSize,
                       min_image_geom(periodic, boxLen), features);
    return check_launch(c, "k_features");
}

int sbx_skin_track(sbx_ctx *c, const float *coord, const float *x0, int n, int periodic, const float *boxLen,
                   float *maxDisp, const float *threshold, int *breachStep, int step) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && (!periodic || boxLen), "skin_track: bad arguments");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && x0 && (maxDisp || threshold), "skin_track: null array");
    SBX_REQUIRE(c, !threshold || breachStep, "skin_track: threshold needs breachStep");
    hipLaunchKernelGGL(k_track, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, x0, n,
                       min_image_geom(periodic, boxLen), maxDisp, threshold, breachStep, step);
    return check_launch(c, "k_track");
}

int sbx_skin_cell_reduce(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells,
                         const float *bodyFloat, const int *bodyInt, int mode, float fill, float *cellOut) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, mode >= 0 && mode <= 2, "skin_cell_reduce: mode 0 (max float), 1 (min int) or 2 (rms float)");
    SBX_REQUIRE(c, nCells >= 0, "skin_cell_reduce: nCells < 0");
    if (nCells == 0) return SBX_OK;
    SBX_REQUIRE(c, coarseOffsets && coarseContents && cellOut && (mode != 1 ? bodyFloat != nullptr : bodyInt != nullptr),
                "skin_cell_reduce: null array");
    hipLaunchKernelGGL(k_cell_reduce, dim3(blocks_for(nCells, T)), dim3(T), 0, c->stream, coarseOffsets,
                       coarseContents, nCells, bodyFloat, bodyInt, mode, fill, cellOut);
    return check_launch(c, "k_cell_reduce");
}

int sbx_skin_cell_scatter(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells,
                          const float *cellVal, float *bodyVal) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nCells >= 0, "skin_cell_scatter: nCells < 0");
    if (nCells == 0) return SBX_OK;
    SBX_REQUIRE(c, coarseOffsets && coarseContents && cellVal && bodyVal, "skin_cell_scatter: null array");
    hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(nCells, T)), dim3(T), 0, c->stream, coarseOffsets,
                       coarseContents, nCells, cellVal, bodyVal);
    return check_launch(c, "k_cell_scatter");
}

}  // extern "C"
