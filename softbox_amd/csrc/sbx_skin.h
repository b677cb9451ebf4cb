// sbx_skin.h -- the skin (Verlet-list) broad phase: device helpers shared by the fused network
// step and the training entry points.
//
// No counterpart in the reference: SpatialGrid rebuilds its grid and re-runs broadPhase every
// step.  Here the candidate slices of SpatialGrid.broadPhase (SpatialGrid.java:340-404) are
// produced from per-body neighbour lists built at a rebuild with reach + s_i + s_j (s = per-body
// skin), filtered every step with the exact sphere predicate on the current centres and put in the
// broad phase's own order (fine-grid stencil position (z, y, x), then ascending index).  A body
// whose centre moved more than its skin since the rebuild goes onto an outside-skin list that every
// body tests, and finds its own candidates from the rebuild's coarse grid, so every pair within
// reach is found: the slices are identical to the per-step rebuild's.  Rebuilds are the host's.
#pragma once
#include "sbx_grid.h"

namespace sbx {

constexpr int SKIN_FEATURES = 12;
// device flag words of the skin state
constexpr int SKF_REBUILD = 0;     // 1 while the host's rebuild runs (the rebuild kernels' gate)
constexpr int SKF_SINCE = 1;       // steps since the last rebuild
constexpr int SKF_REBUILDS = 2;    // rebuilds so far
constexpr int SKF_OVERFLOW = 3;    // bodies whose list overflowed at the last rebuild
constexpr int SKF_ENTRIES = 4;     // list entries at the last rebuild (truncated to 32 bits)
constexpr int SKF_STEPS = 5;       // steps in skin mode so far
constexpr int SKF_EXTRA = 8;       // bodies outside their skins since the last rebuild (> extraCap: overflow)
constexpr int SKF_BRUTE = 9;       // steps in which every body tested every other (outside-skin list overflow)
constexpr int SKF_SELECT = 10;     // body-steps done by a whole block (more than 64 candidates)
constexpr int SKF_GEN = 11;        // generation of the last rebuild (the host's count)
constexpr int SKF_WORDS = 16;

// q = c / d (c >= 0, d > 0, q < 2^20) by a float reciprocal and exact integer correction
__device__ __forceinline__ int div_small(int c, int d, float invD) {
    int q = (int)((float)c * invD);
    while (q > 0 && q * d > c) q--;
    while ((q + 1) * d <= c) q++;
    return q;
}

// (x, y, z) of linear cell c
__device__ __forceinline__ int3 cell_xyz(int c, const GridGeom &g) {
    const int nXY = g.nX * g.nY;
    int z = div_small(c, nXY, 1.0f / (float)nXY);
    int r = c - z * nXY;
    int y = div_small(r, g.nX, 1.0f / (float)g.nX);
    return make_int3(r - y * g.nX, y, z);
}

// stencil position 0..26 of fine cell j relative to cell i in (z, y, x) order, or -1 when j is not
// one of the (up to) 27 cells SpatialGrid.broadPhase visits for i
__device__ __forceinline__ int stencil_pos(int3 ci, int3 j, const GridGeom &g) {
    int dx = j.x - ci.x, dy = j.y - ci.y, dz = j.z - ci.z;
    if (g.periodic) {  // >= 3 bins per axis, so each offset maps to one of -1, 0, +1 at most once
        if (dx == g.nX - 1) dx = -1; else if (dx == 1 - g.nX) dx = 1;
        if (dy == g.nY - 1) dy = -1; else if (dy == 1 - g.nY) dy = 1;
        if (dz == g.nZ - 1) dz = -1; else if (dz == 1 - g.nZ) dz = 1;
    }
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
}

// |min-image(a - b)|^2 of two planar points
__device__ __forceinline__ float disp_sq(const float *__restrict__ a, const float *__restrict__ b, int i, int n,
                                         const GridGeom &g) {
    float dx = a[i] - b[i], dy = a[n + i] - b[n + i], dz = a[2 * n + i] - b[2 * n + i];
    if (g.periodic) {
        dx = min_image(dx, g.Lx, g.iLx);
        dy = min_image(dy, g.Ly, g.iLy);
        dz = min_image(dz, g.Lz, g.iLz);
    }
    return dx * dx + dy * dy + dz * dz;
}

// The 12 skin-model features of one coarse cell (edge cellSize um) from its members' motion over H
// steps:
//  0 log max disp/sqrt(H)   1 log mean disp/sqrt(H)   2 log max disp/H   3 log mean disp/H
//  4 log max |du|/sqrt(H)   5 log mean |du|/sqrt(H)   6 log members per um^3
//  7 fraction of members with an active crosslink     8 mean active crosslinks per member
//  9 log H                 10 log cellSize           11 log(1 + members)
// Empty cells get zeros.  Members are read in CSR order, so the result is deterministic.
__device__ __forceinline__ void skin_cell_features(int c, const int *__restrict__ off, const int *__restrict__ cont,
                                                   int n, const float *__restrict__ cNow,
                                                   const float *__restrict__ cPrev, const float *__restrict__ uNow,
                                                   const float *__restrict__ uPrev, const int *__restrict__ links,
                                                   float H, float cellSize, const GridGeom &g,
                                                   float f[SKIN_FEATURES]) {
    int a = off[c], b = off[c + 1];
    int cnt = b - a;
#pragma unroll
    for (int k = 0; k < SKIN_FEATURES; k++) f[k] = 0.0f;
    if (cnt <= 0) return;
    float sumD = 0.0f, maxD = 0.0f, sumR = 0.0f, maxR = 0.0f, bound = 0.0f, nl = 0.0f;
    for (int p = a; p < b; p++) {
        int i = cont[p];
        float d = sqrtf(disp_sq(cNow, cPrev, i, n, g));
        float ux = uNow[i] - uPrev[i], uy = uNow[n + i] - uPrev[n + i], uz = uNow[2 * n + i] - uPrev[2 * n + i];
        float r = sqrtf(ux * ux + uy * uy + uz * uz);
        sumD += d;
        maxD = fmaxf(maxD, d);
        sumR += r;
        maxR = fmaxf(maxR, r);
        if (links) {
            int l = links[i];
            bound += l > 0 ? 1.0f : 0.0f;
            nl += (float)l;
        }
    }
    float inv = 1.0f / (float)cnt;
    float sq = sqrtf(H), meanD = sumD * inv;
    const float e = 1.0e-6f;
    f[0] = logf(e + maxD / sq);
    f[1] = logf(e + meanD / sq);
    f[2] = logf(e + maxD / H);
    f[3] = logf(e + meanD / H);
    f[4] = logf(e + maxR / sq);
    f[5] = logf(e + sumR * inv / sq);
    f[6] = logf((float)cnt / (cellSize * cellSize * cellSize));
    f[7] = bound * inv;
    f[8] = nl * inv;
    f[9] = logf(H);
    f[10] = logf(cellSize);
    f[11] = logf(1.0f + (float)cnt);
}

}  // namespace sbx
