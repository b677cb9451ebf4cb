// sbx_grid.h -- grid geometry and the per-body broad-phase query (SpatialGrid.java:51-406),
// shared by the stand-alone entry points and the fused network step.
#pragma once
#include <hip/hip_runtime.h>

#include "sbx_internal.h"

namespace sbx {

struct GridGeom {
    float xMin, yMin, zMin, invCell, cutoff;
    int nX, nY, nZ;
    int periodic;
    float Lx, Ly, Lz;  // periodic box lengths (minimum image)
    float iLx, iLy, iLz;  // 1.0f / L (host IEEE division)
};

// Java (int) of a float followed by the reference clamp to [0, nBins-1] (SpatialGrid.java:67-72);
// values below 0 and NaN (Java converts NaN to 0) land on 0.
__device__ __forceinline__ int clamp_cell(float v, int nBins) {
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)nBins) return nBins - 1;
    int c = (int)v;
    return c >= nBins ? nBins - 1 : c;
}

// periodic extension: floor, then wrap into [0, nBins)
__device__ __forceinline__ int wrap_cell(float v, int nBins) {
    if (!(v == v)) return 0;
    float fl = floorf(v);
    double m = fmod((double)fl, (double)nBins);
    if (m < 0) m += nBins;
    int c = (int)m;
    return c >= nBins ? 0 : c;
}

__device__ __forceinline__ int3 cell3_of(float x, float y, float z, const GridGeom &g) {
    float fx = (x - g.xMin) * g.invCell;
    float fy = (y - g.yMin) * g.invCell;
    float fz = (z - g.zMin) * g.invCell;
    if (g.periodic) return make_int3(wrap_cell(fx, g.nX), wrap_cell(fy, g.nY), wrap_cell(fz, g.nZ));
    return make_int3(clamp_cell(fx, g.nX), clamp_cell(fy, g.nY), clamp_cell(fz, g.nZ));
}

__device__ __forceinline__ int cell_of(float x, float y, float z, const GridGeom &g) {
    int3 c = cell3_of(x, y, z, g);
    return c.x + c.y * g.nX + c.z * g.nX * g.nY;
}

// minimum image of a coordinate difference: d - L * rint(d * (1/L)), rint = round half to even
__device__ __forceinline__ float min_image(float d, float L, float invL) { return d - L * rintf(d * invL); }

constexpr int BP_THREADS = 256;
constexpr int BP_BATCH = 4;

// The 27-cell stencil of cell c (SpatialGrid.java:362-380: (z, y, x) order, clamped or, in the
// periodic extension, wrapped) as contents ranges.  The three x-cells of one stencil row are
// adjacent in the cell-sorted contents, so a row is one range [offsets[first], offsets[last + 1])
// -- two when a periodic row wraps.  seg[18] receives the ranges in visiting order (empty ones as
// (0, 0)); all 36 bounds are loaded before any is used, so they are in flight together.  Returns
// the number of contents entries in the stencil.
__device__ __forceinline__ int stencil_ranges(int c, const GridGeom &g, const int *__restrict__ cellOffsets,
                                              int2 *seg) {
    const int nX = g.nX, nY = g.nY, nZ = g.nZ, nXY = nX * nY;
    const int cz = c / nXY;
    const int rem = c - cz * nXY;
    const int cy = rem / nX;
    const int cx = rem - cy * nX;
    // x extent of every row: [xa0, xa1] then (periodic wrap only) [xb0, xb1]
    int xa0, xa1, xb0 = 0, xb1 = -1;
    if (!g.periodic) {
        xa0 = cx > 0 ? cx - 1 : 0;
        xa1 = cx + 1 < nX ? cx + 1 : nX - 1;
    } else if (cx == 0) {
        xa0 = nX - 1; xa1 = nX - 1; xb0 = 0; xb1 = 1;
    } else if (cx == nX - 1) {
        xa0 = nX - 2; xa1 = nX - 1; xb0 = 0; xb1 = 0;
    } else {
        xa0 = cx - 1; xa1 = cx + 1;
    }
    const bool twoSeg = xb1 >= xb0;
    int total = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        int zz = cz + r / 3 - 1, yy = cy + r % 3 - 1;
        bool ok = true;
        if (g.periodic) {
            zz = zz < 0 ? zz + nZ : (zz >= nZ ? zz - nZ : zz);
            yy = yy < 0 ? yy + nY : (yy >= nY ? yy - nY : yy);
        } else {
            ok = zz >= 0 && zz < nZ && yy >= 0 && yy < nY;
        }
        int base = yy * nX + zz * nXY;
        int2 s0 = make_int2(0, 0), s1 = make_int2(0, 0);
        if (ok) {
            s0 = make_int2(cellOffsets[base + xa0], cellOffsets[base + xa1 + 1]);
            if (twoSeg) s1 = make_int2(cellOffsets[base + xb0], cellOffsets[base + xb1 + 1]);
        }
        seg[2 * r] = s0;
        seg[2 * r + 1] = s1;
        total += (s0.y - s0.x) + (s1.y - s1.x);
    }
    return total;
}

// Walks the contents indices of stencil_ranges in order; next() may be called exactly `total`
// times (the ranges live in LDS so they can be indexed dynamically).
struct RangeWalk {
    const int2 *seg;
    int s, idx, end;
    __device__ explicit RangeWalk(const int2 *sg) : seg(sg), s(0), idx(sg[0].x), end(sg[0].y) {}
    __device__ int next() {
        while (idx >= end && s < 17) {
            s++;
            int2 e = seg[s];
            idx = e.x;
            end = e.y;
        }
        return idx++;
    }
};

// SpatialGrid.broadPhase body loop (SpatialGrid.java:362-404): 27-cell stencil in (z, y, x)
// order, within-cell ascending, j > i, inclusive sphere predicate; returns the realised count
// (may exceed maxCand: surplus dropped, the count keeps going, :397-398).  One flat loop over all
// candidates (the wave iterates max over lanes of the candidate total, not the sum over the 18
// ranges of each range's maximum), BP_BATCH candidates at a time so their contents and centre
// loads are in flight together; tests and appends stay in order.
// segLds: 18 int2 of LDS per thread of the block (BP_THREADS x 18).
__device__ __forceinline__ int broad_phase_body(int i, int c, const float *__restrict__ center,
                                                const float *__restrict__ boundingRadius,
                                                const int *__restrict__ cellOffsets,
                                                const int *__restrict__ cellContents, const GridGeom &g, int cap,
                                                int *__restrict__ candPartner, int maxCand, int2 *segLds) {
    const float icx = center[i], icy = center[cap + i], icz = center[2 * cap + i];
    const float ri = boundingRadius[i];
    int2 *seg = segLds + threadIdx.x * 18;
    const int total = stencil_ranges(c, g, cellOffsets, seg);
    int cnt = 0;
    int *slice = candPartner + (size_t)i * (size_t)maxCand;
    RangeWalk w(seg);
    for (int q = 0; q < total; q += BP_BATCH) {
        int jj[BP_BATCH];
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) jj[u] = (q + u < total) ? cellContents[w.next()] : -1;
        float px[BP_BATCH], py[BP_BATCH], pz[BP_BATCH], pr[BP_BATCH];
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) {
            int j = jj[u] > i ? jj[u] : i;  // j <= i is skipped below; read a valid slot meanwhile
            px[u] = center[j];
            py[u] = center[cap + j];
            pz[u] = center[2 * cap + j];
            pr[u] = boundingRadius[j];
        }
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) {
            if (jj[u] <= i) continue;
            float dx = icx - px[u], dy = icy - py[u], dz = icz - pz[u];
            if (g.periodic) {
                dx = min_image(dx, g.Lx, g.iLx);
                dy = min_image(dy, g.Ly, g.iLy);
                dz = min_image(dz, g.Lz, g.iLz);
            }
            float distSq = dx * dx + dy * dy + dz * dz;
            float reach = ri + pr[u] + g.cutoff;
            if (distSq <= reach * reach) {
                if (cnt < maxCand) slice[cnt] = jj[u];
                cnt++;
            }
        }
    }
    return cnt;
}

GridGeom grid_geom(const float *gridParams, const int *gridDims, int periodic, const float *boxLen);
int check_grid(sbx_ctx *c, const float *gridParams, const int *gridDims, int periodic);
int launch_body_cell(sbx_ctx *c, const float *center, int cap, int S, const GridGeom &g, int *bodyCell);
int launch_broad_phase(sbx_ctx *c, const float *center, const float *boundingRadius, const int *bodyCell,
                       const int *cellOffsets, const int *cellContents, const GridGeom &g, int cap, int S,
                       int *candPartner, int *candCount, int maxCand, hipStream_t st = nullptr);

}  // namespace sbx
