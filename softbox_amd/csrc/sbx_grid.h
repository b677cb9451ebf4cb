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

__device__ __forceinline__ int cell_of(float x, float y, float z, const GridGeom &g) {
    float fx = (x - g.xMin) * g.invCell;
    float fy = (y - g.yMin) * g.invCell;
    float fz = (z - g.zMin) * g.invCell;
    int ix, iy, iz;
    if (g.periodic) {
        ix = wrap_cell(fx, g.nX); iy = wrap_cell(fy, g.nY); iz = wrap_cell(fz, g.nZ);
    } else {
        ix = clamp_cell(fx, g.nX); iy = clamp_cell(fy, g.nY); iz = clamp_cell(fz, g.nZ);
    }
    return ix + iy * g.nX + iz * g.nX * g.nY;
}

__device__ __forceinline__ float min_image(float d, float L) { return d - L * rintf(d / L); }

// SpatialGrid.broadPhase body loop (SpatialGrid.java:362-404): 27-cell stencil in (z, y, x)
// order, within-cell ascending, j > i, inclusive sphere predicate; returns the realised count
// (may exceed maxCand: surplus dropped, the count keeps going, :397-398).
__device__ __forceinline__ int broad_phase_body(int i, int c, const float *__restrict__ center,
                                                const float *__restrict__ boundingRadius,
                                                const int *__restrict__ cellOffsets,
                                                const int *__restrict__ cellContents, const GridGeom &g, int cap,
                                                int *__restrict__ candPartner, int maxCand) {
    int nX = g.nX, nY = g.nY, nZ = g.nZ, nXY = nX * nY;
    int cz = c / nXY;
    int rem = c - cz * nXY;
    int cy = rem / nX;
    int cx = rem - cy * nX;
    float icx = center[i], icy = center[cap + i], icz = center[2 * cap + i];
    float ri = boundingRadius[i];
    int x0 = cx - 1, x1 = cx + 1, y0 = cy - 1, y1 = cy + 1, z0 = cz - 1, z1 = cz + 1;
    if (!g.periodic) {
        if (x0 < 0) x0 = 0;
        if (x1 >= nX) x1 = nX - 1;
        if (y0 < 0) y0 = 0;
        if (y1 >= nY) y1 = nY - 1;
        if (z0 < 0) z0 = 0;
        if (z1 >= nZ) z1 = nZ - 1;
    }
    int cnt = 0;
    size_t slice = (size_t)i * (size_t)maxCand;
    for (int zz = z0; zz <= z1; zz++) {
        int wz = g.periodic ? (zz + nZ) % nZ : zz;
        for (int yy = y0; yy <= y1; yy++) {
            int wy = g.periodic ? (yy + nY) % nY : yy;
            for (int xx = x0; xx <= x1; xx++) {
                int wx = g.periodic ? (xx + nX) % nX : xx;
                int cc = wx + wy * nX + wz * nXY;
                int start = cellOffsets[cc];
                int end = cellOffsets[cc + 1];
                for (int idx = start; idx < end; idx++) {
                    int j = cellContents[idx];
                    if (j <= i) continue;
                    float dx = icx - center[j];
                    float dy = icy - center[cap + j];
                    float dz = icz - center[2 * cap + j];
                    if (g.periodic) {
                        dx = min_image(dx, g.Lx);
                        dy = min_image(dy, g.Ly);
                        dz = min_image(dz, g.Lz);
                    }
                    float distSq = dx * dx + dy * dy + dz * dz;
                    float reach = ri + boundingRadius[j] + g.cutoff;
                    if (distSq <= reach * reach) {
                        if (cnt < maxCand) candPartner[slice + cnt] = j;
                        cnt++;
                    }
                }
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
                       int *candPartner, int *candCount, int maxCand);

}  // namespace sbx
