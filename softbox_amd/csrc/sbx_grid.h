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

// One replica of a batched network step: its grid, and where its cells, rods and links start in the
// batch's concatenated arrays (cells are numbered replica after replica); seed keys its draws.
struct RepInfo {
    GridGeom g;
    int cellBase, rodBase, linkBase, seed;
};

// The grid a body is binned in.  SingleGeo: one grid for every body; BatchGeo: the grid of the
// body's replica, its cells offset by the replica's cell base.
struct BodyGeo {
    GridGeom g;
    int cellBase;
};
struct SingleGeo {
    GridGeom g;
    __device__ BodyGeo of(int) const { return BodyGeo{g, 0}; }
};
struct BatchGeo {
    const RepInfo *reps;
    const int *rodRep;
    __device__ BodyGeo of(int i) const {
        const RepInfo &r = reps[rodRep[i]];
        return BodyGeo{r.g, r.cellBase};
    }
};

// Java (int) of a float followed by the reference clamp to [0, nBins-1] (SpatialGrid.java:67-72);
// values below 0 and NaN (Java converts NaN to 0) land on 0.
__device__ __forceinline__ int clamp_cell(float v, int nBins) {
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)nBins) return nBins - 1;
    int c = (int)v;
    return c >= nBins ? nBins - 1 : c;
}

// periodic extension: floor, then wrap into [0, nBins) -- the floor's remainder mod nBins (fmod in
// double, negative remainders shifted up).  Within one period either side of the grid (every pose
// of a periodic run) the remainder is one integer add or subtract; fmod takes the rest.
__device__ __noinline__ inline int wrap_far(float fl, int nBins) {
    double m = fmod((double)fl, (double)nBins);
    if (m < 0) m += nBins;
    int c = (int)m;
    return c >= nBins ? 0 : c;
}

__device__ __forceinline__ int wrap_cell(float v, int nBins) {
    if (!(v == v)) return 0;
    float fl = floorf(v);
    if (fl >= (float)(-nBins) && fl < (float)(2 * nBins)) {  // |fl| < 2^24: the int is exact
        int c = (int)fl;
        return c < 0 ? c + nBins : (c >= nBins ? c - nBins : c);
    }
    return wrap_far(fl, nBins);
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
constexpr int BP_BATCH = 8;  // candidates of one walk loaded together

// The 27-cell stencil of a cell (SpatialGrid.java:362-380: (z, y, x) order, clamped or, in the
// periodic extension, wrapped) as a walk over the cell-sorted contents.  The three x-cells of one
// stencil row are adjacent in the contents, so a row is one range [offsets[first], offsets[last +
// 1]) -- two when a periodic row wraps: 18 ranges in visiting order (empty ones as zero length).
// The walk is held as a position map in registers (two int[18] locals of the caller -- plain
// arrays, which the compiler keeps in registers): entry k (0 <= k < total) is at contents position
// k + shift[r], r the last range with pre[r] <= k (an empty range shares its start with the next
// one, so the last such r is never empty).  Entries of one walk are independent, so a batch of
// them can be loaded at once.
__device__ __forceinline__ int stencil_at(const int *pre, const int *shift, int k) {
    int sh = shift[0];
#pragma unroll
    for (int r = 1; r < 18; r++) sh = (k >= pre[r]) ? shift[r] : sh;
    return k + sh;
}

// stencil_at for a walk whose second segments are all empty (no periodic row wraps: the cell is not
// in the first or last x column): an empty range owns no position, so the 9 first segments decide.
__device__ __forceinline__ int stencil_at9(const int *pre, const int *shift, int k) {
    int sh = shift[0];
#pragma unroll
    for (int r = 1; r < 9; r++) sh = (k >= pre[2 * r]) ? shift[2 * r] : sh;
    return k + sh;
}

// The 18 ranges of cell c into pre/shift (all 36 bounds are loaded before any is used, so they are
// in flight together); returns the walk's length.  *wraps: some row has a second segment.
__device__ __forceinline__ int stencil_map(int c, const GridGeom &g, const int *__restrict__ cellOffsets, int *pre,
                                           int *shift, bool *wraps = nullptr) {
    const int nX = g.nX, nY = g.nY, nZ = g.nZ, nXY = nX * nY;
    const int cz = c / nXY;
    const int rem = c - cz * nXY;
    const int cy = rem / nX;
    const int cx = rem - cy * nX;
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
    if (wraps) *wraps = twoSeg;
    // the three z planes and three y rows of the stencil (wrapped or range-checked once per axis)
    int zBase[3], yBase[3];
    bool zOk[3], yOk[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int zz = cz + d - 1, yy = cy + d - 1;
        if (g.periodic) {
            zz = zz < 0 ? zz + nZ : (zz >= nZ ? zz - nZ : zz);
            yy = yy < 0 ? yy + nY : (yy >= nY ? yy - nY : yy);
            zOk[d] = yOk[d] = true;
        } else {
            zOk[d] = zz >= 0 && zz < nZ;
            yOk[d] = yy >= 0 && yy < nY;
        }
        zBase[d] = zz * nXY;
        yBase[d] = yy * nX;
    }
    int lo[18], hi[18];
#pragma unroll
    for (int r = 0; r < 9; r++) {
        const bool ok = zOk[r / 3] && yOk[r % 3];
        const int base = yBase[r % 3] + zBase[r / 3];
        lo[2 * r] = ok ? cellOffsets[base + xa0] : 0;
        hi[2 * r] = ok ? cellOffsets[base + xa1 + 1] : 0;
        lo[2 * r + 1] = ok && twoSeg ? cellOffsets[base + xb0] : 0;
        hi[2 * r + 1] = ok && twoSeg ? cellOffsets[base + xb1 + 1] : 0;
    }
    int run = 0;
#pragma unroll
    for (int r = 0; r < 18; r++) {
        pre[r] = run;
        shift[r] = lo[r] - run;
        run += hi[r] - lo[r];
    }
    return run;
}

GridGeom grid_geom(const float *gridParams, const int *gridDims, int periodic, const float *boxLen);
int check_grid(sbx_ctx *c, const float *gridParams, const int *gridDims, int periodic);
int launch_body_cell(sbx_ctx *c, const float *center, int cap, int S, const GridGeom &g, int *bodyCell);
int launch_broad_phase(sbx_ctx *c, const float *center, const float *boundingRadius, const int *bodyCell,
                       const int *cellOffsets, const int *cellContents, const GridGeom &g, int cap, int S,
                       int *candPartner, int *candCount, int maxCand, hipStream_t st = nullptr);
// the same from {centre, bounding radius} and cell laid out in contents order (sortedPose[q] and
// sortedCell[q] belong to body cellContents[q]; counting_sort's payload)
// reps == nullptr: every body in grid g; else body i in the grid of replica rodRep[i] (g unused).
// group: lanes per body, 2, 4 or 8; staged: stage each block's neighbourhood in LDS (4 or 8 lanes).
int launch_broad_phase_sorted(sbx_ctx *c, const float4 *sortedPose, const int *sortedCell, const int *cellOffsets,
                              const int *cellContents, const GridGeom &g, const RepInfo *reps, const int *rodRep,
                              int group, int staged, int cap, int S, int *candPartner, int *candCount, int maxCand,
                              hipStream_t st = nullptr);

}  // namespace sbx
