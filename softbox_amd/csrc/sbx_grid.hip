// sbx_grid.hip -- spatial grid binning, stable cell CSR and 27-cell broad phase on gfx950.
//
// Reference: SpatialGrid.java:51-445 (+ FilamentStore.publishToBodyView :275-295 and the CSR
// inverse of CrossBridgeSystem.java:190-298, which shares the counting sort in sbx_core.hip).
// The broad phase walks bodies in cell-sorted order (thread t handles gridCellContents[t]) so that
// a wavefront's 27-cell neighbourhoods overlap and the offsets/contents/centre reads stay in L2.
#include "sbx_device.h"
#include "sbx_grid.h"
#include "sbx_internal.h"

namespace sbx {

__global__ void __launch_bounds__(256) k_publish(const float *__restrict__ coord, const float *__restrict__ segLength,
                                                 float *__restrict__ center, float *__restrict__ boundingRadius,
                                                 int *__restrict__ ownerStore, int *__restrict__ ownerSlot,
                                                 float actinRadius, int n, int cap) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    center[i] = coord[i];
    center[cap + i] = coord[n + i];
    center[2 * cap + i] = coord[2 * n + i];
    boundingRadius[i] = 0.5f * segLength[i] + actinRadius;
    if (ownerStore) ownerStore[i] = 0;
    if (ownerSlot) ownerSlot[i] = i;
}

__global__ void __launch_bounds__(256) k_body_cell(const float *__restrict__ center, int cap, int S, GridGeom g,
                                                   int *__restrict__ bodyCell) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cap) return;
    if (i >= S) { bodyCell[i] = -1; return; }
    bodyCell[i] = cell_of(center[i], center[cap + i], center[2 * cap + i], g);
}

__global__ void __launch_bounds__(BP_THREADS) k_broad_phase(const float *__restrict__ center,
                                                     const float *__restrict__ boundingRadius,
                                                     const int *__restrict__ bodyCell,
                                                     const int *__restrict__ cellOffsets,
                                                     const int *__restrict__ cellContents, GridGeom g, int cap, int S,
                                                     int nSorted, int *__restrict__ candPartner,
                                                     int *__restrict__ candCount, int maxCand) {
    // XCD-aware: each XCD walks one contiguous run of the cell-sorted bodies (a spatial slab), so
    // its L2 holds that slab's cell offsets and centres instead of the whole grid
    int t = xcd_block(blockIdx.x, gridDim.x) * BP_THREADS + threadIdx.x;
    if (t >= cap) return;
    if (t >= nSorted) {  // padding slots [S, cap): count cleared, nothing emitted
        candCount[t] = 0;
        return;
    }
    __shared__ int2 segLds[BP_THREADS * 18];
    int i = cellContents[t];  // cell-sorted order: neighbouring lanes query neighbouring cells
    int c = bodyCell[i];
    if (c < 0) { candCount[i] = 0; return; }
    candCount[i] = broad_phase_body(i, c, center, boundingRadius, cellOffsets, cellContents, g, cap, candPartner,
                                    maxCand, segLds);
}

GridGeom grid_geom(const float *gridParams, const int *gridDims, int periodic, const float *boxLen) {
    GridGeom g;
    g.xMin = gridParams[0]; g.yMin = gridParams[1]; g.zMin = gridParams[2];
    g.invCell = gridParams[4];
    g.cutoff = gridParams[5];
    g.nX = gridDims[0]; g.nY = gridDims[1]; g.nZ = gridDims[2];
    g.periodic = periodic;
    if (periodic) {
        if (boxLen) { g.Lx = boxLen[0]; g.Ly = boxLen[1]; g.Lz = boxLen[2]; }
        else {
            g.Lx = gridParams[3] * g.nX; g.Ly = gridParams[3] * g.nY; g.Lz = gridParams[3] * g.nZ;
        }
        g.iLx = 1.0f / g.Lx; g.iLy = 1.0f / g.Ly; g.iLz = 1.0f / g.Lz;
    } else {
        g.Lx = g.Ly = g.Lz = 0.0f;
        g.iLx = g.iLy = g.iLz = 0.0f;
    }
    return g;
}

int check_grid(sbx_ctx *c, const float *gridParams, const int *gridDims, int periodic) {
    SBX_REQUIRE(c, gridParams && gridDims, "grid: null gridParams/gridDims");
    SBX_REQUIRE(c, gridDims[0] > 0 && gridDims[1] > 0 && gridDims[2] > 0, "grid: non-positive bin counts");
    long long tot = (long long)gridDims[0] * gridDims[1] * gridDims[2];
    SBX_REQUIRE(c, tot == gridDims[3], "grid: gridDims[3]=%d != nX*nY*nZ=%lld", gridDims[3], tot);
    SBX_REQUIRE(c, tot < (1LL << 30), "grid: too many cells (%lld)", tot);
    SBX_REQUIRE(c, gridParams[4] > 0.0f, "grid: invCellSize must be > 0");
    if (periodic)
        SBX_REQUIRE(c, gridDims[0] >= 3 && gridDims[1] >= 3 && gridDims[2] >= 3,
                    "grid: periodic mode needs >= 3 bins per axis");
    return SBX_OK;
}

int launch_body_cell(sbx_ctx *c, const float *center, int cap, int S, const GridGeom &g, int *bodyCell) {
    hipLaunchKernelGGL(k_body_cell, dim3(blocks_for(cap, 256)), dim3(256), 0, c->stream, center, cap, S, g, bodyCell);
    return check_launch(c, "k_body_cell");
}

int launch_broad_phase(sbx_ctx *c, const float *center, const float *boundingRadius, const int *bodyCell,
                       const int *cellOffsets, const int *cellContents, const GridGeom &g, int cap, int S,
                       int *candPartner, int *candCount, int maxCand, hipStream_t st) {
    // bodies with a cell are exactly the first S (counting sort places each live body once)
    hipLaunchKernelGGL(k_broad_phase, dim3(blocks_for(cap, BP_THREADS)), dim3(BP_THREADS), 0, st ? st : c->stream,
                       center,
                       boundingRadius,
                       bodyCell, cellOffsets, cellContents, g, cap, S, S, candPartner, candCount, maxCand);
    return check_launch(c, "k_broad_phase");
}

}  // namespace sbx

using namespace sbx;

extern "C" {

int sbx_publish_filaments(sbx_ctx *c, const float *coord, const float *segLength, float *center,
                          float *boundingRadius, int *ownerStore, int *ownerSlot, const float *viewParams, int n,
                          int cap) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && cap >= n && viewParams, "publish_filaments: need 0 <= n <= cap and viewParams");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && segLength && center && boundingRadius, "publish_filaments: null array");
    hipLaunchKernelGGL(k_publish, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, coord, segLength, center,
                       boundingRadius, ownerStore, ownerSlot, viewParams[0], n, cap);
    return check_launch(c, "k_publish");
}

int sbx_body_cell(sbx_ctx *c, const float *center, const float *gridParams, const int *gridDims, const int *counts,
                  int *bodyCell, int cap, int periodic) {
    SBX_CHECK_CTX(c);
    int r = check_grid(c, gridParams, gridDims, periodic);
    if (r) return r;
    SBX_REQUIRE(c, counts && cap >= 0, "body_cell: null counts or cap < 0");
    int S = counts[1];
    SBX_REQUIRE(c, S >= 0 && S <= cap, "body_cell: S=%d outside [0, cap=%d]", S, cap);
    if (cap == 0) return SBX_OK;
    SBX_REQUIRE(c, center && bodyCell, "body_cell: null array");
    GridGeom g = grid_geom(gridParams, gridDims, periodic, nullptr);
    return launch_body_cell(c, center, cap, S, g, bodyCell);
}

int sbx_grid_build(sbx_ctx *c, const int *bodyCell, const int *counts, const int *gridDims, int *gridCellOffsets,
                   int *gridCellContents, int contentsCap) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, counts && gridDims && gridCellOffsets, "grid_build: null argument");
    int S = counts[1];
    SBX_REQUIRE(c, S >= 0 && contentsCap >= S, "grid_build: need 0 <= S <= contentsCap");
    SBX_REQUIRE(c, gridDims[3] > 0, "grid_build: totalCells must be > 0");
    SBX_REQUIRE(c, S == 0 || (bodyCell && gridCellContents), "grid_build: null array");
    return counting_sort(c, c->api_sort, bodyCell, S, gridDims[3], contentsCap, gridCellOffsets, gridCellContents);
}

int sbx_broad_phase(sbx_ctx *c, const float *center, const float *boundingRadius, const int *bodyCell,
                    const int *gridCellOffsets, const int *gridCellContents, const int *gridDims,
                    const float *gridParams, const int *counts, int *candPartner, int *candCount, int cap,
                    int maxCand, int periodic) {
    SBX_CHECK_CTX(c);
    int r = check_grid(c, gridParams, gridDims, periodic);
    if (r) return r;
    SBX_REQUIRE(c, counts && cap >= 0 && maxCand > 0, "broad_phase: bad counts/cap/maxCand");
    int S = counts[1];
    SBX_REQUIRE(c, S >= 0 && S <= cap, "broad_phase: S outside [0, cap]");
    if (cap == 0) return SBX_OK;
    SBX_REQUIRE(c, center && boundingRadius && bodyCell && gridCellOffsets && gridCellContents && candPartner &&
                       candCount,
                "broad_phase: null array");
    GridGeom g = grid_geom(gridParams, gridDims, periodic, nullptr);
    return launch_broad_phase(c, center, boundingRadius, bodyCell, gridCellOffsets, gridCellContents, g, cap, S,
                              candPartner, candCount, maxCand);
}

int sbx_csr_build(sbx_ctx *c, const int *key, int nItems, int nKeys, int *offsets, int *list) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nItems >= 0 && nKeys >= 0 && offsets, "csr_build: bad sizes or null offsets");
    SBX_REQUIRE(c, nItems == 0 || (key && list), "csr_build: null key/list");
    return counting_sort(c, c->api_sort, key, nItems, nKeys, nItems, offsets, list);
}

}  // extern "C"
