// sbx_grid.hip -- spatial grid binning, stable cell CSR and 27-cell broad phase on gfx950.
//
// Reference: SpatialGrid.java:51-445 (+ FilamentStore.publishToBodyView :275-295 and the CSR
// inverse of CrossBridgeSystem.java:190-298, which shares the counting sort in sbx_core.hip).
// The broad phase walks bodies in cell-sorted order (thread t handles gridCellContents[t]) so that
// a wavefront's 27-cell neighbourhoods overlap and the offsets/contents/centre reads stay in L2.
#include "sbx_device.h"
#include "sbx_grid.h"
#include <type_traits>
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

// Pose of a candidate: at(q, j) = {centre, bounding radius} of body j = contents[q]; own(t, i) and
// cell(t, i) for the querying body i = contents[t].
// PlanarPose: planar centre[3][cap] + radius[cap] and bodyCell, by body index.
struct PlanarPose {
    const float *c, *r;
    const int *bodyCell;
    int cap;
    __device__ float4 at(int, int j) const { return make_float4(c[j], c[cap + j], c[2 * cap + j], r[j]); }
    __device__ float4 own(int, int i) const { return at(0, i); }
    __device__ int cell(int, int i) const { return bodyCell[i]; }
};

// SortedPose: pose and cell already laid out in contents order (by the sort's place pass), so a
// candidate's pose sits at its contents position -- no gather through the body index.
struct SortedPose {
    const float4 *p;
    const int *cellOf;
    __device__ float4 at(int q, int) const { return p[q]; }
    __device__ float4 own(int t, int) const { return p[t]; }
    __device__ int cell(int t, int) const { return cellOf[t]; }
};

// Candidate test of the broad phase (SpatialGrid.java:383-396: minimum image in the periodic
// extension, inclusive sphere predicate on ri + rj + cutoff).
__device__ __forceinline__ bool bp_hit(float4 qi, float4 qj, const GridGeom &g) {
    float dx = qi.x - qj.x, dy = qi.y - qj.y, dz = qi.z - qj.z;
    if (g.periodic) {
        dx = min_image(dx, g.Lx, g.iLx);
        dy = min_image(dy, g.Ly, g.iLy);
        dz = min_image(dz, g.Lz, g.iLz);
    }
    float distSq = dx * dx + dy * dy + dz * dz;
    float reach = qi.w + qj.w + g.cutoff;
    return distSq <= reach * reach;
}

// SpatialGrid.broadPhase (SpatialGrid.java:340-406).  A group of G consecutive lanes takes one
// body of the cell-sorted contents (neighbouring groups query neighbouring cells, so offsets,
// contents and poses stay in cache) and walks its stencil -- 18 contents ranges in the reference's
// visiting order ((z, y, x) cells, ascending index within a cell), mapped to positions in registers
// -- G * U consecutive positions per round: lane l of the group takes positions base + k G + l
// (k < U), so a group's loads are contiguous.  Hits (j > i and the test) are ordered by position
// with a ballot: a hit's slot is the group's count so far plus the hits at earlier positions of the
// round.  Counts keep going past maxCand (SpatialGrid.java:397-398).
constexpr int BP_GROUP = 4;  // lanes per body (stand-alone entry point; the network step chooses 2 or 4)
constexpr int BP_ROUND = 4;  // positions per lane and round

template <typename Pose, typename Geo, int G, int U>
__global__ void __launch_bounds__(BP_THREADS) k_broad_phase(Pose pose, Geo geo, const int *__restrict__ cellOffsets,
                                                            const int *__restrict__ cellContents, int cap,
                                                            int nSorted, int *__restrict__ candPartner,
                                                            int *__restrict__ candCount, int maxCand) {
    static_assert(64 % G == 0, "groups tile a wave");
    // XCD-aware: each XCD walks one contiguous run of the cell-sorted bodies (a spatial slab), so
    // its L2 holds that slab's cell offsets and centres instead of the whole grid
    const int lane = threadIdx.x & 63, gl = lane % G, grp = lane / G;
    const int wave = (xcd_block(blockIdx.x, gridDim.x) * BP_THREADS + threadIdx.x) >> 6;
    const int t = wave * (64 / G) + grp;
    if (t >= cap) return;  // whole groups leave together (t is uniform in a group)
    if (t >= nSorted) {    // padding slots [S, cap): count cleared, nothing emitted
        if (gl == 0) candCount[t] = 0;
        return;
    }
    const int i = cellContents[t];
    const int c = pose.cell(t, i);
    if (c < 0) {
        if (gl == 0) candCount[i] = 0;
        return;
    }
    const float4 qi = pose.own(t, i);
    const BodyGeo bg = geo.of(i);  // the body's grid; its cells start at bg.cellBase
    const GridGeom &g = bg.g;
    int pre[18], shift[18];
    bool wraps;
    const int total = stencil_map(c - bg.cellBase, g, cellOffsets + bg.cellBase, pre, shift, &wraps);
    // wave-uniform: without a wrapped row among the wave's bodies, the 9-segment map gives the same
    // positions for fewer compares
    const bool anyWrap = __ballot(wraps) != 0ull;
    int *slice = candPartner + (size_t)i * (size_t)maxCand;
    const unsigned long long groupBits = G == 64 ? ~0ull : ((1ull << G) - 1) << (grp * G);
    const unsigned long long below = (1ull << lane) - 1;
    int cnt = 0;
    for (int base = 0; base < total; base += G * U) {
        int at[U], jj[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int q = base + k * G + gl;
            at[k] = q < total ? (anyWrap ? stencil_at(pre, shift, q) : stencil_at9(pre, shift, q)) : t;
        }
#pragma unroll
        for (int k = 0; k < U; k++) jj[k] = (base + k * G + gl < total) ? cellContents[at[k]] : -1;
        float4 qj[U];
#pragma unroll
        for (int k = 0; k < U; k++) qj[k] = pose.at(jj[k] > i ? at[k] : t, jj[k] > i ? jj[k] : i);
#pragma unroll
        for (int k = 0; k < U; k++) {
            const bool hit = jj[k] > i && bp_hit(qi, qj[k], g);
            const unsigned long long m = __ballot(hit) & groupBits;
            if (hit) {
                const int slot = cnt + __popcll(m & below);
                if (slot < maxCand) slice[slot] = jj[k];
            }
            cnt += __popcll(m);
        }
    }
    if (gl == 0) candCount[i] = cnt;
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
    PlanarPose pose{center, boundingRadius, bodyCell, cap};
    auto kern = k_broad_phase<PlanarPose, SingleGeo, BP_GROUP, BP_ROUND>;
    hipLaunchKernelGGL(kern, dim3(blocks_for((long long)cap * BP_GROUP, BP_THREADS)), dim3(BP_THREADS), 0,
                       st ? st : c->stream, pose, SingleGeo{g}, cellOffsets, cellContents, cap, S, candPartner,
                       candCount, maxCand);
    return check_launch(c, "k_broad_phase");
}

// LDS-staged form of the same query (SpatialGrid.java:340-406), for dense grids.  A block takes
// BP_THREADS / G consecutive bodies of the sorted contents, whose cells run from cFirst to cLast.
// For each of the 9 (dz, dy) rows of the stencil, every body's row lies in the cells cFirst + dz nXY
// + dy nX - 1 .. cLast + dz nXY + dy nX + 1, one contiguous range of the sorted contents: the block
// copies those 9 ranges (candidate index and packed pose) into LDS once, and each body walks its own
// stencil -- the same 18 ranges in the same order as k_broad_phase -- reading LDS for every range that
// lies inside its row's staged range and global memory for the rest (a periodic wrap or a grid edge
// of another replica).  Hits are ordered exactly as in k_broad_phase.  A block whose ranges exceed
// BPS_CAP entries, or whose bodies span two replicas, walks global memory throughout.
constexpr int BPS_CAP = 1536;  // staged entries per block (20 B each)

template <typename Geo, int G>
__global__ void __launch_bounds__(BP_THREADS) k_broad_phase_staged(SortedPose pose, Geo geo,
                                                                   const int *__restrict__ cellOffsets,
                                                                   const int *__restrict__ cellContents, int cap,
                                                                   int nSorted, int *__restrict__ candPartner,
                                                                   int *__restrict__ candCount, int maxCand) {
    static_assert(64 % G == 0, "groups tile a wave");
    constexpr int U = BP_ROUND;
    constexpr int TB = BP_THREADS / G;  // bodies per block
    __shared__ int sJ[BPS_CAP];
    __shared__ float4 sP[BPS_CAP];
    __shared__ int slabG[9], slabN[9], slabL[9], staged;
    const int tid = threadIdx.x, lane = tid & 63, gl = lane % G;
    const int t0 = xcd_block(blockIdx.x, gridDim.x) * TB;
    const int t = t0 + tid / G;
    const int tEnd = min(nSorted, min(cap, t0 + TB));  // bodies [t0, tEnd) have cells
    // the 9 staged ranges, from the first and the last body's cells (thread k < 9: row (dz, dy))
    if (tid < 9) {
        int g0 = 0, n = 0;
        if (t0 < tEnd) {
            const int iF = cellContents[t0], iL = cellContents[tEnd - 1];
            const BodyGeo bF = geo.of(iF), bL = geo.of(iL);
            if (bF.cellBase == bL.cellBase) {  // one replica
                const int nX = bF.g.nX, nXY = bF.g.nX * bF.g.nY, cells = nXY * bF.g.nZ;
                const int off = (tid / 3 - 1) * nXY + (tid % 3 - 1) * nX;
                int a = pose.cell(t0, iF) + off - 1, b = pose.cell(tEnd - 1, iL) + off + 1;
                a = max(a, bF.cellBase);
                b = min(b, bF.cellBase + cells - 1);
                if (a <= b) {
                    g0 = cellOffsets[a];
                    n = cellOffsets[b + 1] - g0;
                }
            } else {
                n = -1;  // bodies of two replicas: no staging
            }
        }
        slabG[tid] = g0;
        slabN[tid] = n;
    }
    __syncthreads();
    if (tid == 0) {
        int tot = 0, ok = 1;
        for (int k = 0; k < 9; k++) {
            ok = ok && slabN[k] >= 0;
            slabL[k] = tot;
            tot += max(slabN[k], 0);
        }
        staged = ok && tot <= BPS_CAP;
    }
    __syncthreads();
    if (staged) {
        for (int k = 0; k < 9; k++) {
            const int g0 = slabG[k], n = slabN[k], l0 = slabL[k];
            for (int e = tid; e < n; e += BP_THREADS) {
                sJ[l0 + e] = cellContents[g0 + e];
                sP[l0 + e] = pose.p[g0 + e];
            }
        }
    }
    __syncthreads();
    if (t >= cap) return;  // whole groups leave together; no barrier follows
    if (t >= nSorted) {
        if (gl == 0) candCount[t] = 0;
        return;
    }
    const int i = cellContents[t];
    const int c = pose.cell(t, i);
    if (c < 0) {
        if (gl == 0) candCount[i] = 0;
        return;
    }
    const float4 qi = pose.own(t, i);
    const BodyGeo bg = geo.of(i);
    const GridGeom &g = bg.g;
    int pre[18], shift[18];
    bool wraps = false;
    const int total = stencil_map(c - bg.cellBase, g, cellOffsets + bg.cellBase, pre, shift, &wraps);
    // ranges inside their row's staged range read LDS: their shift then maps a walk position to an
    // LDS index (bit r of inLds)
    unsigned inLds = 0;
    if (staged) {
#pragma unroll
        for (int r = 0; r < 18; r++) {
            const int k = r >> 1;
            const int lo = pre[r] + shift[r];
            const int hi = lo + ((r < 17 ? pre[r + 1] : total) - pre[r]);
            if (lo >= slabG[k] && hi <= slabG[k] + slabN[k]) {
                shift[r] += slabL[k] - slabG[k];
                inLds |= 1u << r;
            }
        }
    }
    int *slice = candPartner + (size_t)i * (size_t)maxCand;
    const unsigned long long groupBits = G == 64 ? ~0ull : ((1ull << G) - 1) << ((lane / G) * G);
    const unsigned long long below = (1ull << lane) - 1;
    int cnt = 0, base = 0;
    // No row of any body in the wave wraps (cells off the periodic x faces): the 9 odd ranges are empty
    // and own no position, so the even ranges alone map a position.
    const bool anyWrap = __ballot(wraps) != 0ull;
    // One round: positions base + k G + gl (k < UU) of every body, in ascending position order.
    auto round = [&](auto width) {
        constexpr int UU = decltype(width)::value;
        int at[UU], jj[UU];
        bool fromLds[UU];
#pragma unroll
        for (int k = 0; k < UU; k++) {
            const int q = base + k * G + gl;
            int sh = shift[0], rr = 0;
            if (anyWrap) {
#pragma unroll
                for (int r = 1; r < 18; r++)
                    if (q >= pre[r]) { sh = shift[r]; rr = r; }
            } else {
#pragma unroll
                for (int r = 2; r < 18; r += 2)
                    if (q >= pre[r]) { sh = shift[r]; rr = r; }
            }
            fromLds[k] = q < total && ((inLds >> rr) & 1u);
            at[k] = q < total ? q + sh : t;
        }
        float4 qj[UU];
#pragma unroll
        for (int k = 0; k < UU; k++) {
            const bool valid = base + k * G + gl < total;
            jj[k] = !valid ? -1 : (fromLds[k] ? sJ[at[k]] : cellContents[at[k]]);
        }
#pragma unroll
        for (int k = 0; k < UU; k++) {
            const bool use = jj[k] > i;
            qj[k] = !use ? qi : (fromLds[k] ? sP[at[k]] : pose.at(at[k], jj[k]));
        }
#pragma unroll
        for (int k = 0; k < UU; k++) {
            const bool hit = jj[k] > i && bp_hit(qi, qj[k], g);
            const unsigned long long m = __ballot(hit) & groupBits;
            if (hit) {
                const int slot = cnt + __popcll(m & below);
                if (slot < maxCand) slice[slot] = jj[k];
            }
            cnt += __popcll(m);
        }
        base += G * UU;
    };
    // full rounds while some body of the wave has more than G positions left, then single ones
    while (base < total) {
        if (__ballot(base + G < total) != 0ull) round(std::integral_constant<int, U>{});
        else round(std::integral_constant<int, 1>{});
    }
    if (gl == 0) candCount[i] = cnt;
}

template <typename Geo, int G>
void launch_staged(SortedPose pose, Geo geo, const int *cellOffsets, const int *cellContents, int cap, int S,
                   int *candPartner, int *candCount, int maxCand, hipStream_t st) {
    auto kern = k_broad_phase_staged<Geo, G>;
    hipLaunchKernelGGL(kern, dim3(blocks_for((long long)cap * G, BP_THREADS)), dim3(BP_THREADS), 0, st, pose, geo,
                       cellOffsets, cellContents, cap, S, candPartner, candCount, maxCand);
}

// Row-split form of the same query (SpatialGrid.java:340-406), for sparse grids.  The G lanes of a
// body take whole stencil rows -- lane l the rows l, l + G, ... of the 9 (dz, dy) rows, each the three
// x-cells of one row: one contents range, two at a periodic wrap -- so a lane loads only its own rows'
// bounds and walks them in order, with no position-to-range map.  Pass 1 counts every row's hits; the
// nine counts are exchanged inside the group, which places each row in the slice (rows in (z, y)
// order, contents positions ascending within a row: the broad phase's order); pass 2 walks again only
// the rows that hold hits and writes them.  candCount is the full count, as in k_broad_phase.
template <typename Geo, int G>
__global__ void __launch_bounds__(BP_THREADS) k_broad_phase_rows(SortedPose pose, Geo geo,
                                                                 const int *__restrict__ cellOffsets,
                                                                 const int *__restrict__ cellContents, int cap,
                                                                 int nSorted, int *__restrict__ candPartner,
                                                                 int *__restrict__ candCount, int maxCand) {
    static_assert(G == 4 || G == 8, "4 or 8 lanes per body");
    constexpr int TB = BP_THREADS / G;      // bodies per block
    constexpr int NR = (9 + G - 1) / G;     // rows per lane
    const int tid = threadIdx.x, gl = (tid & 63) % G;
    const int t = xcd_block(blockIdx.x, gridDim.x) * TB + tid / G;
    if (t >= cap) return;  // whole groups leave together: the shuffles below stay inside live groups
    if (t >= nSorted) {
        if (gl == 0) candCount[t] = 0;
        return;
    }
    const int i = cellContents[t];
    const int c = pose.cell(t, i);
    if (c < 0) {
        if (gl == 0) candCount[i] = 0;
        return;
    }
    const float4 qi = pose.own(t, i);
    const BodyGeo bg = geo.of(i);
    const GridGeom &g = bg.g;
    const int *off = cellOffsets + bg.cellBase;
    const int nX = g.nX, nY = g.nY, nZ = g.nZ, nXY = nX * nY;
    const int cl = c - bg.cellBase;
    const int cz = cl / nXY;
    const int rem = cl - cz * nXY;
    const int cy = rem / nX;
    const int cx = rem - cy * nX;
    int xa0, xa1, xb0 = 0, xb1 = -1;  // the row's x-cells: [xa0, xa1] then (periodic wrap) [xb0, xb1]
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
    int lo0[NR], hi0[NR], lo1[NR], hi1[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int r = gl + k * G;
        int zz = cz + r / 3 - 1, yy = cy + r % 3 - 1;
        bool ok = r < 9;
        if (g.periodic) {
            zz = zz < 0 ? zz + nZ : (zz >= nZ ? zz - nZ : zz);
            yy = yy < 0 ? yy + nY : (yy >= nY ? yy - nY : yy);
        } else {
            ok = ok && zz >= 0 && zz < nZ && yy >= 0 && yy < nY;
        }
        const int base = zz * nXY + yy * nX;
        lo0[k] = ok ? off[base + xa0] : 0;
        hi0[k] = ok ? off[base + xa1 + 1] : 0;
        lo1[k] = ok && twoSeg ? off[base + xb0] : 0;
        hi1[k] = ok && twoSeg ? off[base + xb1 + 1] : 0;
    }
    // pass 1: hits per row
    int cnt[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        int h = 0;
        for (int q = lo0[k]; q < hi0[k]; q++) {
            const int j = cellContents[q];
            const float4 qj = pose.at(q, j);
            h += (j > i && bp_hit(qi, qj, g)) ? 1 : 0;
        }
        for (int q = lo1[k]; q < hi1[k]; q++) {
            const int j = cellContents[q];
            const float4 qj = pose.at(q, j);
            h += (j > i && bp_hit(qi, qj, g)) ? 1 : 0;
        }
        cnt[k] = h;
    }
    // each row's place in the slice: row r is counted by lane r % G of the group, in its slot r / G
    int start[NR];
    int total = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        const int v = __shfl(cnt[r / G], r % G, G);
        if (r % G == gl) start[r / G] = total;
        total += v;
    }
#pragma unroll
    for (int r = 9; r < NR * G; r++)
        if (r % G == gl) start[r / G] = total;  // rows past the ninth hold no hit
    // pass 2: the rows with hits write them in order
    int *slice = candPartner + (size_t)i * (size_t)maxCand;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        if (cnt[k] == 0 || start[k] >= maxCand) continue;
        int slot = start[k];
        for (int q = lo0[k]; q < hi0[k] && slot < maxCand; q++) {
            const int j = cellContents[q];
            if (j > i && bp_hit(qi, pose.at(q, j), g)) slice[slot++] = j;
        }
        for (int q = lo1[k]; q < hi1[k] && slot < maxCand; q++) {
            const int j = cellContents[q];
            if (j > i && bp_hit(qi, pose.at(q, j), g)) slice[slot++] = j;
        }
    }
    if (gl == 0) candCount[i] = total;
}

template <typename Geo, int G>
void launch_rows(SortedPose pose, Geo geo, const int *cellOffsets, const int *cellContents, int cap, int S,
                 int *candPartner, int *candCount, int maxCand, hipStream_t st) {
    auto kern = k_broad_phase_rows<Geo, G>;
    hipLaunchKernelGGL(kern, dim3(blocks_for((long long)cap * G, BP_THREADS)), dim3(BP_THREADS), 0, st, pose, geo,
                       cellOffsets, cellContents, cap, S, candPartner, candCount, maxCand);
}

template <typename Geo, int G>
void launch_sorted(sbx_ctx *c, SortedPose pose, Geo geo, const int *cellOffsets, const int *cellContents, int cap,
                   int S, int *candPartner, int *candCount, int maxCand, hipStream_t st) {
    auto kern = k_broad_phase<SortedPose, Geo, G, BP_ROUND>;
    hipLaunchKernelGGL(kern, dim3(blocks_for((long long)cap * G, BP_THREADS)), dim3(BP_THREADS), 0, st, pose, geo,
                       cellOffsets, cellContents, cap, S, candPartner, candCount, maxCand);
}

int launch_broad_phase_sorted(sbx_ctx *c, const float4 *sortedPose, const int *sortedCell, const int *cellOffsets,
                              const int *cellContents, const GridGeom &g, const RepInfo *reps, const int *rodRep,
                              int group, int staged, int cap, int S, int *candPartner, int *candCount, int maxCand,
                              hipStream_t st) {
    SBX_REQUIRE(c, group == 2 || group == 4 || group == 8, "broad phase: %d lanes per body (2, 4 or 8)", group);
    SortedPose pose{sortedPose, sortedCell};
    hipStream_t q = st ? st : c->stream;
    if (staged == 2) {
        SBX_REQUIRE(c, group == 4 || group == 8, "row-split broad phase: 4 or 8 lanes per body");
        if (reps) {
            const BatchGeo geo{reps, rodRep};
            if (group == 8) launch_rows<BatchGeo, 8>(pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                     candCount, maxCand, q);
            else launch_rows<BatchGeo, 4>(pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                          maxCand, q);
        } else {
            const SingleGeo geo{g};
            if (group == 8) launch_rows<SingleGeo, 8>(pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                      candCount, maxCand, q);
            else launch_rows<SingleGeo, 4>(pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                           maxCand, q);
        }
        return check_launch(c, "k_broad_phase_rows");
    }
    if (staged) {
        SBX_REQUIRE(c, group == 4 || group == 8, "staged broad phase: 4 or 8 lanes per body");
        if (reps) {
            const BatchGeo geo{reps, rodRep};
            if (group == 8) launch_staged<BatchGeo, 8>(pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                       candCount, maxCand, q);
            else launch_staged<BatchGeo, 4>(pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                            maxCand, q);
        } else {
            const SingleGeo geo{g};
            if (group == 8) launch_staged<SingleGeo, 8>(pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                        candCount, maxCand, q);
            else launch_staged<SingleGeo, 4>(pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                             maxCand, q);
        }
        return check_launch(c, "k_broad_phase_staged");
    }
    if (reps) {
        const BatchGeo geo{reps, rodRep};
        if (group == 8) launch_sorted<BatchGeo, 8>(c, pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                   candCount, maxCand, q);
        else if (group == 4) launch_sorted<BatchGeo, 4>(c, pose, geo, cellOffsets, cellContents, cap, S,
                                                        candPartner, candCount, maxCand, q);
        else launch_sorted<BatchGeo, 2>(c, pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                        maxCand, q);
    } else {
        const SingleGeo geo{g};
        if (group == 8) launch_sorted<SingleGeo, 8>(c, pose, geo, cellOffsets, cellContents, cap, S, candPartner,
                                                    candCount, maxCand, q);
        else if (group == 4) launch_sorted<SingleGeo, 4>(c, pose, geo, cellOffsets, cellContents, cap, S,
                                                         candPartner, candCount, maxCand, q);
        else launch_sorted<SingleGeo, 2>(c, pose, geo, cellOffsets, cellContents, cap, S, candPartner, candCount,
                                         maxCand, q);
    }
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
