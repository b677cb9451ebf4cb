// sbx_form.hip -- crosslinker formation on the GPU (V2OneXHarness.formationCpu, V2OneXHarness.java:444-464).
//
// One pass at the formation cadence:
//   filID            FilIDSystem.init + rounds(n) pointer-doubling jumps (FilIDSystem.java:31-60)
//   formation grid   segment centres binned and counting-sorted (FormationGrid.buildCpu, FormationGrid.java:87-97)
//   candidates       gridFormCount -> exclusive scan -> gridFormEmit (CrosslinkerSystem.java:461-604): one
//                    lane per segment walks its 27 cells; the scan gives every segment its own contiguous
//                    region, so the list is in (i, j ascending) order with no atomics, identical to the
//                    O(N^2) filFilCandidates
//   formGates        per candidate (:607-676), grid-stride over the request capacity
//   formAdmitReduce  per-segment minimum gate-passing candidate (:680-691): atomicMin, order-free
//   formAdmit        (:713-738); the spacing test walks the two segments' link CSRs (keyed by linkFilA and
//                    linkFilB) instead of every link, with the same outcome
//   allocate         free flags -> scan -> scatter (the free list in slot order) -> scan of the accept flags
//                    (dense ranks) -> allocate + placeOrient (:298-357, :742-760)
// Periodic boxes (no counterpart in the reference) wrap the grid, take centre distances as the minimum
// image in double and shift segment b by its centre's image offset inside the gates.
#include "sbx_device.h"
#include "sbx_form.h"

namespace sbx {
namespace {

constexpr uint32_t XLJ1 = 0x584C4A31u, XLJ2 = 0x584C4A32u, XLFP = 0x584C4650u;

unsigned stride_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

__global__ void __launch_bounds__(256) k_fill_value(int *__restrict__ dst, int v, int n) {
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) dst[k] = v;
}

// the formation grid's cell of each segment centre (SpatialGrid.bodyCell on the formation grid)
__global__ void __launch_bounds__(256) k_cell_of(const float *__restrict__ coord, int n, GridGeom g,
                                                 int *__restrict__ cell) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) cell[i] = cell_of(coord[i], coord[n + i], coord[2 * n + i], g);
}

__global__ void __launch_bounds__(256) k_filid_init(const int *__restrict__ filState, const int *__restrict__ end2,
                                                    int *__restrict__ ptr, int n) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    if (filState && filState[s] < 0) {
        ptr[s] = -s - 2;
        return;
    }
    int nx = end2[s];
    ptr[s] = (nx >= 0 && nx != s && (!filState || filState[nx] >= 0)) ? nx : s;
}

__global__ void __launch_bounds__(256) k_filid_jump(const int *__restrict__ in, int *__restrict__ out,
                                                    const int *__restrict__ filState, int n) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    if (filState && filState[s] < 0) {
        out[s] = in[s];
        return;
    }
    out[s] = in[in[s]];
}

__device__ __forceinline__ double min_image_d(double d, double L) { return d - L * rint(d / L); }

// filFilCandidates' centre distance and coarse capsule bound (CrosslinkerSystem.java:413-423)
__device__ __forceinline__ bool form_pair(int n, double cix, double ciy, double ciz, double hi, int j,
                                          const float *__restrict__ coord, const float *__restrict__ segLength,
                                          double grab, const GridGeom &g) {
    double dx = cix - (double)coord[j], dy = ciy - (double)coord[n + j], dz = ciz - (double)coord[2 * n + j];
    if (g.periodic) {
        dx = min_image_d(dx, (double)g.Lx);
        dy = min_image_d(dy, (double)g.Ly);
        dz = min_image_d(dz, (double)g.Lz);
    }
    double cd = sqrt(dx * dx + dy * dy + dz * dz);
    double bound = hi + 0.5 * (double)segLength[j] + grab;
    return !(cd > bound);
}

// The 27 cells of segment i's cell (clamped in a box, wrapped when periodic) as the broad phase's walk
// over the cell-sorted contents: all 36 range bounds are loaded at once and the partners are read in
// batches, so a lane does not wait on one cell before the next.  Calls visit(j) for every segment j > i
// of another filament within the capsule bound; the visiting order does not matter to either caller.
constexpr int FORM_BATCH = 4;
template <typename F>
__device__ __forceinline__ void form_walk(int i, int n, const float *__restrict__ coord,
                                          const float *__restrict__ segLength, const int *__restrict__ filID,
                                          const int *__restrict__ off, const int *__restrict__ cont,
                                          const GridGeom &g, double grab, F visit) {
    const float fx = coord[i], fy = coord[n + i], fz = coord[2 * n + i];
    int pre[18], shift[18];
    const int total = stencil_map(cell_of(fx, fy, fz, g), g, off, pre, shift);
    const double cix = fx, ciy = fy, ciz = fz, hi = 0.5 * (double)segLength[i];
    const int fi = filID[i];
    for (int q = 0; q < total; q += FORM_BATCH) {
        int jj[FORM_BATCH];
#pragma unroll
        for (int u = 0; u < FORM_BATCH; u++) jj[u] = (q + u < total) ? cont[stencil_at(pre, shift, q + u)] : -1;
        int fj[FORM_BATCH];
#pragma unroll
        for (int u = 0; u < FORM_BATCH; u++) fj[u] = jj[u] > i ? filID[jj[u]] : fi;
#pragma unroll
        for (int u = 0; u < FORM_BATCH; u++) {
            if (jj[u] <= i || fj[u] == fi) continue;
            if (form_pair(n, cix, ciy, ciz, hi, jj[u], coord, segLength, grab, g)) visit(jj[u]);
        }
    }
}

// gridFormCount (CrosslinkerSystem.java:461-513)
__global__ void __launch_bounds__(256) k_form_count(const float *__restrict__ coord, const float *__restrict__ segLength,
                                                    const int *__restrict__ filID, const int *__restrict__ off,
                                                    const int *__restrict__ cont, GridGeom g, double grab, int n,
                                                    int *__restrict__ candCount) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int cnt = 0;
    form_walk(i, n, coord, segLength, filID, off, cont, g, grab, [&](int) { cnt++; });
    candCount[i] = cnt < FORM_MAXC ? cnt : FORM_MAXC;
}

// gridFormScan's total (:524-526) and gridFormEmit (:535-604): segment i writes its partners ascending into
// its own region.  The 256 segments of a block own one contiguous span of the list: when the span fits in
// LDS and no region is capped (FORM_MAXC) or clamped (reqCap), each segment appends its partners to its
// part of the span in LDS, then the block places every entry at its rank within its segment's part (the
// partners of a segment are distinct), spreading long regions over all threads.  Otherwise each segment
// inserts into its region in place, keeping the smallest `cap`, as the reference does.
constexpr int FORM_EMIT_LDS = 8192;
__global__ void __launch_bounds__(256) k_form_emit(const float *__restrict__ coord, const float *__restrict__ segLength,
                                                   const int *__restrict__ filID, const int *__restrict__ off,
                                                   const int *__restrict__ cont, GridGeom g, double grab, int n,
                                                   const int *__restrict__ candBase, const int *__restrict__ candCount,
                                                   int reqCap, int *__restrict__ reqFilA, int *__restrict__ reqFilB,
                                                   int *__restrict__ total) {
    __shared__ int buf[FORM_EMIT_LDS];
    __shared__ unsigned char own[FORM_EMIT_LDS];
    __shared__ int rLo[256], rLen[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        int acc = candBase[n];
        *total = acc < reqCap ? acc : reqCap;
    }
    const int i0 = blockIdx.x * 256, i1 = min(i0 + 256, n);
    const int spanLo = candBase[i0], spanHi = candBase[i1];
    int base = 0, cnt = 0, cap = 0;
    if (i < n) {
        base = candBase[i];
        cnt = candCount[i];
        cap = cnt;
        if (base >= reqCap) cap = 0;
        else if (base + cap > reqCap) cap = reqCap - base;
    }
    const bool whole = cap == cnt && cnt < FORM_MAXC;
    const bool fast = __syncthreads_and(whole) && spanHi - spanLo <= FORM_EMIT_LDS;
    if (fast) {
        const int lo = base - spanLo;
        rLo[threadIdx.x] = lo;
        rLen[threadIdx.x] = cap;
        if (cap > 0) {
            int k = lo;
            form_walk(i, n, coord, segLength, filID, off, cont, g, grab, [&](int j) {
                buf[k] = j;
                own[k] = (unsigned char)threadIdx.x;
                k++;
            });
        }
        __syncthreads();
        for (int p = threadIdx.x; p < spanHi - spanLo; p += 256) {
            const int o = own[p], rl = rLo[o], len = rLen[o], x = buf[p];
            int rank = 0;
            for (int q = rl; q < rl + len; q++) rank += buf[q] < x ? 1 : 0;
            reqFilB[spanLo + rl + rank] = x;
            reqFilA[spanLo + rl + rank] = i0 + o;
        }
        return;
    }
    if (cap <= 0) return;
    int cntIns = 0;
    int *r = reqFilB;
    form_walk(i, n, coord, segLength, filID, off, cont, g, grab, [&](int j) {
        if (cntIns < cap) {
            int p = base + cntIns;
            while (p > base && r[p - 1] > j) {
                r[p] = r[p - 1];
                p--;
            }
            r[p] = j;
            cntIns++;
        } else if (j < r[base + cap - 1]) {
            int p = base + cap - 1;
            while (p > base && r[p - 1] > j) {
                r[p] = r[p - 1];
                p--;
            }
            r[p] = j;
        }
    });
    for (int q = 0; q < cap; q++) reqFilA[base + q] = i;
}

// the unused request tail keyed -1 (gridFormScan :527)
__global__ void __launch_bounds__(256) k_form_tail(int *__restrict__ reqFilA, int *__restrict__ reqFilB, int reqCap,
                                                   const int *__restrict__ total) {
    const int t = *total;
    for (int k = t + blockIdx.x * 256 + threadIdx.x; k < reqCap; k += gridDim.x * 256) {
        reqFilA[k] = -1;
        reqFilB[k] = -1;
    }
}

// formGates (CrosslinkerSystem.java:607-676)
__global__ void __launch_bounds__(256) k_form_gates(const float *__restrict__ coord, const float *__restrict__ uVec,
                                                    const float *__restrict__ end1, const float *__restrict__ end2,
                                                    const float *__restrict__ segLength,
                                                    const int *__restrict__ reqFilA, const int *__restrict__ reqFilB,
                                                    float *__restrict__ reqLoc1, float *__restrict__ reqLoc2,
                                                    int *__restrict__ reqOrient, int *__restrict__ gatePass,
                                                    FormParams fp, int step, int seed, int rng, uint32_t stream,
                                                    GridGeom g, int n, int reqCap) {
    const double maxAngle = fp.maxAngle, grabSq = fp.grabSq, pForm = fp.pForm, jit = fp.jit;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < reqCap; c += gridDim.x * 256) {
        gatePass[c] = 0;
        reqOrient[c] = 0;
        const int a = reqFilA[c], b = reqFilB[c];
        if (a < 0 || b < 0) continue;
        const double uax = uVec[a], uay = uVec[n + a], uaz = uVec[2 * n + a];
        const double ubx = uVec[b], uby = uVec[n + b], ubz = uVec[2 * n + b];
        const double dot = uax * ubx + uay * uby + uaz * ubz;
        const double angT = fast_acos(dot), angTR = fast_acos(-dot);
        bool alignOk;
        if (fp.mode == 1) alignOk = angT <= maxAngle;
        else if (fp.mode == -1) alignOk = angTR <= maxAngle;
        else alignOk = (angT <= maxAngle) || (angTR <= maxAngle);
        if (!alignOk) continue;
        reqOrient[c] = dot >= 0.0 ? 1 : 0;
        const double e1ax = end1[a], e1ay = end1[n + a], e1az = end1[2 * n + a];
        const double e2ax = end2[a], e2ay = end2[n + a], e2az = end2[2 * n + a];
        double e1bx = end1[b], e1by = end1[n + b], e1bz = end1[2 * n + b];
        double e2bx = end2[b], e2by = end2[n + b], e2bz = end2[2 * n + b];
        if (g.periodic) {
            const double rx = (double)coord[b] - (double)coord[a];
            const double ry = (double)coord[n + b] - (double)coord[n + a];
            const double rz = (double)coord[2 * n + b] - (double)coord[2 * n + a];
            const double ox = min_image_d(rx, (double)g.Lx) - rx;
            const double oy = min_image_d(ry, (double)g.Ly) - ry;
            const double oz = min_image_d(rz, (double)g.Lz) - rz;
            e1bx += ox; e1by += oy; e1bz += oz;
            e2bx += ox; e2by += oy; e2bz += oz;
        }
        const double r1x = e2ax - e1ax, r1y = e2ay - e1ay, r1z = e2az - e1az;
        const double r2x = e2bx - e1bx, r2y = e2by - e1by, r2z = e2bz - e1bz;
        const double r3x = e1bx - e1ax, r3y = e1by - e1ay, r3z = e1bz - e1az;
        const double r4x = r1y * r2z - r1z * r2y, r4y = r1z * r2x - r1x * r2z, r4z = r1x * r2y - r1y * r2x;
        const double smallNum = 1e-20;
        if (r4x < smallNum && r4y < smallNum && r4z < smallNum) continue;
        const double denom = r4x * r4x + r4y * r4y + r4z * r4z;
        const double c32x = r3y * r2z - r3z * r2y, c32y = r3z * r2x - r3x * r2z, c32z = r3x * r2y - r3y * r2x;
        const double alpha = (r4x * c32x + r4y * c32y + r4z * c32z) / denom;
        if (alpha < 0.0 || alpha > 1.0) continue;
        const double c31x = r3y * r1z - r3z * r1y, c31y = r3z * r1x - r3x * r1z, c31z = r3x * r1y - r3y * r1x;
        const double beta = (r4x * c31x + r4y * c31y + r4z * c31z) / denom;
        if (beta < 0.0 || beta > 1.0) continue;
        const double p1x = e1ax + alpha * r1x, p1y = e1ay + alpha * r1y, p1z = e1az + alpha * r1z;
        const double p2x = e1bx + beta * r2x, p2y = e1by + beta * r2y, p2z = e1bz + beta * r2z;
        const double ddx = p1x - p2x, ddy = p1y - p2y, ddz = p1z - p2z;
        const double conDistSq = ddx * ddx + ddy * ddy + ddz * ddz;
        if (conDistSq >= grabSq) continue;
        const double da = sqrt((p1x - e1ax) * (p1x - e1ax) + (p1y - e1ay) * (p1y - e1ay) + (p1z - e1az) * (p1z - e1az));
        const double db = sqrt((p2x - e1bx) * (p2x - e1bx) + (p2y - e1by) * (p2y - e1by) + (p2z - e1bz) * (p2z - e1bz));
        const float uj1 = uniform01(event_bits(rng, (uint32_t)c, step, seed, stream, XLJ1));
        const float uj2 = uniform01(event_bits(rng, (uint32_t)c, step, seed, stream, XLJ2));
        double loc1 = da + (2.0 * uj1 - 1.0) * jit;
        double loc2 = db + (2.0 * uj2 - 1.0) * jit;
        const double lenA = segLength[a], lenB = segLength[b];
        if (loc1 > lenA) loc1 = lenA;
        if (loc1 < 0.0) loc1 = 0.0;
        if (loc2 > lenB) loc2 = lenB;
        if (loc2 < 0.0) loc2 = 0.0;
        reqLoc1[c] = (float)loc1;
        reqLoc2[c] = (float)loc2;
        const float up = uniform01(event_bits(rng, (uint32_t)c, step, seed, stream, XLFP));
        if (up < pForm) gatePass[c] = 1;
    }
}

// formAdmitReduce (:680-691): the minimum over candidates is independent of the atomic order
__global__ void __launch_bounds__(256) k_form_admit_reduce(const int *__restrict__ reqFilA,
                                                           const int *__restrict__ reqFilB,
                                                           const int *__restrict__ gatePass, int *__restrict__ minCand,
                                                           int reqCap) {
    for (int c = blockIdx.x * 256 + threadIdx.x; c < reqCap; c += gridDim.x * 256) {
        if (gatePass[c] == 0) continue;
        atomicMin(&minCand[reqFilA[c]], c);
        atomicMin(&minCand[reqFilB[c]], c);
    }
}

// a formed link at `loc` keeps minSep from every ACTIVE link end on the same segment (:726-733)
__device__ __forceinline__ bool spaced(int seg, double loc, const int *__restrict__ off, const int *__restrict__ idx,
                                       const float *__restrict__ linkLoc, const int *__restrict__ linkState,
                                       double minSep) {
    for (int p = off[seg]; p < off[seg + 1]; p++) {
        const int L = idx[p];
        if (linkState[L] < 0) continue;
        if (fabs(loc - (double)linkLoc[L]) < minSep) return false;
    }
    return true;
}

// formAdmit (:713-738)
__global__ void __launch_bounds__(256) k_form_admit(const int *__restrict__ reqFilA, const int *__restrict__ reqFilB,
                                                    const float *__restrict__ reqLoc1, const float *__restrict__ reqLoc2,
                                                    const int *__restrict__ gatePass, const int *__restrict__ minCand,
                                                    const int *__restrict__ activeLinkCount,
                                                    const int *__restrict__ linkState, const int *__restrict__ offA,
                                                    const int *__restrict__ idxA, const int *__restrict__ offB,
                                                    const int *__restrict__ idxB, const float *__restrict__ loc1,
                                                    const float *__restrict__ loc2, int *__restrict__ acceptFlag,
                                                    FormParams fp, int reqCap) {
    const double minSep = fp.minSep;
    const int maxLinks = (int)fp.maxLinks;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < reqCap; c += gridDim.x * 256) {
        acceptFlag[c] = 0;
        if (gatePass[c] == 0) continue;
        const int segA = reqFilA[c], segB = reqFilB[c];
        if (minCand[segA] != c || minCand[segB] != c) continue;
        if (activeLinkCount[segA] >= maxLinks || activeLinkCount[segB] >= maxLinks) continue;
        const double la = reqLoc1[c], lb = reqLoc2[c];
        const bool ok = spaced(segA, la, offA, idxA, loc1, linkState, minSep) &&
                        spaced(segA, la, offB, idxB, loc2, linkState, minSep) &&
                        spaced(segB, lb, offA, idxA, loc1, linkState, minSep) &&
                        spaced(segB, lb, offB, idxB, loc2, linkState, minSep);
        if (ok) acceptFlag[c] = 1;
    }
}

// freeFlags (:298-303)
__global__ void __launch_bounds__(256) k_free_flags(const int *__restrict__ linkState, int *__restrict__ freeCount,
                                                    int C) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C) freeCount[s] = linkState[s] < 0 ? 1 : 0;
}

// freeScatter (:308-316): each free slot writes its own prefix position (slot order)
__global__ void __launch_bounds__(256) k_free_scatter(const int *__restrict__ linkState,
                                                      const int *__restrict__ freeOffsets, int *__restrict__ freeList,
                                                      int C) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C && linkState[s] < 0) freeList[freeOffsets[s]] = s;
}

// allocate + placeOrient (:323-357, :742-760): request r of dense rank < nFree takes freeList[rank]
__global__ void __launch_bounds__(256) k_form_allocate(const int *__restrict__ reqFilA, const int *__restrict__ reqFilB,
                                                       const float *__restrict__ reqLoc1,
                                                       const float *__restrict__ reqLoc2,
                                                       const int *__restrict__ reqOrient,
                                                       const int *__restrict__ rankOffsets,
                                                       const int *__restrict__ freeList,
                                                       const int *__restrict__ freeOffsets, int *__restrict__ linkFilA,
                                                       int *__restrict__ linkFilB, float *__restrict__ loc1,
                                                       float *__restrict__ loc2, int *__restrict__ linkState,
                                                       float *__restrict__ strainHist, int *__restrict__ strainPlace,
                                                       int *__restrict__ linkOrientSame,
                                                       float *__restrict__ torqueMagHist,
                                                       int *__restrict__ torqueMagPlace, int *__restrict__ formed,
                                                       int K, int C) {
    const int nFree = freeOffsets[C];
    for (int r = blockIdx.x * 256 + threadIdx.x; r < K; r += gridDim.x * 256) {
        const int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        const int slot = freeList[rank];
        linkFilA[slot] = reqFilA[r];
        linkFilB[slot] = reqFilB[r];
        loc1[slot] = reqLoc1[r];
        loc2[slot] = reqLoc2[r];
        for (int j = 0; j < SBX_STRAIN_WIN; j++) strainHist[(size_t)slot * SBX_STRAIN_WIN + j] = 0.0f;
        strainPlace[slot] = 0;
        linkState[slot] = 0;  // CrosslinkerStore.LINK_ACTIVE
        linkOrientSame[slot] = reqOrient[r];
        torqueMagPlace[slot] = 0;
        for (int j = 0; j < SBX_TORQUE_WIN; j++) torqueMagHist[(size_t)slot * SBX_TORQUE_WIN + j] = 0.0f;
        if (formed) atomicAdd(formed, 1);
    }
}

}  // namespace

FormParams form_params(const float *f, int mode) {
    FormParams p;
    p.maxAngle = f[0];
    p.grabSq = f[1];
    p.minSep = f[2];
    p.maxLinks = f[3];
    p.pForm = f[4];
    p.jit = f[5];
    p.grab = f[6];
    p.mode = mode;
    return p;
}

int form_filid(sbx_ctx *c, const int *filState, const int *end2NbrSlot, int *filID, int *scratch, int n,
               hipStream_t q) {
    if (n <= 0) return SBX_OK;
    hipLaunchKernelGGL(k_filid_init, dim3(blocks_for(n, 256)), dim3(256), 0, q, filState, end2NbrSlot, filID, n);
    // FilIDSystem.rounds: ceil(log2 n) rounded up to even, so the last jump lands in filID
    int k = 0;
    if (n <= 1) k = 1;
    else for (unsigned v = (unsigned)(n - 1); v; v >>= 1) k++;
    if (k % 2) k++;
    int *a = filID, *b = scratch;
    for (int r = 0; r < k; r++) {
        hipLaunchKernelGGL(k_filid_jump, dim3(blocks_for(n, 256)), dim3(256), 0, q, a, b, filState, n);
        int *t = a;
        a = b;
        b = t;
    }
    return check_launch(c, "filid");
}

int form_candidates(sbx_ctx *c, SortScratch &s, const float *coord, const float *segLength, const int *filID,
                    const int *gridOff, const int *gridCont, const GridGeom &g, double grab,
                    int *candCount, int *candBase, int *reqFilA, int *reqFilB, int *total, int nSeg, int reqCap,
                    hipStream_t q) {
    if (nSeg > 0) {
        hipLaunchKernelGGL(k_form_count, dim3(blocks_for(nSeg, 256)), dim3(256), 0, q, coord, segLength, filID,
                           gridOff, gridCont, g, grab, nSeg, candCount);
    }
    int r = exclusive_scan(c, s, candCount, nSeg, candBase, q);
    if (r) return r;
    hipLaunchKernelGGL(k_form_emit, dim3(blocks_for(nSeg > 0 ? nSeg : 1, 256)), dim3(256), 0, q, coord, segLength,
                       filID, gridOff, gridCont, g, grab, nSeg, candBase, candCount, reqCap, reqFilA, reqFilB, total);
    hipLaunchKernelGGL(k_form_tail, dim3(stride_blocks(reqCap)), dim3(256), 0, q, reqFilA, reqFilB, reqCap, total);
    return check_launch(c, "form candidates");
}

int form_gates(sbx_ctx *c, const float *coord, const float *uVec, const float *end1, const float *end2,
               const float *segLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
               int *reqOrient, int *gatePass, const FormParams &fp, int step, int seed, int rng, uint32_t stream,
               const GridGeom &g, int nSeg, int reqCap, hipStream_t q) {
    hipLaunchKernelGGL(k_form_gates, dim3(stride_blocks(reqCap)), dim3(256), 0, q, coord, uVec, end1, end2, segLength,
                       reqFilA, reqFilB, reqLoc1, reqLoc2, reqOrient, gatePass, fp, step, seed, rng, stream, g, nSeg,
                       reqCap);
    return check_launch(c, "k_form_gates");
}

int form_admit_reduce(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const int *gatePass, int *minCand,
                      int reqCap, int nSeg, hipStream_t q) {
    hipLaunchKernelGGL(k_fill_value, dim3(stride_blocks(nSeg)), dim3(256), 0, q, minCand, reqCap, nSeg);
    hipLaunchKernelGGL(k_form_admit_reduce, dim3(stride_blocks(reqCap)), dim3(256), 0, q, reqFilA, reqFilB, gatePass,
                       minCand, reqCap);
    return check_launch(c, "k_form_admit_reduce");
}

int form_admit(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
               const int *gatePass, const int *minCand, const int *activeLinkCount, const int *linkState,
               const int *offA, const int *idxA, const int *offB, const int *idxB, const float *loc1,
               const float *loc2, int *acceptFlag, const FormParams &fp, int reqCap, hipStream_t q) {
    hipLaunchKernelGGL(k_form_admit, dim3(stride_blocks(reqCap)), dim3(256), 0, q, reqFilA, reqFilB, reqLoc1, reqLoc2,
                       gatePass, minCand, activeLinkCount, linkState, offA, idxA, offB, idxB, loc1, loc2, acceptFlag,
                       fp, reqCap);
    return check_launch(c, "k_form_admit");
}

int form_allocate(sbx_ctx *c, SortScratch &s, const int *reqFilA, const int *reqFilB, const float *reqLoc1,
                  const float *reqLoc2, const int *reqOrient, const int *acceptFlag, int *freeCount, int *freeOffsets,
                  int *freeList, int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1, float *loc2,
                  int *linkState, float *strainHist, int *strainPlace, int *linkOrientSame, float *torqueMagHist,
                  int *torqueMagPlace, int *formed, int K, int C, hipStream_t q) {
    if (C > 0) hipLaunchKernelGGL(k_free_flags, dim3(blocks_for(C, 256)), dim3(256), 0, q, linkState, freeCount, C);
    int r = exclusive_scan(c, s, freeCount, C, freeOffsets, q);
    if (r) return r;
    if (C > 0)
        hipLaunchKernelGGL(k_free_scatter, dim3(blocks_for(C, 256)), dim3(256), 0, q, linkState, freeOffsets, freeList,
                           C);
    if ((r = exclusive_scan(c, s, acceptFlag, K, rankOffsets, q))) return r;
    hipLaunchKernelGGL(k_form_allocate, dim3(stride_blocks(K)), dim3(256), 0, q, reqFilA, reqFilB, reqLoc1, reqLoc2,
                       reqOrient, rankOffsets, freeList, freeOffsets, linkFilA, linkFilB, loc1, loc2, linkState,
                       strainHist, strainPlace, linkOrientSame, torqueMagHist, torqueMagPlace, formed, K, C);
    return check_launch(c, "form allocate");
}

int form_pass(sbx_ctx *c, SortScratch &s, const FormPass &f, int step, bool filIDValid, hipStream_t q) {
    int r;
    const int n = f.nSeg;
    if (!filIDValid && (r = form_filid(c, f.filState, f.end2NbrSlot, f.filID, f.filIDScratch, n, q))) return r;
    const int cells = f.geom.nX * f.geom.nY * f.geom.nZ;
    if (n > 0) {
        hipLaunchKernelGGL(k_cell_of, dim3(blocks_for(n, 256)), dim3(256), 0, q, f.coord, n, f.geom, f.bodyCell);
        if ((r = check_launch(c, "formation grid cells"))) return r;
    }
    if ((r = counting_sort(c, s, f.bodyCell, n, cells, n, f.gridOff, f.gridCont, nullptr, q))) return r;
    if ((r = form_candidates(c, s, f.coord, f.segLength, f.filID, f.gridOff, f.gridCont, f.geom,
                             (double)f.fp.grab, f.candCount, f.candBase, f.reqFilA, f.reqFilB, f.stats, n, f.reqCap,
                             q)))
        return r;
    if ((r = form_gates(c, f.coord, f.uVec, f.end1, f.end2, f.segLength, f.reqFilA, f.reqFilB, f.reqLoc1, f.reqLoc2,
                        f.reqOrient, f.gatePass, f.fp, step, f.seed, f.rng, f.stream, f.geom, n, f.reqCap, q)))
        return r;
    if ((r = form_admit_reduce(c, f.reqFilA, f.reqFilB, f.gatePass, f.minCand, f.reqCap, n, q))) return r;
    if ((r = form_admit(c, f.reqFilA, f.reqFilB, f.reqLoc1, f.reqLoc2, f.gatePass, f.minCand, f.activeLinkCount,
                        f.linkState, f.offA, f.idxA, f.offB, f.idxB, f.loc1, f.loc2, f.acceptFlag, f.fp, f.reqCap, q)))
        return r;
    hipLaunchKernelGGL(k_fill_value, dim3(1), dim3(256), 0, q, f.stats + 1, 0, 1);
    return form_allocate(c, s, f.reqFilA, f.reqFilB, f.reqLoc1, f.reqLoc2, f.reqOrient, f.acceptFlag, f.freeCount,
                         f.freeOffsets, f.freeList, f.rankOffsets, f.linkFilA, f.linkFilB, f.loc1, f.loc2,
                         f.linkState, f.strainHist, f.strainPlace, f.linkOrientSame, f.torqueMagHist,
                         f.torqueMagPlace, f.stats + 1, f.reqCap, f.C, q);
}

}  // namespace sbx

using namespace sbx;

// ------------------------------------------------------------------ C ABI: the per-System entry points
extern "C" {

int sbx_filid(sbx_ctx *c, const int *filState, const int *end2NbrSlot, int *filID, int *filIDScratch, int n) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0, "filid: n < 0");
    SBX_REQUIRE(c, n == 0 || (end2NbrSlot && filID && filIDScratch), "filid: null array");
    return form_filid(c, filState, end2NbrSlot, filID, filIDScratch, n, c->stream);
}

int sbx_grid_form_candidates(sbx_ctx *c, const float *filCoord, const float *filSegLength, const int *filID,
                             const float *gridParams, const int *gridDims, const int *gridCellOffsets,
                             const int *gridCellContents, int *candCountSeg, int *candBaseSeg, int *reqFilA,
                             int *reqFilB, const float *formParams, int *formCounts, int nSeg, int reqCap,
                             int periodic, const float *boxLen) {
    SBX_CHECK_CTX(c);
    int r = check_grid(c, gridParams, gridDims, periodic);
    if (r) return r;
    SBX_REQUIRE(c, formParams && formCounts && nSeg >= 0 && reqCap > 0, "grid_form_candidates: bad arguments");
    SBX_REQUIRE(c, !periodic || (boxLen && gridDims[0] >= 3 && gridDims[1] >= 3 && gridDims[2] >= 3),
                "grid_form_candidates: periodic needs boxLen and >= 3 cells per axis");
    SBX_REQUIRE(c, nSeg == 0 || (filCoord && filSegLength && filID && gridCellOffsets && gridCellContents &&
                                 candCountSeg && candBaseSeg && reqFilA && reqFilB),
                "grid_form_candidates: null array");
    GridGeom g = grid_geom(gridParams, gridDims, periodic, boxLen);
    return form_candidates(c, c->api_sort, filCoord, filSegLength, filID, gridCellOffsets, gridCellContents,
                           g, (double)formParams[6], candCountSeg, candBaseSeg, reqFilA, reqFilB, formCounts, nSeg,
                           reqCap, c->stream);
}

int sbx_form_gates(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1, const float *filEnd2,
                   const float *filSegLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
                   int *reqOrient, int *gatePass, const float *formParams, const int *formCounts, int nSeg, int reqCap,
                   int rng, uint32_t stream, int periodic, const float *boxLen) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, formParams && formCounts && nSeg >= 0 && reqCap >= 0, "form_gates: bad arguments");
    SBX_REQUIRE(c, !periodic || boxLen, "form_gates: periodic needs boxLen");
    if (reqCap == 0) return SBX_OK;
    SBX_REQUIRE(c, filCoord && filUVec && filEnd1 && filEnd2 && filSegLength && reqFilA && reqFilB && reqLoc1 &&
                       reqLoc2 && reqOrient && gatePass,
                "form_gates: null array");
    GridGeom g{};
    g.periodic = periodic;
    if (periodic) { g.Lx = boxLen[0]; g.Ly = boxLen[1]; g.Lz = boxLen[2]; }
    return form_gates(c, filCoord, filUVec, filEnd1, filEnd2, filSegLength, reqFilA, reqFilB, reqLoc1, reqLoc2,
                      reqOrient, gatePass, form_params(formParams, formCounts[3]), formCounts[1], formCounts[2], rng,
                      stream, g, nSeg, reqCap, c->stream);
}

int sbx_form_admit_reduce(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const int *gatePass, int *minCand,
                          int reqCap, int nSeg) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, reqCap >= 0 && nSeg >= 0, "form_admit_reduce: bad sizes");
    SBX_REQUIRE(c, (reqCap == 0 || (reqFilA && reqFilB && gatePass)) && (nSeg == 0 || minCand),
                "form_admit_reduce: null array");
    return form_admit_reduce(c, reqFilA, reqFilB, gatePass, minCand, reqCap, nSeg, c->stream);
}

int sbx_form_admit(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
                   const int *gatePass, const int *minCand, const int *activeLinkCount, const int *linkState,
                   const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2, int *acceptFlag,
                   const float *formParams, int reqCap, int nSeg, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, formParams && reqCap >= 0 && nSeg >= 0 && C >= 0, "form_admit: bad arguments");
    if (reqCap == 0) return SBX_OK;
    SBX_REQUIRE(c, reqFilA && reqFilB && reqLoc1 && reqLoc2 && gatePass && minCand && activeLinkCount && acceptFlag,
                "form_admit: null array");
    SBX_REQUIRE(c, C == 0 || (linkState && linkFilA && linkFilB && loc1 && loc2), "form_admit: null link array");
    // the spacing test walks each segment's links: CSRs of the links by linkFilA and by linkFilB
    long long need = 2LL * (nSeg + 1) + 2LL * (C > 0 ? C : 1);
    if (need > c->form_csr_ints) {
        if (c->form_csr) hipFree(c->form_csr);
        c->form_csr = nullptr;
        c->form_csr_ints = 0;
        int r = check_hip(c, hipMalloc(&c->form_csr, sizeof(int) * (size_t)need), "hipMalloc form csr");
        if (r) return r;
        c->form_csr_ints = need;
    }
    int *offA = c->form_csr, *offB = offA + nSeg + 1, *idxA = offB + nSeg + 1, *idxB = idxA + (C > 0 ? C : 1);
    int r = counting_sort(c, c->api_sort, linkFilA, C, nSeg, C, offA, idxA);
    if (!r) r = counting_sort(c, c->api_sort, linkFilB, C, nSeg, C, offB, idxB);
    if (r) return r;
    return form_admit(c, reqFilA, reqFilB, reqLoc1, reqLoc2, gatePass, minCand, activeLinkCount, linkState, offA,
                      idxA, offB, idxB, loc1, loc2, acceptFlag, form_params(formParams, 0), reqCap, c->stream);
}

int sbx_form_allocate(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
                      const int *reqOrient, const int *acceptFlag, int *freeCount, int *freeOffsets, int *freeList,
                      int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1, float *loc2, int *linkState,
                      float *strainHist, int *strainPlace, int *linkOrientSame, float *torqueMagHist,
                      int *torqueMagPlace, int K, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, K >= 0 && C >= 0 && freeOffsets && rankOffsets, "form_allocate: bad arguments");
    SBX_REQUIRE(c, K == 0 || (reqFilA && reqFilB && reqLoc1 && reqLoc2 && reqOrient && acceptFlag),
                "form_allocate: null request array");
    SBX_REQUIRE(c, C == 0 || (freeCount && freeList && linkFilA && linkFilB && loc1 && loc2 && linkState &&
                              strainHist && strainPlace && linkOrientSame && torqueMagHist && torqueMagPlace),
                "form_allocate: null link array");
    return form_allocate(c, c->api_sort, reqFilA, reqFilB, reqLoc1, reqLoc2, reqOrient, acceptFlag, freeCount,
                         freeOffsets, freeList, rankOffsets, linkFilA, linkFilB, loc1, loc2, linkState, strainHist,
                         strainPlace, linkOrientSame, torqueMagHist, torqueMagPlace, nullptr, K, C, c->stream);
}

}  // extern "C"
