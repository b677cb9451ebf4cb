// sbx_turnover.hip -- actin turnover on the GPU (gfx950): aging, pointed-end depolymerisation and death,
// barbed-end growth and splits, cofilin severing, the filament-slot allocator and the actin pool.
//
// Every rule is a per-slot (or per-request) self-write of the reference's (AgingSystem.java,
// DepolySystem.java, GrowthSystem.java, SeveringSystem.java, FilamentBirthSystem.java), so one cadence
// runs as a handful of passes over the slots instead of the reference's twenty: a slot's later rules read
// only what its own earlier rules wrote, except the chain link break of a death, which every slot applies
// to itself by looking at its neighbours' death flags (the reference scatters it from the dying slot; on
// reciprocal chains, which every operation here keeps, the two are the same).  Event totals for the actin
// pool are atomic integer sums; the pool itself (a double, ActinPool.java) is advanced by one thread in the
// reference's order of takes and puts.  All integer decisions are bit-exact with the oracle; the pose and
// drag arithmetic keeps the reference's operation order.
#include <new>

#include "sbx_turnover.h"

namespace sbx {
namespace {

enum { TL_GREW = 0, TL_RET, TL_SEV, TL_DEATHS, TL_DISSOLVES, TL_SPLITS, TL_COUNT };

// ---------------------------------------------------------------- per-slot rules
// AgingSystem.age body (AgingSystem.java:53-63)
__device__ __forceinline__ void age_one(int s, int C, const int *filState, float *nucFrac, float pH, float pD) {
    if (filState[s] < 0) return;
    const float fATP = nucFrac[s], fADPPi = nucFrac[C + s];
    const float nATP = fATP * (1.0f - pH);
    const float nADPPi = fADPPi * (1.0f - pD) + fATP * pH;
    const float nADP = 1.0f - nATP - nADPPi;
    nucFrac[s] = nATP;
    nucFrac[C + s] = nADPPi;
    nucFrac[2 * C + s] = nADP;
}

// DepolySystem.depoly / depolyProxy body (DepolySystem.java:79-100, 124-149): monomers returned; *death set
// when the segment is below the floor
__device__ __forceinline__ int depoly_one(int s, int C, const int *filState, int *monomerCount, float *coord,
                                          const float *uVec, const int *e1s, const float *nucFrac,
                                          const float *depolyParams3, float pATP, float pADP, int step, int seed,
                                          int rng, uint32_t stream, int *death) {
    *death = 0;
    if (filState[s] < 0) return 0;
    if (e1s[s] >= 0) return 0;
    const int M = monomerCount[s];
    const float halfmono = depolyParams3[1];
    const int floorMon = (int)depolyParams3[2];
    float P = depolyParams3[0];
    if (nucFrac) {
        const float adp = nucFrac[2 * C + s];
        P = pATP * (1.0f - adp) + pADP * adp;
    }
    if (M >= floorMon) {
        const float u = uniform01(event_bits(rng, (uint32_t)s, step, seed, stream, SALT_DEPO));
        if (u >= P) return 0;
        monomerCount[s] = M - 1;
        const float h = 0.5f * halfmono;
        coord[s] = coord[s] + h * uVec[s];
        coord[C + s] = coord[C + s] + h * uVec[C + s];
        coord[2 * C + s] = coord[2 * C + s] + h * uVec[2 * C + s];
        return 1;
    }
    *death = 1;
    return M;
}

// DepolySystem.applyDeath (DepolySystem.java:163-186) for slot s: a dying slot frees itself and drops its
// links; a live slot drops its links to dying neighbours
__device__ __forceinline__ void death_one(int s, const int *flag, int *filState, int *monomerCount, int *seedNode,
                                          int *e1s, int *e1d, int *e2s, int *e2d, float *bts, float *brs) {
    const int n1 = e1s[s], n2 = e2s[s];
    if (flag[s]) {
        filState[s] = FIL_FREE;
        bts[s] = 0.0f;
        brs[s] = 0.0f;
        monomerCount[s] = 0;
        if (n2 >= 0) { e2s[s] = -1; e2d[s] = -1; }
        if (n1 >= 0) { e1s[s] = -1; e1d[s] = -1; }
        if (seedNode[s] >= 0) seedNode[s] = -1;
        return;
    }
    if (n2 >= 0 && flag[n2]) { e2s[s] = -1; e2d[s] = -1; }
    if (n1 >= 0 && flag[n1]) { e1s[s] = -1; e1d[s] = -1; }
}

// GrowthSystem.grow body (GrowthSystem.java:83-104): 1 when the tip grew
__device__ __forceinline__ int grow_one(int s, int C, const int *seedNode, int *monomerCount, float *coord,
                                        const float *uVec, float P, float halfmono, int maxMon, int step, int seed,
                                        int rng, uint32_t stream) {
    if (seedNode[s] < 0) return 0;
    if (monomerCount[s] >= maxMon) return 0;
    const float u = uniform01(event_bits(rng, (uint32_t)s, step, seed, stream, SALT_GROW));
    if (u >= P) return 0;
    monomerCount[s] = monomerCount[s] + 1;
    const float h = 0.5f * halfmono;
    coord[s] = coord[s] - h * uVec[s];
    coord[C + s] = coord[C + s] - h * uVec[C + s];
    coord[2 * C + s] = coord[2 * C + s] - h * uVec[2 * C + s];
    return 1;
}

// SeveringSystem.cofilinAccumulate body (SeveringSystem.java:43-49)
__device__ __forceinline__ void cof_acc_one(int s, int C, const int *filState, const float *nucFrac, float *cofFrac,
                                            float pCof) {
    if (filState[s] < 0) return;
    const float fADP = nucFrac[2 * C + s];
    const float fCof = cofFrac[s];
    cofFrac[s] = fCof + (fADP - fCof) * pCof;
}

// GrowthSystem.markSplits body (GrowthSystem.java:124-143); requests at stride RC
__device__ __forceinline__ int mark_split_one(int s, int C, int RC, const int *seedNode, const int *monomerCount,
                                              const float *coord, const float *uVec, const float *yVec,
                                              float *reqCoord, float *reqUVec, float *reqYVec, double halfmono,
                                              int splitThresh, int childMon) {
    if (seedNode[s] < 0) return 0;
    const int M = monomerCount[s];
    if (M < splitThresh) return 0;
    const int parentNewMon = M - childMon;
    const double Lold = (M + 1) * halfmono;
    const double Lpar = (parentNewMon + 1) * halfmono;
    const double Lchild = (childMon + 1) * halfmono;
    const double ux = uVec[s], uy = uVec[C + s], uz = uVec[2 * C + s];
    const double shift = 0.5 * Lold - Lpar - 0.5 * Lchild + halfmono;
    reqCoord[s] = (float)(coord[s] + shift * ux);
    reqCoord[RC + s] = (float)(coord[C + s] + shift * uy);
    reqCoord[2 * RC + s] = (float)(coord[2 * C + s] + shift * uz);
    reqUVec[s] = (float)ux;
    reqUVec[RC + s] = (float)uy;
    reqUVec[2 * RC + s] = (float)uz;
    reqYVec[s] = yVec[s];
    reqYVec[RC + s] = yVec[C + s];
    reqYVec[2 * RC + s] = yVec[2 * C + s];
    return 1;
}

// FilamentBirthSystem.allocate body (FilamentBirthSystem.java:76-84) for request r into slot
__device__ __forceinline__ void alloc_one(int r, int slot, int C, int RC, const float *reqCoord, const float *reqUVec,
                                          const float *reqYVec, float *coord, float *uVec, float *yVec, float *bts,
                                          float *brs, int *filState, float bt, float br) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        coord[a * C + slot] = reqCoord[a * RC + r];
        uVec[a * C + slot] = reqUVec[a * RC + r];
        yVec[a * C + slot] = reqYVec[a * RC + r];
    }
    bts[slot] = bt;
    brs[slot] = br;
    filState[slot] = FIL_ACTIVE;
}

// GrowthSystem.splitWire body (GrowthSystem.java:168-194): parent Gs, child Cs
__device__ __forceinline__ void split_wire_one(int Gs, int Cs, int C, int *monomerCount, float *coord,
                                               const float *uVec, int *e1s, int *e1d, int *e2s, int *e2d,
                                               int *seedNode, double halfmono, int childMon) {
    const int M = monomerCount[Gs];
    const int parentNewMon = M - childMon;
    const int Mold = e1s[Gs], MoldSide = e1d[Gs];
    const double Lold = (M + 1) * halfmono, Lpar = (parentNewMon + 1) * halfmono;
    const double ux = uVec[Gs], uy = uVec[C + Gs], uz = uVec[2 * C + Gs];
    const double shift = 0.5 * Lold - 0.5 * Lpar;
    coord[Gs] = (float)(coord[Gs] + shift * ux);
    coord[C + Gs] = (float)(coord[C + Gs] + shift * uy);
    coord[2 * C + Gs] = (float)(coord[2 * C + Gs] + shift * uz);
    monomerCount[Gs] = parentNewMon;
    monomerCount[Cs] = childMon;
    seedNode[Cs] = -1;
    e1s[Gs] = Cs;
    e1d[Gs] = 1;
    e2s[Cs] = Gs;
    e2d[Cs] = 0;
    e1s[Cs] = Mold;
    e1d[Cs] = Mold >= 0 ? MoldSide : -1;
    if (Mold >= 0) {
        if (MoldSide == 0) e1s[Mold] = Cs;
        else e2s[Mold] = Cs;
    }
}

// GrowthSystem.recomputeDrag body (GrowthSystem.java:211-225)
__device__ __forceinline__ void drag_one(int i, int C, const int *monomerCount, float *segLength, const int *e1s,
                                         const int *e2s, float *bTransGam, float *bRotGam, float *bTransDiff,
                                         float *bRotDiff, const float *dp) {
    const double halfmono = dp[0], aeta = dp[1], radius = dp[2], kT = dp[3];
    const double aPar = dp[4], aOrth = dp[5], aTurn = dp[6];
    const int stdSeg = (int)dp[7], minMon = (int)dp[8];
    const int mc = monomerCount[i];
    const double length = (mc + 1) * halfmono;
    segLength[i] = (float)length;
    const bool atEnd = (e1s[i] < 0) || (e2s[i] < 0);
    const double minLength = (atEnd ? stdSeg : minMon) * halfmono;
    const double asIf = length < minLength ? minLength : length;
    const double LM = 1.0e-6 * asIf, RM = radius * 1.0e-6;
    const double logT = log(LM / (2.0 * RM));
    const double bTGx = (2.0 * SBX_PI_D * aeta * LM) / (logT + aPar);
    const double bTGy = (4.0 * SBX_PI_D * aeta * LM) / (logT + aOrth);
    const double bRGx = 4.0 * SBX_PI_D * aeta * RM * RM * LM;
    const double bRGy = (SBX_PI_D * aeta * (LM * LM * LM)) / (3.0 * (logT + aTurn));
    const int iy = C + i, iz = 2 * C + i;
    bTransGam[i] = (float)bTGx;
    bTransGam[iy] = (float)bTGy;
    bTransGam[iz] = (float)bTGy;
    bRotGam[i] = (float)bRGx;
    bRotGam[iy] = (float)bRGy;
    bRotGam[iz] = (float)bRGy;
    if (bTransDiff) {
        bTransDiff[i] = (float)(kT / bTGx);
        bTransDiff[iy] = (float)(kT / bTGy);
        bTransDiff[iz] = (float)(kT / bTGy);
    }
    if (bRotDiff) {
        bRotDiff[i] = (float)(kT / bRGx);
        bRotDiff[iy] = (float)(kT / bRGy);
        bRotDiff[iz] = (float)(kT / bRGy);
    }
}

// an accepted request r whose rank finds a free slot: its slot, else -1
__device__ __forceinline__ int granted_slot(int r, const int *rankOffsets, const int *freeList, int nFree) {
    const int rank = rankOffsets[r];
    if (!(rankOffsets[r + 1] > rank)) return -1;
    if (rank >= nFree) return -1;
    return freeList[rank];
}

__device__ __forceinline__ void tally_add(int *t, int v) {
    if (v) atomicAdd(t, v);
}

// ---------------------------------------------------------------- cadence passes
// pass 1: age + depoly(Proxy)
__global__ void __launch_bounds__(256) k_turn_depoly(TurnArrays a, sbx_turnover_params p, int *deathFlag, int *tally,
                                                    int step, int rng, uint32_t stream) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= a.C) return;
    age_one(s, a.C, a.filState, a.nucFrac, p.agingParams[0], p.agingParams[1]);
    int death = 0;
    const int ret = depoly_one(s, a.C, a.filState, a.monomerCount, a.coord, a.uVec, a.e1s, p.proxy ? a.nucFrac : nullptr,
                               p.depolyParams, p.depolyRateParams[0], p.depolyRateParams[1], step, p.seed, rng, stream,
                               &death);
    deathFlag[s] = death;
    tally_add(tally + TL_RET, ret);
    tally_add(tally + TL_DEATHS, death);
}

// pass 2: applyDeath (depoly) + grow + cofilinAccumulate + cofilinDissolve
__global__ void __launch_bounds__(256) k_turn_grow(TurnArrays a, sbx_turnover_params p, const int *deathFlag,
                                                  int *severFlag, int *tally, int step, int rng, uint32_t stream) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= a.C) return;
    const int C = a.C;
    death_one(s, deathFlag, a.filState, a.monomerCount, a.seedNode, a.e1s, a.e1d, a.e2s, a.e2d, a.bts, a.brs);
    // GrowthStore.refreshRate (GrowthStore.java:89-92) from the pool as the cadence found it
    const float P = p.growthOn ? (float)(p.onRate * a.pool[0] * p.biochemDeltaT) : 0.0f;
    const int grew = grow_one(s, C, a.seedNode, a.monomerCount, a.coord, a.uVec, P, p.growParams[1],
                              (int)p.growParams[2], step, p.seed, rng, stream);
    tally_add(tally + TL_GREW, grew);
    cof_acc_one(s, C, a.filState, a.nucFrac, a.cofFrac, p.cofilinParams[0]);
    int sev = 0, ret = 0;
    if (a.filState[s] >= 0 && a.cofFrac[s] > p.cofilinParams[1]) {  // SeveringSystem.java:64-71
        sev = 1;
        ret = a.monomerCount[s];
    }
    severFlag[s] = sev;
    tally_add(tally + TL_SEV, ret);
    tally_add(tally + TL_DISSOLVES, sev);
}

// pass 3: applyDeath (dissolve) + markSplits + freeFlags
__global__ void __launch_bounds__(256) k_turn_mark(TurnArrays a, sbx_turnover_params p, const int *severFlag,
                                                  int *acceptFlag, int *freeCount, float *reqCoord, float *reqUVec,
                                                  float *reqYVec) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= a.C) return;
    const int C = a.C;
    death_one(s, severFlag, a.filState, a.monomerCount, a.seedNode, a.e1s, a.e1d, a.e2s, a.e2d, a.bts, a.brs);
    acceptFlag[s] = mark_split_one(s, C, C, a.seedNode, a.monomerCount, a.coord, a.uVec, a.yVec, reqCoord, reqUVec,
                                   reqYVec, (double)p.splitParams[0], (int)p.splitParams[1], (int)p.splitParams[2]);
    freeCount[s] = a.filState[s] < 0 ? 1 : 0;
}

// FilamentBirthSystem.freeScatter (FilamentBirthSystem.java:56-63), one slot per thread
__global__ void __launch_bounds__(256) k_free_scatter_fil(const int *filState, const int *freeOffsets, int *freeList,
                                                         int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C && filState[s] < 0) freeList[freeOffsets[s]] = s;
}

// pass 4: allocate + splitWire + splitInheritNuc + nucleateFreshCofilin per granted split
__global__ void __launch_bounds__(256) k_turn_split(TurnArrays a, sbx_turnover_params p, const int *rankOffsets,
                                                   const int *freeList, const int *freeOffsets, const float *reqCoord,
                                                   const float *reqUVec, const float *reqYVec, int *tally) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int C = a.C;
    if (r >= C) return;
    const int Cs = granted_slot(r, rankOffsets, freeList, freeOffsets[C]);
    if (Cs < 0) return;
    alloc_one(r, Cs, C, C, reqCoord, reqUVec, reqYVec, a.coord, a.uVec, a.yVec, a.bts, a.brs, a.filState,
              p.birthParams[0], p.birthParams[1]);
    split_wire_one(r, Cs, C, a.monomerCount, a.coord, a.uVec, a.e1s, a.e1d, a.e2s, a.e2d, a.seedNode,
                   (double)p.splitParams[0], (int)p.splitParams[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) a.nucFrac[k * C + Cs] = a.nucFrac[k * C + r];
    a.cofFrac[Cs] = 0.0f;
    atomicAdd(tally + TL_SPLITS, 1);
}

// pass 5: recomputeDrag over every slot
__global__ void __launch_bounds__(256) k_turn_drag(TurnArrays a, sbx_turnover_params p) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.C) return;
    drag_one(i, a.C, a.monomerCount, a.segLength, a.e1s, a.e2s, a.bTransGam, a.bRotGam, a.bTransDiff, a.bRotDiff,
             p.dragParams);
}

// the pool (ActinPool.take / put, FullSystemDemoHarness.java:662-672 order) and the cadence's counters
__global__ void k_turn_pool(TurnArrays a, sbx_turnover_params p, int *tally) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int nGrew = tally[TL_GREW], nRet = tally[TL_RET], nSev = tally[TL_SEV];
    double conc = a.pool[0];
    conc += nRet * p.uMPerMonomer;  // put(returned)
    a.ledger[1] += nRet;
    if (nGrew > 0) {  // depletePoolForGrows -> take(nGrew)
        conc -= nGrew * p.uMPerMonomer;
        if (conc < 0.0) conc = 0.0;
        a.ledger[0] += nGrew;
    }
    conc += nSev * p.uMPerMonomer;  // put(dissolved)
    a.ledger[1] += nSev;
    a.pool[0] = conc;
    for (int k = 0; k < TL_COUNT; k++) {
        a.stats[k] = tally[k];
        tally[k] = 0;
    }
    a.stats[6] += 1;
}

// ---------------------------------------------------------------- stand-alone System kernels
__global__ void __launch_bounds__(256) k_age(const int *filState, float *nucFrac, float pH, float pD, int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C) age_one(s, C, filState, nucFrac, pH, pD);
}

__global__ void __launch_bounds__(256) k_growth_atp(const int *grewFlag, const int *monomerCount, float *nucFrac,
                                                   int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C || grewFlag[s] == 0) return;
    const int M = monomerCount[s];
    if (M <= 0) return;
    const float w = (float)(M - 1) / (float)M;  // AgingSystem.java:77-82
    const float inv = 1.0f / (float)M;
    nucFrac[s] = nucFrac[s] * w + inv;
    nucFrac[C + s] = nucFrac[C + s] * w;
    nucFrac[2 * C + s] = nucFrac[2 * C + s] * w;
}

__global__ void __launch_bounds__(256) k_split_inherit(const int *rankOffsets, const int *freeList,
                                                      const int *freeOffsets, float *nucFrac, int C, int K) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    const int Cs = granted_slot(r, rankOffsets, freeList, freeOffsets[C]);
    if (Cs < 0) return;
#pragma unroll
    for (int k = 0; k < 3; k++) nucFrac[k * C + Cs] = nucFrac[k * C + r];
}

__global__ void __launch_bounds__(256) k_depoly(const int *filState, int *monomerCount, float *coord,
                                               const float *uVec, const int *e1s, const float *nucFrac,
                                               int *returnedMon, int *deathFlag, float P, float halfmono,
                                               float floorMon, float pATP, float pADP, int step, int seed, int fires,
                                               int C, int rng, uint32_t stream) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C) return;
    returnedMon[s] = 0;
    deathFlag[s] = 0;
    if (fires == 0) return;
    const float dp[3] = {P, halfmono, floorMon};
    int death = 0;
    const int ret = depoly_one(s, C, filState, monomerCount, coord, uVec, e1s, nucFrac, dp, pATP, pADP, step, seed,
                               rng, stream, &death);
    returnedMon[s] = ret;
    deathFlag[s] = death;
}

__global__ void __launch_bounds__(256) k_apply_death(int *filState, int *monomerCount, int *seedNode, int *e1s,
                                                    int *e1d, int *e2s, int *e2d, float *bts, float *brs,
                                                    const int *deathFlag, int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C) death_one(s, deathFlag, filState, monomerCount, seedNode, e1s, e1d, e2s, e2d, bts, brs);
}

__global__ void __launch_bounds__(256) k_grow(const int *seedNode, int *monomerCount, float *coord, const float *uVec,
                                             int *grewFlag, float P, float halfmono, int maxMon, int step, int seed,
                                             int fires, int C, int rng, uint32_t stream) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C) return;
    grewFlag[s] = 0;
    if (fires == 0) return;
    grewFlag[s] = grow_one(s, C, seedNode, monomerCount, coord, uVec, P, halfmono, maxMon, step, seed, rng, stream);
}

__global__ void __launch_bounds__(256) k_mark_splits(const int *seedNode, const int *monomerCount, const float *coord,
                                                    const float *uVec, const float *yVec, int *acceptFlag,
                                                    float *reqCoord, float *reqUVec, float *reqYVec, float halfmono,
                                                    int splitThresh, int childMon, int fires, int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C) return;
    acceptFlag[s] = 0;
    if (fires == 0) return;
    acceptFlag[s] = mark_split_one(s, C, C, seedNode, monomerCount, coord, uVec, yVec, reqCoord, reqUVec, reqYVec,
                                   (double)halfmono, splitThresh, childMon);
}

__global__ void __launch_bounds__(256) k_split_wire(const int *rankOffsets, const int *freeList,
                                                   const int *freeOffsets, int *monomerCount, float *coord,
                                                   const float *uVec, int *e1s, int *e1d, int *e2s, int *e2d,
                                                   int *seedNode, float halfmono, int childMon, int C, int K) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    const int Cs = granted_slot(r, rankOffsets, freeList, freeOffsets[C]);
    if (Cs < 0) return;
    split_wire_one(r, Cs, C, monomerCount, coord, uVec, e1s, e1d, e2s, e2d, seedNode, (double)halfmono, childMon);
}

struct DragArgs {
    float dp[9];
};

__global__ void __launch_bounds__(256) k_recompute_drag(const int *monomerCount, float *segLength, const int *e1s,
                                                       const int *e2s, float *bTransGam, float *bRotGam,
                                                       float *bTransDiff, float *bRotDiff, DragArgs d, int C) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < C)
        drag_one(i, C, monomerCount, segLength, e1s, e2s, bTransGam, bRotGam, bTransDiff, bRotDiff, d.dp);
}

__global__ void __launch_bounds__(256) k_cof_acc(const int *filState, const float *nucFrac, float *cofFrac, float pCof,
                                                int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C) cof_acc_one(s, C, filState, nucFrac, cofFrac, pCof);
}

__global__ void __launch_bounds__(256) k_cof_dissolve(const int *filState, const int *monomerCount,
                                                     const float *cofFrac, int *severDeathFlag, int *severReturnedMon,
                                                     float ratio, int fires, int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C) return;
    severDeathFlag[s] = 0;
    severReturnedMon[s] = 0;
    if (fires == 0 || filState[s] < 0) return;
    if (cofFrac[s] > ratio) {
        severDeathFlag[s] = 1;
        severReturnedMon[s] = monomerCount[s];
    }
}

__global__ void __launch_bounds__(256) k_cof_fresh(const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                                  float *cofFrac, int C, int K) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    const int Cs = granted_slot(r, rankOffsets, freeList, freeOffsets[C]);
    if (Cs >= 0) cofFrac[Cs] = 0.0f;
}

__global__ void __launch_bounds__(256) k_free_flags_fil(const int *filState, int *freeCount, int C) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < C) freeCount[s] = filState[s] < 0 ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_birth_allocate(const int *rankOffsets, const int *freeList,
                                                       const int *freeOffsets, const float *reqCoord,
                                                       const float *reqUVec, const float *reqYVec, float *coord,
                                                       float *uVec, float *yVec, float *bts, float *brs,
                                                       int *filState, float bt, float br, int C, int K, int RC) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    const int slot = granted_slot(r, rankOffsets, freeList, freeOffsets[C]);
    if (slot >= 0) alloc_one(r, slot, C, RC, reqCoord, reqUVec, reqYVec, coord, uVec, yVec, bts, brs, filState, bt, br);
}

}  // namespace

int turn_scratch_reserve(sbx_ctx *c, TurnScratch &s, int C) {
    if (C <= s.cap) return SBX_OK;
    turn_scratch_free(s);
    const size_t n = (size_t)C;
    int r = SBX_OK;
    int **ints[] = {&s.deathFlag, &s.severFlag, &s.acceptFlag, &s.freeCount, &s.freeList};
    for (int **p : ints)
        if (!r) r = check_hip(c, hipMalloc(p, sizeof(int) * n), "hipMalloc turnover scratch");
    if (!r) r = check_hip(c, hipMalloc(&s.freeOffsets, sizeof(int) * (n + 1)), "hipMalloc turnover scratch");
    if (!r) r = check_hip(c, hipMalloc(&s.rankOffsets, sizeof(int) * (n + 1)), "hipMalloc turnover scratch");
    if (!r) r = check_hip(c, hipMalloc(&s.tally, sizeof(int) * 8), "hipMalloc turnover scratch");
    float **flts[] = {&s.reqCoord, &s.reqUVec, &s.reqYVec};
    for (float **p : flts)
        if (!r) r = check_hip(c, hipMalloc(p, sizeof(float) * 3 * n), "hipMalloc turnover scratch");
    if (!r) r = fill_i32(c, s.tally, 0, 8);
    if (r) return r;
    s.cap = C;
    return SBX_OK;
}

void turn_scratch_destroy(TurnScratch *s) {
    if (!s) return;
    turn_scratch_free(*s);
    delete s;
}

void turn_scratch_free(TurnScratch &s) {
    void *ps[] = {s.deathFlag, s.severFlag, s.acceptFlag, s.freeCount, s.freeList, s.freeOffsets, s.rankOffsets,
                  s.tally, s.reqCoord, s.reqUVec, s.reqYVec};
    for (void *p : ps)
        if (p) hipFree(p);
    sort_scratch_free(s.scan);
    s = TurnScratch();
}

int turnover_cadence(sbx_ctx *c, const TurnArrays &a, const sbx_turnover_params &p, TurnScratch &s, int step, int rng,
                     uint32_t stream, hipStream_t q) {
    const int C = a.C;
    if (C <= 0) return SBX_OK;
    const unsigned nb = blocks_for(C, 256);
    hipLaunchKernelGGL(k_turn_depoly, dim3(nb), dim3(256), 0, q, a, p, s.deathFlag, s.tally, step, rng, stream);
    hipLaunchKernelGGL(k_turn_grow, dim3(nb), dim3(256), 0, q, a, p, s.deathFlag, s.severFlag, s.tally, step, rng,
                       stream);
    hipLaunchKernelGGL(k_turn_mark, dim3(nb), dim3(256), 0, q, a, p, s.severFlag, s.acceptFlag, s.freeCount,
                       s.reqCoord, s.reqUVec, s.reqYVec);
    int r = check_launch(c, "turnover passes 1-3");
    if (r) return r;
    if ((r = exclusive_scan(c, s.scan, s.freeCount, C, s.freeOffsets, q))) return r;
    if ((r = exclusive_scan(c, s.scan, s.acceptFlag, C, s.rankOffsets, q))) return r;
    hipLaunchKernelGGL(k_free_scatter_fil, dim3(nb), dim3(256), 0, q, a.filState, s.freeOffsets, s.freeList, C);
    hipLaunchKernelGGL(k_turn_split, dim3(nb), dim3(256), 0, q, a, p, s.rankOffsets, s.freeList, s.freeOffsets,
                       s.reqCoord, s.reqUVec, s.reqYVec, s.tally);
    hipLaunchKernelGGL(k_turn_drag, dim3(nb), dim3(256), 0, q, a, p);
    hipLaunchKernelGGL(k_turn_pool, dim3(1), dim3(64), 0, q, a, p, s.tally);
    return check_launch(c, "turnover passes 4-6");
}

}  // namespace sbx

using namespace sbx;

// ------------------------------------------------------------------ C ABI
extern "C" {

int sbx_aging_age(sbx_ctx *c, const int *filState, float *nucFrac, const float *agingParams, int fires, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && agingParams, "aging_age: bad arguments");
    if (C == 0 || !fires) return SBX_OK;
    SBX_REQUIRE(c, filState && nucFrac, "aging_age: null array");
    hipLaunchKernelGGL(k_age, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, filState, nucFrac, agingParams[0],
                       agingParams[1], C);
    return check_launch(c, "k_age");
}

int sbx_aging_growth_atp(sbx_ctx *c, const int *grewFlag, const int *monomerCount, float *nucFrac, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0, "aging_growth_atp: C < 0");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, grewFlag && monomerCount && nucFrac, "aging_growth_atp: null array");
    hipLaunchKernelGGL(k_growth_atp, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, grewFlag, monomerCount,
                       nucFrac, C);
    return check_launch(c, "k_growth_atp");
}

int sbx_aging_split_inherit_nuc(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                float *nucFrac, int C, int K) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0 && K <= C, "aging_split_inherit_nuc: bad sizes");
    if (K == 0) return SBX_OK;
    SBX_REQUIRE(c, rankOffsets && freeList && freeOffsets && nucFrac, "aging_split_inherit_nuc: null array");
    hipLaunchKernelGGL(k_split_inherit, dim3(blocks_for(K, 256)), dim3(256), 0, c->stream, rankOffsets, freeList,
                       freeOffsets, nucFrac, C, K);
    return check_launch(c, "k_split_inherit");
}

int sbx_depoly(sbx_ctx *c, const int *filState, int *monomerCount, float *coord, const float *uVec,
               const int *end1NbrSlot, const float *nucFrac, int *returnedMon, int *deathFlag,
               const float *depolyParams, const float *depolyRateParams, int step, int seed, int fires, int C, int rng,
               uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && depolyParams, "depoly: bad arguments");
    SBX_REQUIRE(c, !nucFrac || depolyRateParams, "depoly: the proxy needs depolyRateParams");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "depoly: unknown rng %d", rng);
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, filState && monomerCount && coord && uVec && end1NbrSlot && returnedMon && deathFlag,
                "depoly: null array");
    const float pATP = depolyRateParams ? depolyRateParams[0] : 0.0f, pADP = depolyRateParams ? depolyRateParams[1] : 0.0f;
    hipLaunchKernelGGL(k_depoly, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, filState, monomerCount, coord,
                       uVec, end1NbrSlot, nucFrac, returnedMon, deathFlag, depolyParams[0], depolyParams[1],
                       depolyParams[2], pATP, pADP, step, seed, fires, C, rng, stream);
    return check_launch(c, "k_depoly");
}

int sbx_apply_death(sbx_ctx *c, int *filState, int *monomerCount, int *seedNode, int *end1NbrSlot, int *end1NbrSide,
                    int *end2NbrSlot, int *end2NbrSide, float *brownTransScale, float *brownRotScale,
                    const int *deathFlag, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0, "apply_death: C < 0");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, filState && monomerCount && seedNode && end1NbrSlot && end1NbrSide && end2NbrSlot && end2NbrSide &&
                       brownTransScale && brownRotScale && deathFlag,
                "apply_death: null array");
    hipLaunchKernelGGL(k_apply_death, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, filState, monomerCount,
                       seedNode, end1NbrSlot, end1NbrSide, end2NbrSlot, end2NbrSide, brownTransScale, brownRotScale,
                       deathFlag, C);
    return check_launch(c, "k_apply_death");
}

int sbx_growth_grow(sbx_ctx *c, const int *seedNode, int *monomerCount, float *coord, const float *uVec,
                    int *grewFlag, const float *growParams, int step, int seed, int fires, int C, int rng,
                    uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && growParams, "growth_grow: bad arguments");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "growth_grow: unknown rng %d", rng);
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, seedNode && monomerCount && coord && uVec && grewFlag, "growth_grow: null array");
    hipLaunchKernelGGL(k_grow, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, seedNode, monomerCount, coord, uVec,
                       grewFlag, growParams[0], growParams[1], (int)growParams[2], step, seed, fires, C, rng, stream);
    return check_launch(c, "k_grow");
}

int sbx_growth_mark_splits(sbx_ctx *c, const int *seedNode, const int *monomerCount, const float *coord,
                           const float *uVec, const float *yVec, int *acceptFlag, float *reqCoord, float *reqUVec,
                           float *reqYVec, const float *splitParams, int fires, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && splitParams, "growth_mark_splits: bad arguments");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, seedNode && monomerCount && coord && uVec && yVec && acceptFlag && reqCoord && reqUVec && reqYVec,
                "growth_mark_splits: null array");
    hipLaunchKernelGGL(k_mark_splits, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, seedNode, monomerCount, coord,
                       uVec, yVec, acceptFlag, reqCoord, reqUVec, reqYVec, splitParams[0], (int)splitParams[1],
                       (int)splitParams[2], fires, C);
    return check_launch(c, "k_mark_splits");
}

int sbx_growth_split_wire(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                          int *monomerCount, float *coord, const float *uVec, int *end1NbrSlot, int *end1NbrSide,
                          int *end2NbrSlot, int *end2NbrSide, int *seedNode, const float *splitParams, int C, int K) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0 && K <= C && splitParams, "growth_split_wire: bad arguments");
    if (K == 0) return SBX_OK;
    SBX_REQUIRE(c, rankOffsets && freeList && freeOffsets && monomerCount && coord && uVec && end1NbrSlot &&
                       end1NbrSide && end2NbrSlot && end2NbrSide && seedNode,
                "growth_split_wire: null array");
    hipLaunchKernelGGL(k_split_wire, dim3(blocks_for(K, 256)), dim3(256), 0, c->stream, rankOffsets, freeList,
                       freeOffsets, monomerCount, coord, uVec, end1NbrSlot, end1NbrSide, end2NbrSlot, end2NbrSide,
                       seedNode, splitParams[0], (int)splitParams[2], C, K);
    return check_launch(c, "k_split_wire");
}

int sbx_growth_recompute_drag(sbx_ctx *c, const int *monomerCount, float *segLength, const int *end1NbrSlot,
                              const int *end2NbrSlot, float *bTransGam, float *bRotGam, float *bTransDiff,
                              float *bRotDiff, const float *dragParams, int fires, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && dragParams, "growth_recompute_drag: bad arguments");
    if (C == 0 || !fires) return SBX_OK;
    SBX_REQUIRE(c, monomerCount && segLength && end1NbrSlot && end2NbrSlot && bTransGam && bRotGam,
                "growth_recompute_drag: null array");
    DragArgs d;
    for (int k = 0; k < 9; k++) d.dp[k] = dragParams[k];
    hipLaunchKernelGGL(k_recompute_drag, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, monomerCount, segLength,
                       end1NbrSlot, end2NbrSlot, bTransGam, bRotGam, bTransDiff, bRotDiff, d, C);
    return check_launch(c, "k_recompute_drag");
}

int sbx_cofilin_accumulate(sbx_ctx *c, const int *filState, const float *nucFrac, float *cofFrac,
                           const float *cofilinParams, int fires, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && cofilinParams, "cofilin_accumulate: bad arguments");
    if (C == 0 || !fires) return SBX_OK;
    SBX_REQUIRE(c, filState && nucFrac && cofFrac, "cofilin_accumulate: null array");
    hipLaunchKernelGGL(k_cof_acc, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, filState, nucFrac, cofFrac,
                       cofilinParams[0], C);
    return check_launch(c, "k_cof_acc");
}

int sbx_cofilin_dissolve(sbx_ctx *c, const int *filState, const int *monomerCount, const float *cofFrac,
                         int *severDeathFlag, int *severReturnedMon, const float *cofilinParams, int fires, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && cofilinParams, "cofilin_dissolve: bad arguments");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, filState && monomerCount && cofFrac && severDeathFlag && severReturnedMon,
                "cofilin_dissolve: null array");
    hipLaunchKernelGGL(k_cof_dissolve, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, filState, monomerCount,
                       cofFrac, severDeathFlag, severReturnedMon, cofilinParams[1], fires, C);
    return check_launch(c, "k_cof_dissolve");
}

int sbx_cofilin_fresh(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets, float *cofFrac,
                      int C, int K) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0 && K <= C, "cofilin_fresh: bad sizes");
    if (K == 0) return SBX_OK;
    SBX_REQUIRE(c, rankOffsets && freeList && freeOffsets && cofFrac, "cofilin_fresh: null array");
    hipLaunchKernelGGL(k_cof_fresh, dim3(blocks_for(K, 256)), dim3(256), 0, c->stream, rankOffsets, freeList,
                       freeOffsets, cofFrac, C, K);
    return check_launch(c, "k_cof_fresh");
}

int sbx_birth_alloc(sbx_ctx *c, int *filState, int *freeCount, int *freeOffsets, int *freeList, int *acceptFlag,
                    int *rankOffsets, const float *reqCoord, const float *reqUVec, const float *reqYVec, float *coord,
                    float *uVec, float *yVec, float *brownTransScale, float *brownRotScale, const float *birthParams,
                    int C, int K, int RC) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0 && RC >= K && birthParams, "birth_alloc: bad arguments");
    SBX_REQUIRE(c, filState && freeCount && freeOffsets && freeList && acceptFlag && rankOffsets,
                "birth_alloc: null array");
    SBX_REQUIRE(c, K == 0 || (reqCoord && reqUVec && reqYVec && coord && uVec && yVec && brownTransScale &&
                              brownRotScale),
                "birth_alloc: null array");
    hipStream_t q = c->stream;
    if (C > 0) hipLaunchKernelGGL(k_free_flags_fil, dim3(blocks_for(C, 256)), dim3(256), 0, q, filState, freeCount, C);
    int r = exclusive_scan(c, c->api_sort, freeCount, C, freeOffsets, q);
    if (r) return r;
    if (C > 0)
        hipLaunchKernelGGL(k_free_scatter_fil, dim3(blocks_for(C, 256)), dim3(256), 0, q, filState, freeOffsets,
                           freeList, C);
    if ((r = exclusive_scan(c, c->api_sort, acceptFlag, K, rankOffsets, q))) return r;
    if (K > 0)
        hipLaunchKernelGGL(k_birth_allocate, dim3(blocks_for(K, 256)), dim3(256), 0, q, rankOffsets, freeList,
                           freeOffsets, reqCoord, reqUVec, reqYVec, coord, uVec, yVec, brownTransScale, brownRotScale,
                           filState, birthParams[0], birthParams[1], C, K, RC);
    if ((r = check_launch(c, "birth allocate"))) return r;
    // csrScan's cursor side effect (CrossBridgeSystem.java:202): both counted arrays come back cleared
    if ((r = fill_i32(c, freeCount, 0, C))) return r;
    return fill_i32(c, acceptFlag, 0, K);
}

int sbx_turnover_cadence(sbx_ctx *c, const sbx_filaments *fil, const sbx_turnover *t, const sbx_turnover_params *p,
                         int step, int rng, uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, fil && t && p, "turnover_cadence: null argument");
    SBX_REQUIRE(c, t->C == fil->n, "turnover_cadence: turnover C %d != filament n %d", t->C, fil->n);
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "turnover_cadence: unknown rng %d", rng);
    SBX_REQUIRE(c, t->filState && t->monomerCount && t->seedNode && t->nucFrac && t->cofFrac && t->pool && t->ledger &&
                       t->stats,
                "turnover_cadence: null turnover array");
    SBX_REQUIRE(c, fil->end1NbrSlot && fil->end1NbrSide && fil->end2NbrSlot && fil->end2NbrSide,
                "turnover_cadence: chain topology arrays null");
    int r = set_device(c);
    if (r) return r;
    if (!c->turn) {
        c->turn = new (std::nothrow) TurnScratch();
        SBX_REQUIRE(c, c->turn, "turnover_cadence: out of host memory");
    }
    if ((r = turn_scratch_reserve(c, *c->turn, t->C))) return r;
    TurnArrays a{};
    a.C = t->C;
    a.coord = fil->coord; a.uVec = fil->uVec; a.yVec = fil->yVec; a.segLength = fil->segLength;
    a.bTransGam = fil->bTransGam; a.bRotGam = fil->bRotGam; a.bts = fil->brownTransScale; a.brs = fil->brownRotScale;
    a.e1s = fil->end1NbrSlot; a.e1d = fil->end1NbrSide; a.e2s = fil->end2NbrSlot; a.e2d = fil->end2NbrSide;
    a.filState = t->filState; a.monomerCount = t->monomerCount; a.seedNode = t->seedNode;
    a.nucFrac = t->nucFrac; a.cofFrac = t->cofFrac; a.bTransDiff = t->bTransDiff; a.bRotDiff = t->bRotDiff;
    a.pool = t->pool; a.ledger = t->ledger; a.stats = t->stats;
    return turnover_cadence(c, a, *p, *c->turn, step, rng, stream, c->stream);
}

}  // extern "C"
