/*
 * softbox_oracle_turnover.c -- CPU restatement of SoftBox's actin turnover (test infrastructure).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see softbox_oracle.h).
 *
 * Per-segment systems, each a restatement of the cited Java method:
 *   AgingSystem.age / growthAtp / splitInheritNuc / nucleateFreshAtp   (AgingSystem.java:44-124)
 *   DepolySystem.depoly / depolyProxy / applyDeath                     (DepolySystem.java:70-187)
 *   GrowthSystem.grow / markSplits / splitWire / recomputeDrag         (GrowthSystem.java:73-227)
 *   SeveringSystem.cofilinAccumulate / cofilinDissolve / nucleateFreshCofilin (SeveringSystem.java:37-97)
 *   FilamentBirthSystem.freeFlags / freeScatter / allocate             (FilamentBirthSystem.java:41-86)
 *   CrossBridgeSystem.csrScan (exclusive prefix that clears its input, CrossBridgeSystem.java:198-205)
 *   ActinPool.take / put (ActinPool.java)
 * and the biochem cadence that composes them in FullSystemDemoHarness.cpuStep's order
 * (FullSystemDemoHarness.java:650-684; one growth pass, no node nucleation).  The random draws of grow and
 * depoly use the network's stream: Wang keys by the reference's (slot, step, seed) base with the salts GROW
 * and DEPO, or Philox.
 */
#include <math.h>
#include <string.h>

#include "softbox_oracle.h"

#define SALT_GROW 0x47524F57u
#define SALT_DEPO 0x4445504Fu

/* ------------------------------------------------------------------ aging */
/* AgingSystem.age (AgingSystem.java:44-64): forward-Euler cascade, f_ADP as the remainder */
void sbo_aging_age(int C, const int *filState, float *nucFrac, const float agingParams[2], int fires) {
    float pH = agingParams[0], pD = agingParams[1];
    for (int s = 0; s < C; s++) {
        if (fires == 0) continue;
        if (filState[s] < 0) continue;
        float fATP = nucFrac[s], fADPPi = nucFrac[C + s];
        float nATP = fATP * (1.0f - pH);
        float nADPPi = fADPPi * (1.0f - pD) + fATP * pH;
        float nADP = 1.0f - nATP - nADPPi;
        nucFrac[s] = nATP;
        nucFrac[C + s] = nADPPi;
        nucFrac[2 * C + s] = nADP;
    }
}

/* AgingSystem.growthAtp (AgingSystem.java:72-84): mix one fresh ATP monomer into an M-monomer segment */
void sbo_aging_growth_atp(int C, const int *grewFlag, const int *monomerCount, float *nucFrac) {
    for (int s = 0; s < C; s++) {
        if (grewFlag[s] == 0) continue;
        int M = monomerCount[s];
        if (M <= 0) continue;
        float w = (float)(M - 1) / (float)M;
        float inv = 1.0f / (float)M;
        nucFrac[s] = nucFrac[s] * w + inv;
        nucFrac[C + s] = nucFrac[C + s] * w;
        nucFrac[2 * C + s] = nucFrac[2 * C + s] * w;
    }
}

/* AgingSystem.splitInheritNuc (AgingSystem.java:94-108): the split child copies its parent's composition */
void sbo_aging_split_inherit_nuc(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                 float *nucFrac) {
    int nFree = freeOffsets[C];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int Gs = r, Cs = freeList[rank];
        nucFrac[Cs] = nucFrac[Gs];
        nucFrac[C + Cs] = nucFrac[C + Gs];
        nucFrac[2 * C + Cs] = nucFrac[2 * C + Gs];
    }
}

/* AgingSystem.nucleateFreshAtp (AgingSystem.java:112-124): a nucleated seed is pure ATP */
void sbo_aging_nucleate_fresh_atp(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                  float *nucFrac) {
    int nFree = freeOffsets[C];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int slot = freeList[rank];
        nucFrac[slot] = 1.0f;
        nucFrac[C + slot] = 0.0f;
        nucFrac[2 * C + slot] = 0.0f;
    }
}

/* ------------------------------------------------------------------ depolymerisation + death */
/* DepolySystem.depoly (DepolySystem.java:70-101) with nucFrac == NULL, else DepolySystem.depolyProxy
 * (:115-150): the pointed tip (end1 free) loses one monomer with probability P (fixed, or interpolated
 * by the segment's ADP fraction), end2 held; below the floor the segment dies en masse. */
void sbo_depoly(int C, const int *filState, int *monomerCount, float *coord, const float *uVec,
                const int *end1NbrSlot, const float *nucFrac, int *returnedMon, int *deathFlag,
                const float depolyParams[3], const float depolyRateParams[2], int32_t step, int32_t seed, int rng,
                uint32_t stream, int fires) {
    float halfmono = depolyParams[1];
    int floorMon = (int)depolyParams[2];
    for (int s = 0; s < C; s++) {
        returnedMon[s] = 0;
        deathFlag[s] = 0;
        if (fires == 0) continue;
        if (filState[s] < 0) continue;
        if (end1NbrSlot[s] >= 0) continue;
        int M = monomerCount[s];
        float P = depolyParams[0];
        if (nucFrac) {
            float adp = nucFrac[2 * C + s];
            P = depolyRateParams[0] * (1.0f - adp) + depolyRateParams[1] * adp;
        }
        if (M >= floorMon) {
            float u = sbo_uniform(sbo_event_bits(rng, (uint32_t)s, step, seed, stream, SALT_DEPO));
            if (u >= P) continue;
            monomerCount[s] = M - 1;
            float ux = uVec[s], uy = uVec[C + s], uz = uVec[2 * C + s];
            coord[s] = coord[s] + 0.5f * halfmono * ux;
            coord[C + s] = coord[C + s] + 0.5f * halfmono * uy;
            coord[2 * C + s] = coord[2 * C + s] + 0.5f * halfmono * uz;
            returnedMon[s] = 1;
        } else {
            deathFlag[s] = 1;
            returnedMon[s] = M;
        }
    }
}

/* DepolySystem.applyDeath (DepolySystem.java:159-187): markFree, break both chain links (clear the
 * neighbours' back-pointers), clear the node tether */
void sbo_apply_death(int C, int *filState, int *monomerCount, int *seedNode, int *e1s, int *e1d, int *e2s, int *e2d,
                     float *bts, float *brs, const int *deathFlag) {
    for (int s = 0; s < C; s++) {
        if (deathFlag[s] == 0) continue;
        filState[s] = SBO_FIL_FREE;
        bts[s] = 0.0f;
        brs[s] = 0.0f;
        monomerCount[s] = 0;
        int n2 = e2s[s], n2side = e2d[s];
        if (n2 >= 0) {
            if (n2side == 0) { e1s[n2] = -1; e1d[n2] = -1; }
            else { e2s[n2] = -1; e2d[n2] = -1; }
            e2s[s] = -1; e2d[s] = -1;
        }
        int n1 = e1s[s], n1side = e1d[s];
        if (n1 >= 0) {
            if (n1side == 0) { e1s[n1] = -1; e1d[n1] = -1; }
            else { e2s[n1] = -1; e2d[n1] = -1; }
            e1s[s] = -1; e1d[s] = -1;
        }
        if (seedNode[s] >= 0) seedNode[s] = -1;
    }
}

/* ------------------------------------------------------------------ growth */
/* GrowthSystem.grow (GrowthSystem.java:73-106): a node-held tip below maxMon gains one monomer with
 * probability P, end2 held, end1 extending outward */
void sbo_grow(int C, const int *seedNode, int *monomerCount, float *coord, const float *uVec, int *grewFlag,
              const float growParams[3], int32_t step, int32_t seed, int rng, uint32_t stream, int fires) {
    float P = growParams[0], halfmono = growParams[1];
    int maxMon = (int)growParams[2];
    for (int s = 0; s < C; s++) {
        grewFlag[s] = 0;
        if (fires == 0) continue;
        if (seedNode[s] < 0) continue;
        if (monomerCount[s] >= maxMon) continue;
        float u = sbo_uniform(sbo_event_bits(rng, (uint32_t)s, step, seed, stream, SALT_GROW));
        if (u >= P) continue;
        monomerCount[s] = monomerCount[s] + 1;
        float ux = uVec[s], uy = uVec[C + s], uz = uVec[2 * C + s];
        coord[s] = coord[s] - 0.5f * halfmono * ux;
        coord[C + s] = coord[C + s] - 0.5f * halfmono * uy;
        coord[2 * C + s] = coord[2 * C + s] - 0.5f * halfmono * uz;
        grewFlag[s] = 1;
    }
}

/* GrowthSystem.markSplits (GrowthSystem.java:115-145): a tip at splitThresh stages its outer half as a
 * birth request at index s (request stride RC) */
void sbo_mark_splits(int C, const int *seedNode, const int *monomerCount, const float *coord, const float *uVec,
                     const float *yVec, int *acceptFlag, float *reqCoord, float *reqUVec, float *reqYVec, int RC,
                     const float splitParams[3], int fires) {
    double halfmono = splitParams[0];
    int splitThresh = (int)splitParams[1];
    int childMon = (int)splitParams[2];
    for (int s = 0; s < C; s++) {
        acceptFlag[s] = 0;
        if (fires == 0) continue;
        if (seedNode[s] < 0) continue;
        int M = monomerCount[s];
        if (M < splitThresh) continue;
        int parentNewMon = M - childMon;
        double Lold = (M + 1) * halfmono;
        double Lpar = (parentNewMon + 1) * halfmono;
        double Lchild = (childMon + 1) * halfmono;
        double ux = uVec[s], uy = uVec[C + s], uz = uVec[2 * C + s];
        double shift = 0.5 * Lold - Lpar - 0.5 * Lchild + halfmono;
        reqCoord[s] = (float)(coord[s] + shift * ux);
        reqCoord[RC + s] = (float)(coord[C + s] + shift * uy);
        reqCoord[2 * RC + s] = (float)(coord[2 * C + s] + shift * uz);
        reqUVec[s] = (float)ux; reqUVec[RC + s] = (float)uy; reqUVec[2 * RC + s] = (float)uz;
        reqYVec[s] = yVec[s]; reqYVec[RC + s] = yVec[C + s]; reqYVec[2 * RC + s] = yVec[2 * C + s];
        acceptFlag[s] = 1;
    }
}

/* GrowthSystem.splitWire (GrowthSystem.java:155-196): shrink the parent (end2 held), finish the child,
 * rewire the three slots parent - child - old outward neighbour */
void sbo_split_wire(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                    int *monomerCount, float *coord, const float *uVec, int *e1s, int *e1d, int *e2s, int *e2d,
                    int *seedNode, const float splitParams[3]) {
    int nFree = freeOffsets[C];
    double halfmono = splitParams[0];
    int childMon = (int)splitParams[2];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int Gs = r, Cs = freeList[rank];
        int M = monomerCount[Gs];
        int parentNewMon = M - childMon;
        int Mold = e1s[Gs], MoldSide = e1d[Gs];
        double Lold = (M + 1) * halfmono, Lpar = (parentNewMon + 1) * halfmono;
        double ux = uVec[Gs], uy = uVec[C + Gs], uz = uVec[2 * C + Gs];
        double shift = 0.5 * Lold - 0.5 * Lpar;
        coord[Gs] = (float)(coord[Gs] + shift * ux);
        coord[C + Gs] = (float)(coord[C + Gs] + shift * uy);
        coord[2 * C + Gs] = (float)(coord[2 * C + Gs] + shift * uz);
        monomerCount[Gs] = parentNewMon;
        monomerCount[Cs] = childMon;
        seedNode[Cs] = -1;
        e1s[Gs] = Cs; e1d[Gs] = 1;
        e2s[Cs] = Gs; e2d[Cs] = 0;
        e1s[Cs] = Mold;
        e1d[Cs] = Mold >= 0 ? MoldSide : -1;
        if (Mold >= 0) {
            if (MoldSide == 0) e1s[Mold] = Cs;
            else e2s[Mold] = Cs;
        }
    }
}

/* GrowthSystem.recomputeDrag (GrowthSystem.java:206-227): segLength from monomerCount, the min-length
 * clamp (chain ends: stdSegLength monomers, interior: minMonomerCt), slender-rod drag and diffusion */
void sbo_recompute_drag(int C, const int *monomerCount, float *segLength, const int *end1NbrSlot,
                        const int *end2NbrSlot, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff,
                        const float dragParams[9], int fires) {
    double halfmono = dragParams[0], aeta = dragParams[1], radius = dragParams[2], kT = dragParams[3];
    double aPar = dragParams[4], aOrth = dragParams[5], aTurn = dragParams[6];
    int stdSeg = (int)dragParams[7], minMon = (int)dragParams[8];
    const double PI = 3.141592653589793;
    for (int i = 0; i < C; i++) {
        if (fires == 0) continue;
        int mc = monomerCount[i];
        double length = (mc + 1) * halfmono;
        segLength[i] = (float)length;
        int atEnd = (end1NbrSlot[i] < 0) || (end2NbrSlot[i] < 0);
        double minLength = (atEnd ? stdSeg : minMon) * halfmono;
        double asIf = length < minLength ? minLength : length;
        double LM = 1.0e-6 * asIf, RM = radius * 1.0e-6;
        double logT = log(LM / (2.0 * RM));
        double bTGx = (2.0 * PI * aeta * LM) / (logT + aPar);
        double bTGy = (4.0 * PI * aeta * LM) / (logT + aOrth);
        double bRGx = 4.0 * PI * aeta * RM * RM * LM;
        double bRGy = (PI * aeta * (LM * LM * LM)) / (3.0 * (logT + aTurn));
        int iy = C + i, iz = 2 * C + i;
        bTransGam[i] = (float)bTGx; bTransGam[iy] = (float)bTGy; bTransGam[iz] = (float)bTGy;
        bRotGam[i] = (float)bRGx; bRotGam[iy] = (float)bRGy; bRotGam[iz] = (float)bRGy;
        if (bTransDiff) {
            bTransDiff[i] = (float)(kT / bTGx); bTransDiff[iy] = (float)(kT / bTGy); bTransDiff[iz] = (float)(kT / bTGy);
        }
        if (bRotDiff) {
            bRotDiff[i] = (float)(kT / bRGx); bRotDiff[iy] = (float)(kT / bRGy); bRotDiff[iz] = (float)(kT / bRGy);
        }
    }
}

/* ------------------------------------------------------------------ severing */
/* SeveringSystem.cofilinAccumulate (SeveringSystem.java:37-50): f_cof += (f_ADP - f_cof) p_cof */
void sbo_cofilin_accumulate(int C, const int *filState, const float *nucFrac, float *cofFrac,
                            const float cofilinParams[2], int fires) {
    float pCof = cofilinParams[0];
    for (int s = 0; s < C; s++) {
        if (fires == 0) continue;
        if (filState[s] < 0) continue;
        float fADP = nucFrac[2 * C + s];
        float fCof = cofFrac[s];
        cofFrac[s] = fCof + (fADP - fCof) * pCof;
    }
}

/* SeveringSystem.cofilinDissolve (SeveringSystem.java:59-73): a segment past cofilinRatio dissolves */
void sbo_cofilin_dissolve(int C, const int *filState, const int *monomerCount, const float *cofFrac,
                          int *severDeathFlag, int *severReturnedMon, const float cofilinParams[2], int fires) {
    float ratio = cofilinParams[1];
    for (int s = 0; s < C; s++) {
        severDeathFlag[s] = 0;
        severReturnedMon[s] = 0;
        if (fires == 0) continue;
        if (filState[s] < 0) continue;
        if (cofFrac[s] > ratio) {
            severDeathFlag[s] = 1;
            severReturnedMon[s] = monomerCount[s];
        }
    }
}

/* SeveringSystem.nucleateFreshCofilin (SeveringSystem.java:87-97): a born slot starts with no cofilin */
void sbo_nucleate_fresh_cofilin(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                float *cofFrac) {
    int nFree = freeOffsets[C];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        cofFrac[freeList[rank]] = 0.0f;
    }
}

/* ------------------------------------------------------------------ filament birth allocator */
/* CrossBridgeSystem.csrScan (CrossBridgeSystem.java:198-205): exclusive prefix, the counts cleared */
int sbo_csr_scan(int n, int *counts, int *offsets) {
    int acc = 0;
    for (int s = 0; s < n; s++) {
        offsets[s] = acc;
        acc += counts[s];
        counts[s] = 0;
    }
    offsets[n] = acc;
    return acc;
}

/* FullSystemDemoHarness.allocCpu (FullSystemDemoHarness.java:950-957): FilamentBirthSystem.freeFlags,
 * csrScan, freeScatter, csrScan of the requests, allocate (FilamentBirthSystem.java:41-86) */
void sbo_birth_alloc(int C, int K, int *filState, int *freeCount, int *freeOffsets, int *freeList, int *acceptFlag,
                     int *rankOffsets, const float *reqCoord, const float *reqUVec, const float *reqYVec, int RC,
                     float *coord, float *uVec, float *yVec, float *bts, float *brs, const float birthParams[2]) {
    for (int s = 0; s < C; s++) freeCount[s] = filState[s] < 0 ? 1 : 0;
    sbo_csr_scan(C, freeCount, freeOffsets);
    for (int s = 0; s < C; s++)
        if (filState[s] < 0) freeList[freeOffsets[s]] = s;
    sbo_csr_scan(K, acceptFlag, rankOffsets);
    int nFree = freeOffsets[C];
    float bt = birthParams[0], br = birthParams[1];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int slot = freeList[rank];
        coord[slot] = reqCoord[r]; coord[C + slot] = reqCoord[RC + r]; coord[2 * C + slot] = reqCoord[2 * RC + r];
        uVec[slot] = reqUVec[r]; uVec[C + slot] = reqUVec[RC + r]; uVec[2 * C + slot] = reqUVec[2 * RC + r];
        yVec[slot] = reqYVec[r]; yVec[C + slot] = reqYVec[RC + r]; yVec[2 * C + slot] = reqYVec[2 * RC + r];
        bts[slot] = bt;
        brs[slot] = br;
        filState[slot] = SBO_FIL_ACTIVE;
    }
}

/* ------------------------------------------------------------------ the actin pool */
/* ActinPool.take / put (ActinPool.java): concentration in uM, the integer ledger exact */
static void pool_take(sbo_turnover *t, int monomers) {
    t->conc -= monomers * t->uMPerMonomer;
    if (t->conc < 0.0) t->conc = 0.0;
    t->taken += monomers;
}

static void pool_put(sbo_turnover *t, int monomers) {
    t->conc += monomers * t->uMPerMonomer;
    t->returned += monomers;
}

/* ------------------------------------------------------------------ the biochem cadence */
/* FullSystemDemoHarness.cpuStep turnover block (FullSystemDemoHarness.java:650-684) on a cadence step:
 * GrowthStore.refreshRate, age, depoly(Proxy), csrScan + put, applyDeath, grow, csrScan, growthAtp,
 * depletePoolForGrows, cofilinAccumulate, cofilinDissolve, csrScan, applyDeath, put, markSplits,
 * allocCpu, splitWire, splitInheritNuc, nucleateFreshCofilin, recomputeDrag.  csrScan clears the flags it
 * sums, so growthAtp (reading grewFlag after its scan) writes nothing, as in the reference. */
void sbo_turnover_cadence(sbo_turnover *t, sbo_net *net, int32_t step) {
    const int C = t->C;
    float growParams[3] = {t->growParams[0], t->growParams[1], t->growParams[2]};
    growParams[0] = t->growthOn ? (float)(t->onRate * t->conc * t->biochemDeltaT) : 0.0f;
    sbo_aging_age(C, t->filState, t->nucFrac, t->agingParams, 1);
    sbo_depoly(C, t->filState, t->monomerCount, net->coord, net->uVec, net->end1NbrSlot, t->proxy ? t->nucFrac : NULL,
               t->returnedMon, t->deathFlag, t->depolyParams, t->depolyRateParams, step, t->seed, net->rng,
               net->stream, 1);
    int nRet = sbo_csr_scan(C, t->returnedMon, t->scanOffsets);
    int nDeath = 0;
    for (int s = 0; s < C; s++) nDeath += t->deathFlag[s];
    sbo_apply_death(C, t->filState, t->monomerCount, t->seedNode, net->end1NbrSlot, net->end1NbrSide,
                    net->end2NbrSlot, net->end2NbrSide, net->brownTransScale, net->brownRotScale, t->deathFlag);
    pool_put(t, nRet);
    sbo_grow(C, t->seedNode, t->monomerCount, net->coord, net->uVec, t->grewFlag, growParams, step, t->seed, net->rng,
             net->stream, 1);
    int nGrew = sbo_csr_scan(C, t->grewFlag, t->scanOffsets);
    sbo_aging_growth_atp(C, t->grewFlag, t->monomerCount, t->nucFrac);
    if (nGrew > 0) pool_take(t, nGrew);
    sbo_cofilin_accumulate(C, t->filState, t->nucFrac, t->cofFrac, t->cofilinParams, 1);
    sbo_cofilin_dissolve(C, t->filState, t->monomerCount, t->cofFrac, t->severDeathFlag, t->severReturnedMon,
                         t->cofilinParams, 1);
    int nSev = sbo_csr_scan(C, t->severReturnedMon, t->scanOffsets);
    int nSevered = 0;
    for (int s = 0; s < C; s++) nSevered += t->severDeathFlag[s];
    sbo_apply_death(C, t->filState, t->monomerCount, t->seedNode, net->end1NbrSlot, net->end1NbrSide,
                    net->end2NbrSlot, net->end2NbrSide, net->brownTransScale, net->brownRotScale, t->severDeathFlag);
    pool_put(t, nSev);
    sbo_mark_splits(C, t->seedNode, t->monomerCount, net->coord, net->uVec, net->yVec, t->acceptFlag, t->reqCoord,
                    t->reqUVec, t->reqYVec, C, t->splitParams, 1);
    sbo_birth_alloc(C, C, t->filState, t->freeCount, t->freeOffsets, t->freeList, t->acceptFlag, t->rankOffsets,
                    t->reqCoord, t->reqUVec, t->reqYVec, C, net->coord, net->uVec, net->yVec, net->brownTransScale,
                    net->brownRotScale, t->birthParams);
    sbo_split_wire(C, C, t->rankOffsets, t->freeList, t->freeOffsets, t->monomerCount, net->coord, net->uVec,
                   net->end1NbrSlot, net->end1NbrSide, net->end2NbrSlot, net->end2NbrSide, t->seedNode, t->splitParams);
    sbo_aging_split_inherit_nuc(C, C, t->rankOffsets, t->freeList, t->freeOffsets, t->nucFrac);
    sbo_nucleate_fresh_cofilin(C, C, t->rankOffsets, t->freeList, t->freeOffsets, t->cofFrac);
    sbo_recompute_drag(C, t->monomerCount, net->segLength, net->end1NbrSlot, net->end2NbrSlot, net->bTransGam,
                       net->bRotGam, t->bTransDiff, t->bRotDiff, t->dragParams, 1);
    int nAcc = t->rankOffsets[C], nFree = t->freeOffsets[C];
    t->last[0] = nGrew;
    t->last[1] = nRet;
    t->last[2] = nSev;
    t->last[3] = nDeath;
    t->last[4] = nSevered;
    t->last[5] = nAcc < nFree ? nAcc : nFree;
    t->cadences += 1;
}
