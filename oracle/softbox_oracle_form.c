/*
 * softbox_oracle_form.c -- CPU restatement of SoftBox's crosslinker formation (test infrastructure).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see softbox_oracle.h).
 *
 * Follows V2OneXHarness.formationCpu (V2OneXHarness.java:444-464):
 *   FilIDSystem.init + jump rounds (FilIDSystem.java:31-60)      -> filID (chain terminal per segment)
 *   countActiveLinks (CrosslinkerSystem.java:694-706)            -> start-of-formation saturation counts
 *   FormationGrid.buildCpu (FormationGrid.java:87-97)            -> a grid of segment centres
 *   gridFormCount / gridFormScan / gridFormEmit (:461-604)       -> candidates (i, j > i) in (i, j) order
 *   formGates (:607-676) -> formAdmitReduce (:680-691) -> formAdmit (:713-738)
 *   freeFlags / csrScan / freeScatter / csrScan / allocate / placeOrient (:298-357, :742-760)
 * Periodic boxes (no counterpart in the reference): cells wrap, centre distances take the minimum image
 * in double, and the gates see segment b shifted by the image offset of its centre.  Box mode is the
 * reference verbatim.  The random draws use the network's stream (Wang: the reference keying by the
 * candidate index, salts XLJ1 / XLJ2 / XLFP; or Philox).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "softbox_oracle.h"

static const double PI_F = 3.141592653589793;
#define XLJ1 0x584C4A31u
#define XLJ2 0x584C4A32u
#define XLFP 0x584C4650u

/* ------------------------------------------------------------------ FilIDSystem */
/* FilIDSystem.rounds, FilIDSystem.java:31-34: ceil(log2 n) rounded up to even */
int sbo_filid_rounds(int n) {
    int k = 0;
    if (n <= 1) k = 1;
    else {
        unsigned v = (unsigned)(n - 1);
        while (v) { k++; v >>= 1; }
    }
    return (k % 2 == 0) ? k : k + 1;
}

/* FilIDSystem.init, FilIDSystem.java:38-47 (filState NULL: every slot ACTIVE) */
void sbo_filid_init(int n, const int *filState, const int *end2NbrSlot, int *ptr) {
    for (int seg = 0; seg < n; seg++) {
        if (filState && filState[seg] < 0) { ptr[seg] = -seg - 2; continue; }
        int nx = end2NbrSlot[seg];
        ptr[seg] = (nx >= 0 && nx != seg && (!filState || filState[nx] >= 0)) ? nx : seg;
    }
}

/* FilIDSystem.jump, FilIDSystem.java:52-59 */
void sbo_filid_jump(int n, const int *ptrIn, int *ptrOut, const int *filState) {
    for (int seg = 0; seg < n; seg++) {
        if (filState && filState[seg] < 0) { ptrOut[seg] = ptrIn[seg]; continue; }
        ptrOut[seg] = ptrIn[ptrIn[seg]];
    }
}

/* V2OneXHarness.computeFilID (V2OneXHarness.java:466-470): init, then rounds(n) ping-pong jumps (even,
 * so the result lands back in filID) */
void sbo_filid(int n, const int *filState, const int *end2NbrSlot, int *filID, int *scratch) {
    sbo_filid_init(n, filState, end2NbrSlot, filID);
    int *a = filID, *b = scratch;
    int rounds = sbo_filid_rounds(n);
    for (int k = 0; k < rounds; k++) {
        sbo_filid_jump(n, a, b, filState);
        int *t = a; a = b; b = t;
    }
}

/* ------------------------------------------------------------------ candidate predicate */
static double min_image_d(double d, double L) { return d - L * rint(d / L); }

/* centre distance and the coarse capsule bound of CrosslinkerSystem.filFilCandidates (:413-423) */
static int form_pair(int nSeg, int i, int j, const float *coord, const float *segLength, double grab,
                     int periodic, const float *boxLen) {
    double cix = coord[i], ciy = coord[nSeg + i], ciz = coord[2 * nSeg + i];
    double hi = 0.5 * segLength[i];
    double dx = cix - coord[j], dy = ciy - coord[nSeg + j], dz = ciz - coord[2 * nSeg + j];
    if (periodic) {
        dx = min_image_d(dx, boxLen[0]);
        dy = min_image_d(dy, boxLen[1]);
        dz = min_image_d(dz, boxLen[2]);
    }
    double cd = sqrt(dx * dx + dy * dy + dz * dz);
    double bound = hi + 0.5 * segLength[j] + grab;
    return !(cd > bound);
}

/* CrosslinkerSystem.filFilCandidates, CrosslinkerSystem.java:405-431 (the O(N^2) reference) */
void sbo_fil_fil_candidates(int nSeg, const float *coord, const float *segLength, const int *filID, int reqCap,
                            int *reqFilA, int *reqFilB, const float *formParams, int *formCounts, int periodic,
                            const float *boxLen) {
    double grab = formParams[6];
    int c = 0;
    for (int i = 0; i < nSeg; i++) {
        for (int j = i + 1; j < nSeg; j++) {
            if (filID[i] == filID[j]) continue;
            if (!form_pair(nSeg, i, j, coord, segLength, grab, periodic, boxLen)) continue;
            if (c < reqCap) { reqFilA[c] = i; reqFilB[c] = j; c++; }
        }
    }
    formCounts[0] = c;
    for (int k = c; k < reqCap; k++) { reqFilA[k] = -1; reqFilB[k] = -1; }
}

/* the cell index of one axis as gridFormCount computes it (:472-478): Java (int) truncation of the
 * float, then the clamp to [0, nBins-1] (a value below 0, NaN included, lands on 0); periodic: floor,
 * then wrapped into [0, nBins) */
static int form_axis(float v, int nBins, int periodic) {
    if (periodic) {
        if (!(v == v)) return 0;
        double m = fmod((double)floorf(v), (double)nBins);
        if (m < 0) m += nBins;
        int c = (int)m;
        return c >= nBins ? 0 : c;
    }
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)nBins) return nBins - 1;
    int c = (int)v;
    return c >= nBins ? nBins - 1 : c;
}

static void form_cell(float v0, float v1, float v2, const float *gridParams, const int *gridDims, int periodic,
                      int *cx, int *cy, int *cz) {
    *cx = form_axis((v0 - gridParams[0]) * gridParams[4], gridDims[0], periodic);
    *cy = form_axis((v1 - gridParams[1]) * gridParams[4], gridDims[1], periodic);
    *cz = form_axis((v2 - gridParams[2]) * gridParams[4], gridDims[2], periodic);
}

typedef void (*form_visit)(void *ctx, int j);

/* the 27-cell walk of gridFormCount / gridFormEmit (:479-513, :565-596): every body j of the stencil */
static void form_walk(int i, const float *coord, int nSeg, const float *gridParams, const int *gridDims,
                      const int *offsets, const int *contents, int periodic, form_visit visit, void *ctx) {
    int nX = gridDims[0], nY = gridDims[1], nZ = gridDims[2], nXY = nX * nY;
    int cx, cy, cz;
    form_cell(coord[i], coord[nSeg + i], coord[2 * nSeg + i], gridParams, gridDims, periodic, &cx, &cy, &cz);
    int x0 = cx - 1, x1 = cx + 1, y0 = cy - 1, y1 = cy + 1, z0 = cz - 1, z1 = cz + 1;
    if (!periodic) {
        if (x0 < 0) x0 = 0;
        if (x1 >= nX) x1 = nX - 1;
        if (y0 < 0) y0 = 0;
        if (y1 >= nY) y1 = nY - 1;
        if (z0 < 0) z0 = 0;
        if (z1 >= nZ) z1 = nZ - 1;
    }
    for (int zz = z0; zz <= z1; zz++) {
        int wz = periodic ? (zz + nZ) % nZ : zz;
        for (int yy = y0; yy <= y1; yy++) {
            int wy = periodic ? (yy + nY) % nY : yy;
            for (int xx = x0; xx <= x1; xx++) {
                int wx = periodic ? (xx + nX) % nX : xx;
                int cc = wx + wy * nX + wz * nXY;
                for (int idx = offsets[cc]; idx < offsets[cc + 1]; idx++) visit(ctx, contents[idx]);
            }
        }
    }
}

typedef struct {
    int i, nSeg, periodic, cnt, cap, base;
    const float *coord, *segLength, *boxLen;
    const int *filID;
    double grab;
    int *reqFilB;
} form_walk_ctx;

static int form_qualifies(form_walk_ctx *w, int j) {
    if (j <= w->i) return 0;
    if (w->filID[j] == w->filID[w->i]) return 0;
    return form_pair(w->nSeg, w->i, j, w->coord, w->segLength, w->grab, w->periodic, w->boxLen);
}

static void visit_count(void *ctx, int j) {
    form_walk_ctx *w = (form_walk_ctx *)ctx;
    if (form_qualifies(w, j)) w->cnt++;
}

/* insertion of j into the region, ascending, keeping the smallest `cap` (gridFormEmit :583-593) */
static void visit_emit(void *ctx, int j) {
    form_walk_ctx *w = (form_walk_ctx *)ctx;
    if (!form_qualifies(w, j)) return;
    int *r = w->reqFilB;
    if (w->cnt < w->cap) {
        int p = w->base + w->cnt;
        while (p > w->base && r[p - 1] > j) { r[p] = r[p - 1]; p--; }
        r[p] = j;
        w->cnt++;
    } else if (j < r[w->base + w->cap - 1]) {
        int p = w->base + w->cap - 1;
        while (p > w->base && r[p - 1] > j) { r[p] = r[p - 1]; p--; }
        r[p] = j;
    }
}

/* CrosslinkerSystem.gridFormCount, CrosslinkerSystem.java:461-513 */
void sbo_grid_form_count(int nSeg, const float *coord, const float *segLength, const int *filID,
                         const float *gridParams, const int *gridDims, const int *offsets, const int *contents,
                         int *candCountSeg, const float *formParams, int periodic, const float *boxLen) {
    for (int i = 0; i < nSeg; i++) {
        form_walk_ctx w = {i, nSeg, periodic, 0, 0, 0, coord, segLength, boxLen, filID, formParams[6], NULL};
        form_walk(i, coord, nSeg, gridParams, gridDims, offsets, contents, periodic, visit_count, &w);
        candCountSeg[i] = w.cnt < SBO_FORM_MAXC ? w.cnt : SBO_FORM_MAXC;
    }
}

/* CrosslinkerSystem.gridFormScan, CrosslinkerSystem.java:516-529 */
void sbo_grid_form_scan(int nSeg, const int *candCountSeg, int *candBaseSeg, int reqCap, int *reqFilA, int *reqFilB,
                        int *formCounts) {
    int acc = 0;
    for (int s = 0; s < nSeg; s++) { candBaseSeg[s] = acc; acc += candCountSeg[s]; }
    candBaseSeg[nSeg] = acc;
    int total = acc < reqCap ? acc : reqCap;
    formCounts[0] = total;
    for (int k = total; k < reqCap; k++) { reqFilA[k] = -1; reqFilB[k] = -1; }
}

/* CrosslinkerSystem.gridFormEmit, CrosslinkerSystem.java:535-604 */
void sbo_grid_form_emit(int nSeg, const float *coord, const float *segLength, const int *filID,
                        const float *gridParams, const int *gridDims, const int *offsets, const int *contents,
                        const int *candBaseSeg, const int *candCountSeg, int reqCap, int *reqFilA, int *reqFilB,
                        const float *formParams, int periodic, const float *boxLen) {
    for (int i = 0; i < nSeg; i++) {
        int base = candBaseSeg[i];
        int cap = candCountSeg[i];
        if (base >= reqCap) cap = 0;
        else if (base + cap > reqCap) cap = reqCap - base;
        if (cap <= 0) continue;
        form_walk_ctx w = {i, nSeg, periodic, 0, cap, base, coord, segLength, boxLen, filID, formParams[6], reqFilB};
        form_walk(i, coord, nSeg, gridParams, gridDims, offsets, contents, periodic, visit_emit, &w);
        for (int k = 0; k < cap; k++) reqFilA[base + k] = i;
    }
}

/* CrosslinkerSystem.fastAcos, CrosslinkerSystem.java:372-376 */
static double form_fast_acos(double dot) {
    if (dot > 0.95) { double t = 1.0 - dot; if (t < 0) t = 0; return sqrt(2.0 * t); }
    if (dot < -0.95) { double t = 1.0 + dot; if (t < 0) t = 0; return PI_F - sqrt(2.0 * t); }
    return sbo_accurate_acos(dot);
}

/* CrosslinkerSystem.formGates, CrosslinkerSystem.java:607-676 (candidates c < nCand; periodic: b shifted
 * by the image offset of its centre relative to a's) */
void sbo_form_gates(int nSeg, int nCand, const float *coord, const float *uVec, const float *end1, const float *end2,
                    const float *segLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
                    int *reqOrient, int *gatePass, const float *formParams, int32_t step, int32_t seed, int mode,
                    int rng, uint32_t stream, int periodic, const float *boxLen) {
    double maxAngle = formParams[0], grabSq = formParams[1];
    double pForm = formParams[4], jit = formParams[5];
    for (int c = 0; c < nCand; c++) {
        gatePass[c] = 0;
        reqOrient[c] = 0;
        int a = reqFilA[c], b = reqFilB[c];
        if (a < 0 || b < 0) continue;
        double uax = uVec[a], uay = uVec[nSeg + a], uaz = uVec[2 * nSeg + a];
        double ubx = uVec[b], uby = uVec[nSeg + b], ubz = uVec[2 * nSeg + b];
        double dot = uax * ubx + uay * uby + uaz * ubz;
        double angT = form_fast_acos(dot), angTR = form_fast_acos(-dot);
        int alignOk;
        if (mode == 1) alignOk = angT <= maxAngle;
        else if (mode == -1) alignOk = angTR <= maxAngle;
        else alignOk = (angT <= maxAngle) || (angTR <= maxAngle);
        if (!alignOk) continue;
        reqOrient[c] = dot >= 0.0 ? 1 : 0;
        double e1ax = end1[a], e1ay = end1[nSeg + a], e1az = end1[2 * nSeg + a];
        double e2ax = end2[a], e2ay = end2[nSeg + a], e2az = end2[2 * nSeg + a];
        double e1bx = end1[b], e1by = end1[nSeg + b], e1bz = end1[2 * nSeg + b];
        double e2bx = end2[b], e2by = end2[nSeg + b], e2bz = end2[2 * nSeg + b];
        if (periodic) {
            double d[3];
            for (int k = 0; k < 3; k++) {
                double raw = (double)coord[k * nSeg + b] - (double)coord[k * nSeg + a];
                d[k] = min_image_d(raw, boxLen[k]) - raw;
            }
            e1bx += d[0]; e1by += d[1]; e1bz += d[2];
            e2bx += d[0]; e2by += d[1]; e2bz += d[2];
        }
        double r1x = e2ax - e1ax, r1y = e2ay - e1ay, r1z = e2az - e1az;
        double r2x = e2bx - e1bx, r2y = e2by - e1by, r2z = e2bz - e1bz;
        double r3x = e1bx - e1ax, r3y = e1by - e1ay, r3z = e1bz - e1az;
        double r4x = r1y * r2z - r1z * r2y, r4y = r1z * r2x - r1x * r2z, r4z = r1x * r2y - r1y * r2x;
        double smallNum = 1e-20;
        if (r4x < smallNum && r4y < smallNum && r4z < smallNum) continue;
        double denom = r4x * r4x + r4y * r4y + r4z * r4z;
        double c32x = r3y * r2z - r3z * r2y, c32y = r3z * r2x - r3x * r2z, c32z = r3x * r2y - r3y * r2x;
        double alpha = (r4x * c32x + r4y * c32y + r4z * c32z) / denom;
        if (alpha < 0.0 || alpha > 1.0) continue;
        double c31x = r3y * r1z - r3z * r1y, c31y = r3z * r1x - r3x * r1z, c31z = r3x * r1y - r3y * r1x;
        double beta = (r4x * c31x + r4y * c31y + r4z * c31z) / denom;
        if (beta < 0.0 || beta > 1.0) continue;
        double p1x = e1ax + alpha * r1x, p1y = e1ay + alpha * r1y, p1z = e1az + alpha * r1z;
        double p2x = e1bx + beta * r2x, p2y = e1by + beta * r2y, p2z = e1bz + beta * r2z;
        double ddx = p1x - p2x, ddy = p1y - p2y, ddz = p1z - p2z;
        double conDistSq = ddx * ddx + ddy * ddy + ddz * ddz;
        if (conDistSq >= grabSq) continue;
        double da = sqrt((p1x - e1ax) * (p1x - e1ax) + (p1y - e1ay) * (p1y - e1ay) + (p1z - e1az) * (p1z - e1az));
        double db = sqrt((p2x - e1bx) * (p2x - e1bx) + (p2y - e1by) * (p2y - e1by) + (p2z - e1bz) * (p2z - e1bz));
        float uj1 = sbo_uniform(sbo_event_bits(rng, (uint32_t)c, step, seed, stream, XLJ1));
        float uj2 = sbo_uniform(sbo_event_bits(rng, (uint32_t)c, step, seed, stream, XLJ2));
        double loc1 = da + (2.0 * uj1 - 1.0) * jit;
        double loc2 = db + (2.0 * uj2 - 1.0) * jit;
        double lenA = segLength[a], lenB = segLength[b];
        if (loc1 > lenA) loc1 = lenA;
        if (loc1 < 0.0) loc1 = 0.0;
        if (loc2 > lenB) loc2 = lenB;
        if (loc2 < 0.0) loc2 = 0.0;
        reqLoc1[c] = (float)loc1;
        reqLoc2[c] = (float)loc2;
        float up = sbo_uniform(sbo_event_bits(rng, (uint32_t)c, step, seed, stream, XLFP));
        if (up < pForm) gatePass[c] = 1;
    }
}

/* CrosslinkerSystem.formAdmitReduce, CrosslinkerSystem.java:680-691 (sentinel reqCap) */
void sbo_form_admit_reduce(int nCand, int reqCap, int nSeg, const int *reqFilA, const int *reqFilB,
                           const int *gatePass, int *minCand) {
    for (int s = 0; s < nSeg; s++) minCand[s] = reqCap;
    for (int c = 0; c < nCand; c++) {
        if (gatePass[c] == 0) continue;
        int a = reqFilA[c], b = reqFilB[c];
        if (c < minCand[a]) minCand[a] = c;
        if (c < minCand[b]) minCand[b] = c;
    }
}

/* CrosslinkerSystem.formAdmit, CrosslinkerSystem.java:713-738 */
void sbo_form_admit(int nCand, int C, const int *reqFilA, const int *reqFilB, const float *reqLoc1,
                    const float *reqLoc2, const int *gatePass, const int *minCand, const int *activeLinkCount,
                    const int *linkState, const int *linkFilA, const int *linkFilB, const float *loc1,
                    const float *loc2, int *acceptFlag, const float *formParams) {
    double minSep = formParams[2];
    int maxLinks = (int)formParams[3];
    for (int c = 0; c < nCand; c++) {
        acceptFlag[c] = 0;
        if (gatePass[c] == 0) continue;
        int segA = reqFilA[c], segB = reqFilB[c];
        if (minCand[segA] != c || minCand[segB] != c) continue;
        if (activeLinkCount[segA] >= maxLinks || activeLinkCount[segB] >= maxLinks) continue;
        double la = reqLoc1[c], lb = reqLoc2[c];
        int ok = 1;
        for (int L = 0; L < C; L++) {
            if (linkState[L] < 0) continue;
            int fa = linkFilA[L], fb = linkFilB[L];
            if (fa == segA && fabs(la - loc1[L]) < minSep) ok = 0;
            if (fb == segA && fabs(la - loc2[L]) < minSep) ok = 0;
            if (fa == segB && fabs(lb - loc1[L]) < minSep) ok = 0;
            if (fb == segB && fabs(lb - loc2[L]) < minSep) ok = 0;
        }
        if (ok) acceptFlag[c] = 1;
    }
}

/* CrossBridgeSystem.csrScan, CrossBridgeSystem.java:198-205 (the exclusive scan; the cursor reset of
 * the count array is not needed by the callers here) */
static void scan_excl(int n, const int *in, int *out) {
    int acc = 0;
    for (int s = 0; s < n; s++) { out[s] = acc; acc += in[s]; }
    out[n] = acc;
}

/* CrosslinkerSystem.freeFlags + csrScan + freeScatter + csrScan(rank) + allocate + placeOrient
 * (CrosslinkerSystem.java:298-357, 742-760): accepted request r with dense rank < nFree claims
 * freeList[rank]; K requests (the reference sets K = reqCap; requests at or past nCand are never
 * accepted, so K = nCand gives the same ranks).  Returns the number of links formed. */
int sbo_form_allocate(int K, int C, const int *acceptFlag, const int *reqFilA, const int *reqFilB,
                      const float *reqLoc1, const float *reqLoc2, const int *reqOrient, int *freeCount,
                      int *freeOffsets, int *freeList, int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1,
                      float *loc2, int *linkState, float *strainHist, int *strainPlace, int *linkOrientSame,
                      float *torqueMagHist, int *torqueMagPlace) {
    for (int s = 0; s < C; s++) freeCount[s] = linkState[s] < 0 ? 1 : 0;
    scan_excl(C, freeCount, freeOffsets);
    for (int s = 0; s < C; s++)
        if (linkState[s] < 0) freeList[freeOffsets[s]] = s;
    scan_excl(K, acceptFlag, rankOffsets);
    int nFree = freeOffsets[C];
    int formed = 0;
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int slot = freeList[rank];
        linkFilA[slot] = reqFilA[r];
        linkFilB[slot] = reqFilB[r];
        loc1[slot] = reqLoc1[r];
        loc2[slot] = reqLoc2[r];
        for (int j = 0; j < SBO_STRAIN_WIN; j++) strainHist[slot * SBO_STRAIN_WIN + j] = 0.0f;
        strainPlace[slot] = 0;
        linkState[slot] = 0;
        linkOrientSame[slot] = reqOrient[r];
        torqueMagPlace[slot] = 0;
        for (int j = 0; j < SBO_TORQUE_WIN; j++) torqueMagHist[slot * SBO_TORQUE_WIN + j] = 0.0f;
        formed++;
    }
    return formed;
}

/* the whole formation pass (V2OneXHarness.formationCpu, V2OneXHarness.java:444-464) */
int sbo_formation(sbo_form *f, int32_t step, int32_t seed, int rng, uint32_t stream) {
    int n = f->nSeg;
    sbo_filid(n, f->filState, f->end2NbrSlot, f->filID, f->filIDScratch);
    sbo_count_active_links(f->C, n, f->linkState, f->linkFilA, f->linkFilB, f->activeLinkCount);
    /* FormationGrid.buildCpu: segment centres into the formation grid (publish + bodyCell + sort) */
    for (int a = 0; a < 3; a++) memcpy(f->center + (size_t)a * n, f->coord + (size_t)a * n, sizeof(float) * n);
    sbo_body_cell(n, n, f->center, f->gridParams, f->gridDims, f->periodic, f->bodyCell);
    sbo_grid_build_serial(n, n, f->bodyCell, f->gridDims, f->cellCount, f->gridCellOffsets, f->gridCellContents);
    sbo_grid_form_count(n, f->coord, f->segLength, f->filID, f->gridParams, f->gridDims, f->gridCellOffsets,
                        f->gridCellContents, f->candCountSeg, f->formParams, f->periodic, f->boxLen);
    sbo_grid_form_scan(n, f->candCountSeg, f->candBaseSeg, f->reqCap, f->reqFilA, f->reqFilB, f->formCounts);
    sbo_grid_form_emit(n, f->coord, f->segLength, f->filID, f->gridParams, f->gridDims, f->gridCellOffsets,
                       f->gridCellContents, f->candBaseSeg, f->candCountSeg, f->reqCap, f->reqFilA, f->reqFilB,
                       f->formParams, f->periodic, f->boxLen);
    int nCand = f->formCounts[0];
    sbo_form_gates(n, nCand, f->coord, f->uVec, f->end1, f->end2, f->segLength, f->reqFilA, f->reqFilB, f->reqLoc1,
                   f->reqLoc2, f->reqOrient, f->gatePass, f->formParams, step, seed, f->mode, rng, stream, f->periodic,
                   f->boxLen);
    sbo_form_admit_reduce(nCand, f->reqCap, n, f->reqFilA, f->reqFilB, f->gatePass, f->minCand);
    sbo_form_admit(nCand, f->C, f->reqFilA, f->reqFilB, f->reqLoc1, f->reqLoc2, f->gatePass, f->minCand,
                   f->activeLinkCount, f->linkState, f->linkFilA, f->linkFilB, f->loc1, f->loc2, f->acceptFlag,
                   f->formParams);
    return sbo_form_allocate(nCand, f->C, f->acceptFlag, f->reqFilA, f->reqFilB, f->reqLoc1, f->reqLoc2,
                             f->reqOrient, f->freeCount, f->freeOffsets, f->freeList, f->rankOffsets, f->linkFilA,
                             f->linkFilB, f->loc1, f->loc2, f->linkState, f->strainHist, f->strainPlace,
                             f->linkOrientSame, f->torqueMagHist, f->torqueMagPlace);
}
