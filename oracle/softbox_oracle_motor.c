/*
 * softbox_oracle_motor.c -- CPU restatement of the myosin side of SoftBox's mechanics hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see softbox_oracle.h.  Same rules as softbox_oracle.c:
 * the reference's float/double types and operation order, -ffp-contract=off.
 *
 * Motor body layout: sub-body 3m = rod, 3m+1 = lever, 3m+2 = head (MotorStore.java:45-60), planar
 * SoA with stride nB = 3*nMotors.  Per-motor arrays (head, uVec, rodUVec) have stride nMotors.
 */
#include <math.h>
#include <string.h>

#include "softbox_oracle.h"

#define PI_D 3.141592653589793
#define NUC_NONE 0
#define NUC_ATP 1
#define NUC_ADPPI 2
#define NUC_ADP 3
#define FREE_BINDABLE (-1)
#define FREE_COOLDOWN (-2)
#define XB_STRIDE 13
#define MINI_STRIDE 6

/* MotorStore.publishHeadFromBody, MotorStore.java:300-314 */
void sbo_publish_head_from_body(int nM, const float *bodyCoord, const float *bodyUVec, const float *bodySegLength,
                                float *head, float *uVec, float *rodUVec) {
    int nB = 3 * nM;
    for (int m = 0; m < nM; m++) {
        int h = 3 * m + 2, r = 3 * m;
        float hl = bodySegLength[h];
        head[m] = bodyCoord[h] + 0.5f * hl * bodyUVec[h];
        head[nM + m] = bodyCoord[nB + h] + 0.5f * hl * bodyUVec[nB + h];
        head[2 * nM + m] = bodyCoord[2 * nB + h] + 0.5f * hl * bodyUVec[2 * nB + h];
        uVec[m] = bodyUVec[h]; uVec[nM + m] = bodyUVec[nB + h]; uVec[2 * nM + m] = bodyUVec[2 * nB + h];
        rodUVec[m] = bodyUVec[r]; rodUVec[nM + m] = bodyUVec[nB + r]; rodUVec[2 * nM + m] = bodyUVec[2 * nB + r];
    }
}

/* MotorStore.publishToBodyView, MotorStore.java:326-347 (ownerStore 1 = motor) */
void sbo_publish_motors(int nM, int cap, int base, const float *head, const float *reach, float *center,
                        float *boundingRadius, int *ownerStore, int *ownerSlot) {
    for (int m = 0; m < nM; m++) {
        int b = base + m;
        center[b] = head[m];
        center[cap + b] = head[nM + m];
        center[2 * cap + b] = head[2 * nM + m];
        boundingRadius[b] = reach[m];
        ownerStore[b] = 1;
        ownerSlot[b] = m;
    }
}

/* BindingDetectionSystem.reachTestDistSq, BindingDetectionSystem.java:59-85: conDist^2 or -1 */
static float reach_test(float mx, float my, float mz, float mux, float muy, float muz, float rux, float ruy, float ruz,
                        float e1x, float e1y, float e1z, float e2x, float e2y, float e2z, float myoColTol,
                        float alignTol) {
    float r1x = e2x - e1x, r1y = e2y - e1y, r1z = e2z - e1z;
    float denom = r1x * r1x + r1y * r1y + r1z * r1z;
    if (denom <= 0.0f) return -1.0f;
    float r2x = mx - e1x, r2y = my - e1y, r2z = mz - e1z;
    float numer = r2x * r1x + r2y * r1y + r2z * r1z;
    float alpha = numer / denom;
    if (alpha < 0.0f || alpha > 1.0f) return -1.0f;
    float cpx = e1x + alpha * r1x, cpy = e1y + alpha * r1y, cpz = e1z + alpha * r1z;
    float dx = cpx - mx, dy = cpy - my, dz = cpz - mz;
    float conDistSq = dx * dx + dy * dy + dz * dz;
    if (conDistSq >= myoColTol * myoColTol) return -1.0f;
    float inv = 1.0f / (float)sqrt((double)denom);
    float fux = r1x * inv, fuy = r1y * inv, fuz = r1z * inv;
    float motDotFil = mux * fux + muy * fuy + muz * fuz;
    if (motDotFil < alignTol) return -1.0f;
    float rodDotFil = rux * fux + ruy * fuy + ruz * fuz;
    if (rodDotFil < 0.0f) return -1.0f;
    return conDistSq;
}

static int clamp_bin(float v, int n) {
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)n) return n - 1;
    int c = (int)v;
    return c >= n ? n - 1 : c;
}

/* BindingDetectionSystem.gridReachable, BindingDetectionSystem.java:143-228 (the fused per-motor grid
 * query; the reachTestDistSq predicate inlined there is identical to reach_test) */
void sbo_grid_reachable(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                        const float *segEnd1, const float *segEnd2, const float *gridParams, const int *gridDims,
                        const int *gridCellOffsets, const int *gridCellContents, const int *ownerStore,
                        const int *ownerSlot, int maxCand, int *reachSeg, int *reachCount, const float *kinParams) {
    float myoColTol = kinParams[7], alignTol = kinParams[8];
    float xMin = gridParams[0], yMin = gridParams[1], zMin = gridParams[2], invCell = gridParams[4];
    int nX = gridDims[0], nY = gridDims[1], nZ = gridDims[2], nXY = nX * nY;
    for (int m = 0; m < nM; m++) {
        float mx = head[m], my = head[nM + m], mz = head[2 * nM + m];
        float mux = uVec[m], muy = uVec[nM + m], muz = uVec[2 * nM + m];
        float rux = rodUVec[m], ruy = rodUVec[nM + m], ruz = rodUVec[2 * nM + m];
        int cx = clamp_bin((mx - xMin) * invCell, nX);
        int cy = clamp_bin((my - yMin) * invCell, nY);
        int cz = clamp_bin((mz - zMin) * invCell, nZ);
        int x0 = cx - 1 < 0 ? 0 : cx - 1, x1 = cx + 1 >= nX ? nX - 1 : cx + 1;
        int y0 = cy - 1 < 0 ? 0 : cy - 1, y1 = cy + 1 >= nY ? nY - 1 : cy + 1;
        int z0 = cz - 1 < 0 ? 0 : cz - 1, z1 = cz + 1 >= nZ ? nZ - 1 : cz + 1;
        int out = 0;
        for (int zz = z0; zz <= z1; zz++)
            for (int yy = y0; yy <= y1; yy++)
                for (int xx = x0; xx <= x1; xx++) {
                    int cc = xx + yy * nX + zz * nXY;
                    for (int idx = gridCellOffsets[cc]; idx < gridCellOffsets[cc + 1]; idx++) {
                        int j = gridCellContents[idx];
                        if (ownerStore[j] != 0) continue;
                        int s = ownerSlot[j];
                        float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, segEnd1[s], segEnd1[nSeg + s],
                                             segEnd1[2 * nSeg + s], segEnd2[s], segEnd2[nSeg + s],
                                             segEnd2[2 * nSeg + s], myoColTol, alignTol);
                        if (d >= 0.0f) {
                            if (out < maxCand) reachSeg[(size_t)m * maxCand + out] = s;
                            out++;
                        }
                    }
                }
        reachCount[m] = out;
    }
}

/* BindingDetectionSystem.bruteReachable, BindingDetectionSystem.java:260-283 */
static void brute_reachable(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                            const float *segEnd1, const float *segEnd2, int maxCand, int *reachSeg, int *reachCount,
                            const int *seedNode, const float *kinParams) {
    float myoColTol = kinParams[7], alignTol = kinParams[8];
    for (int m = 0; m < nM; m++) {
        int out = 0;
        for (int s = 0; s < nSeg; s++) {
            if (seedNode && seedNode[s] >= 0) continue;   /* node-held tip (BindingDetectionSystem.java:426) */
            float d = reach_test(head[m], head[nM + m], head[2 * nM + m], uVec[m], uVec[nM + m], uVec[2 * nM + m],
                                 rodUVec[m], rodUVec[nM + m], rodUVec[2 * nM + m], segEnd1[s], segEnd1[nSeg + s],
                                 segEnd1[2 * nSeg + s], segEnd2[s], segEnd2[nSeg + s], segEnd2[2 * nSeg + s],
                                 myoColTol, alignTol);
            if (d >= 0.0f) {
                if (out < maxCand) reachSeg[(size_t)m * maxCand + out] = s;
                out++;
            }
        }
        reachCount[m] = out;
    }
}

void sbo_brute_reachable(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                         const float *segEnd1, const float *segEnd2, int maxCand, int *reachSeg, int *reachCount,
                         const float *kinParams) {
    brute_reachable(nM, nSeg, head, uVec, rodUVec, segEnd1, segEnd2, maxCand, reachSeg, reachCount, NULL, kinParams);
}

/* BindingDetectionSystem.bruteReachableNodeAware (BindingDetectionSystem.java:410-436): segments held by a
 * node (seedNode >= 0) are never candidates */
void sbo_brute_reachable_node_aware(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                                    const float *segEnd1, const float *segEnd2, int maxCand, int *reachSeg,
                                    int *reachCount, const int *seedNode, const float *kinParams) {
    brute_reachable(nM, nSeg, head, uVec, rodUVec, segEnd1, segEnd2, maxCand, reachSeg, reachCount, seedNode,
                    kinParams);
}

static void bind_nearest(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                         const float *segEnd1, const float *segEnd2, const int *reachSeg, const int *reachCount,
                         int maxCand, int *boundSeg, float *bindArc, const int *seedNode, const float *kinParams);

/* BindingDetectionSystem.bindNearest, BindingDetectionSystem.java:356-389 */
void sbo_bind_nearest(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                      const float *segEnd1, const float *segEnd2, const int *reachSeg, const int *reachCount,
                      int maxCand, int *boundSeg, float *bindArc, const float *kinParams) {
    bind_nearest(nM, nSeg, head, uVec, rodUVec, segEnd1, segEnd2, reachSeg, reachCount, maxCand, boundSeg, bindArc,
                 NULL, kinParams);
}

/* BindingDetectionSystem.bindNearestNodeAware (BindingDetectionSystem.java:440-473) */
void sbo_bind_nearest_node_aware(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                                 const float *segEnd1, const float *segEnd2, const int *reachSeg,
                                 const int *reachCount, int maxCand, int *boundSeg, float *bindArc,
                                 const int *seedNode, const float *kinParams) {
    bind_nearest(nM, nSeg, head, uVec, rodUVec, segEnd1, segEnd2, reachSeg, reachCount, maxCand, boundSeg, bindArc,
                 seedNode, kinParams);
}

static void bind_nearest(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                         const float *segEnd1, const float *segEnd2, const int *reachSeg, const int *reachCount,
                         int maxCand, int *boundSeg, float *bindArc, const int *seedNode, const float *kinParams) {
    float myoColTol = kinParams[7], alignTol = kinParams[8];
    for (int m = 0; m < nM; m++) {
        if (boundSeg[m] != FREE_BINDABLE) continue;
        float mx = head[m], my = head[nM + m], mz = head[2 * nM + m];
        float mux = uVec[m], muy = uVec[nM + m], muz = uVec[2 * nM + m];
        float rux = rodUVec[m], ruy = rodUVec[nM + m], ruz = rodUVec[2 * nM + m];
        int cnt = reachCount[m];
        if (cnt > maxCand) cnt = maxCand;
        int bestSeg = -1;
        float bestD = 1.0e30f, bestArc = 0.0f;
        for (int k = 0; k < cnt; k++) {
            int s = reachSeg[(size_t)m * maxCand + k];
            if (seedNode && seedNode[s] >= 0) continue;
            float e1x = segEnd1[s], e1y = segEnd1[nSeg + s], e1z = segEnd1[2 * nSeg + s];
            float e2x = segEnd2[s], e2y = segEnd2[nSeg + s], e2z = segEnd2[2 * nSeg + s];
            float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, e1x, e1y, e1z, e2x, e2y, e2z, myoColTol,
                                 alignTol);
            if (d >= 0.0f && d < bestD) {
                bestD = d;
                bestSeg = s;
                float r1x = e2x - e1x, r1y = e2y - e1y, r1z = e2z - e1z;
                float denom = r1x * r1x + r1y * r1y + r1z * r1z;
                float numer = (mx - e1x) * r1x + (my - e1y) * r1y + (mz - e1z) * r1z;
                bestArc = numer / (float)sqrt((double)denom);
            }
        }
        if (bestSeg >= 0) {
            boundSeg[m] = bestSeg;
            bindArc[m] = bestArc;
        }
    }
}

/* NucleotideCycleSystem.catchSlipRelease, NucleotideCycleSystem.java:77-135 */
void sbo_catch_slip_release(int nM, int *boundSeg, const float *forceDotFil, const float *forceMag, int *cooldown,
                            int *stats, int *capStats, const float *kinParams, int32_t step, int32_t seed, int rng,
                            uint32_t stream) {
    float kOff = kinParams[0], aCatch = kinParams[1], aSlip = kinParams[2];
    float xCatch = kinParams[3], xSlip = kinParams[4], kT = kinParams[5], dt = kinParams[6];
    int refractorySteps = (int)kinParams[10];
    float breakForceN = kinParams[11];
    int capOn = kinParams[12] > 0.5f;
    float blockProb = kinParams[13];
    for (int m = 0; m < nM; m++) {
        int bs = boundSeg[m];
        if (bs >= 0) {
            stats[2 * m] = stats[2 * m] + 1;
            int enterRefractory = refractorySteps > 0;
            if (blockProb < 1.0f) {
                float ru = sbo_uniform(sbo_event_bits(rng, (uint32_t)m, step, seed, stream, 0x52465241u));
                enterRefractory = enterRefractory && (ru < blockProb);
            }
            if (capOn && forceMag[m] > breakForceN) {
                capStats[m] = capStats[m] + 1;
                if (enterRefractory) { boundSeg[m] = FREE_COOLDOWN; cooldown[m] = refractorySteps; }
                else boundSeg[m] = FREE_BINDABLE;
                continue;
            }
            float F = forceDotFil[m];
            float rate = kOff * (aCatch * (float)exp((double)(-F * xCatch / kT)) +
                                 aSlip * (float)exp((double)(F * xSlip / kT)));
            float u = sbo_uniform(sbo_event_bits(rng, (uint32_t)m, step, seed, stream, 0x4D54u));
            if (u < rate * dt) {
                stats[2 * m + 1] = stats[2 * m + 1] + 1;
                if (enterRefractory) { boundSeg[m] = FREE_COOLDOWN; cooldown[m] = refractorySteps; }
                else boundSeg[m] = FREE_BINDABLE;
            }
        } else if (bs == FREE_COOLDOWN) {
            int c = cooldown[m] - 1;
            if (c <= 0) boundSeg[m] = FREE_BINDABLE;
            else cooldown[m] = c;
        }
    }
}

/* NucleotideCycleSystem.cycle, NucleotideCycleSystem.java:33-70 */
void sbo_nucleotide_cycle(int nM, int *nucleotideState, const int *boundSeg, const float *forceDotHist,
                          const float *nucParams, int32_t step, int32_t seed, int rng, uint32_t stream) {
    float dt = nucParams[0], atpOn = nucParams[1];
    float onATP = nucParams[2], offATP = nucParams[3];
    float onPi = nucParams[4], offPi = nucParams[5];
    float onADP = nucParams[6], offADP = nucParams[7];
    for (int m = 0; m < nM; m++) {
        int state = nucleotideState[m];
        int bound = boundSeg[m] >= 0;
        float u = sbo_uniform(sbo_event_bits(rng, (uint32_t)m, step, seed, stream, 0x4E55u));
        if (state == NUC_NONE) {
            if (u < atpOn * dt) state = NUC_ATP;
        } else if (state == NUC_ATP) {
            float rate = bound ? onATP : offATP;
            if (u < rate * dt) state = NUC_ADPPI;
        } else if (state == NUC_ADPPI) {
            float rate = bound ? onPi : offPi;
            if (u < rate * dt) state = NUC_ADP;
        } else {
            float avg = 0.0f;
            int b = m * 10;
            for (int k = 0; k < 10; k++) avg += forceDotHist[b + k];
            avg *= 0.1f;
            if (avg <= 0.0f) {
                float rate = bound ? onADP : offADP;
                if (u < rate * dt) state = NUC_NONE;
            }
        }
        nucleotideState[m] = state;
    }
}

/* CrossBridgeSystem.bondForces, CrossBridgeSystem.java:58-142 */
void sbo_bond_forces(int nM, int nSeg, const float *motorCoord, const float *motorUVec, const float *motorYVec,
                     const float *motorBRotGam, const float *filCoord, const float *filUVec, const float *filYVec,
                     const float *filBRotGam, const float *filSegLength, const int *boundSeg, const float *bindArc,
                     const int *nucleotideState, float *bondData, const float *xbParams) {
    int nB = 3 * nM;
    double myoSpring = xbParams[0], j1FMT = xbParams[2];
    double dt = xbParams[3], headLen = xbParams[4];
    double xbias = xbParams[5];
    const double DEG2RAD = PI_D / 180.0, RAD2DEG = 180.0 / PI_D;
    for (int m = 0; m < nM; m++) {
        int d = m * XB_STRIDE;
        for (int k = 0; k < XB_STRIDE; k++) bondData[d + k] = 0.0f;
        int s = boundSeg[m];
        if (s < 0) continue;
        int h = 3 * m + 2;
        double hcx = motorCoord[h], hcy = motorCoord[nB + h], hcz = motorCoord[2 * nB + h];
        double hux = motorUVec[h], huy = motorUVec[nB + h], huz = motorUVec[2 * nB + h];
        double hyx = motorYVec[h], hyy = motorYVec[nB + h], hyz = motorYVec[2 * nB + h];
        double hbRGx = motorBRotGam[h], hbRGy = motorBRotGam[nB + h];
        double htipx = hcx + 0.5 * headLen * hux, htipy = hcy + 0.5 * headLen * huy, htipz = hcz + 0.5 * headLen * huz;
        double scx = filCoord[s], scy = filCoord[nSeg + s], scz = filCoord[2 * nSeg + s];
        double sux = filUVec[s], suy = filUVec[nSeg + s], suz = filUVec[2 * nSeg + s];
        double syx = filYVec[s], syy = filYVec[nSeg + s], syz = filYVec[2 * nSeg + s];
        double sbRGx = filBRotGam[s], sbRGy = filBRotGam[nSeg + s];
        double slen = filSegLength[s];
        double aOff = bindArc[m] - 0.5 * slen;
        double apx = scx + aOff * sux, apy = scy + aOff * suy, apz = scz + aOff * suz;
        double dx = apx - htipx, dy = apy - htipy, dz = apz - htipz;
        double dist = sqrt(dx * dx + dy * dy + dz * dz);
        double fmag = myoSpring * dist;
        double Fx = 0, Fy = 0, Fz = 0;
        if (dist > 0.0) { double inv = fmag / dist; Fx = inv * dx; Fy = inv * dy; Fz = inv * dz; }
        double RHx = (htipx - hcx) * 1e-6, RHy = (htipy - hcy) * 1e-6, RHz = (htipz - hcz) * 1e-6;
        double THx = RHy * Fz - RHz * Fy, THy = RHz * Fx - RHx * Fz, THz = RHx * Fy - RHy * Fx;
        double RSx = (apx - scx) * 1e-6, RSy = (apy - scy) * 1e-6, RSz = (apz - scz) * 1e-6;
        double nFx = -Fx, nFy = -Fy, nFz = -Fz;
        double TSx = RSy * nFz - RSz * nFy, TSy = RSz * nFx - RSx * nFz, TSz = RSx * nFy - RSy * nFx;
        double restF9 = (nucleotideState[m] != NUC_ADPPI) ? 120.0 : 90.0;
        double t9x = suy * huz - suz * huy, t9y = suz * hux - sux * huz, t9z = sux * huy - suy * hux;
        double m9 = t9x * t9x + t9y * t9y + t9z * t9z;
        double T9x = 0, T9y = 0, T9z = 0;
        if (m9 > 1.0e-30) {
            double im = 1.0 / sqrt(m9);
            t9x *= im; t9y *= im; t9z *= im;
            double dot = sux * hux + suy * huy + suz * huz;
            if (dot > 1) dot = 1;
            if (dot < -1) dot = -1;
            double angD = sbo_accurate_acos(dot) * RAD2DEG - restF9;
            double tm = j1FMT * DEG2RAD * angD / ((1.0 / hbRGy + 1.0 / sbRGy) * dt);
            T9x = tm * t9x; T9y = tm * t9y; T9z = tm * t9z;
        }
        double t10x = syy * hyz - syz * hyy, t10y = syz * hyx - syx * hyz, t10z = syx * hyy - syy * hyx;
        double m10 = t10x * t10x + t10y * t10y + t10z * t10z;
        double T10x = 0, T10y = 0, T10z = 0;
        if (m10 > 1.0e-30) {
            double im = 1.0 / sqrt(m10);
            t10x *= im; t10y *= im; t10z *= im;
            double dot = syx * hyx + syy * hyy + syz * hyz;
            if (dot > 1) dot = 1;
            if (dot < -1) dot = -1;
            double ang = sbo_accurate_acos(dot) * RAD2DEG;
            double tm = j1FMT * DEG2RAD * ang / ((1.0 / hbRGx + 1.0 / sbRGx) * dt);
            T10x = tm * t10x; T10y = tm * t10y; T10z = tm * t10z;
        }
        bondData[d] = (float)Fx; bondData[d + 1] = (float)Fy; bondData[d + 2] = (float)Fz;
        bondData[d + 3] = (float)(THx - T9x - T10x);
        bondData[d + 4] = (float)(THy - T9y - T10y);
        bondData[d + 5] = (float)(THz - T9z - T10z);
        bondData[d + 6] = (float)(nFx - xbias); bondData[d + 7] = (float)nFy; bondData[d + 8] = (float)nFz;
        bondData[d + 9] = (float)(TSx + T9x + T10x);
        bondData[d + 10] = (float)(TSy + T9y + T10y);
        bondData[d + 11] = (float)(TSz + T9z + T10z);
        bondData[d + 12] = (float)(Fx * sux + Fy * suy + Fz * suz);
    }
}

/* CrossBridgeSystem.applyHeadForce, CrossBridgeSystem.java:145-157 */
void sbo_apply_head_force(int nM, const float *bondData, float *bodyForceSum, float *bodyTorqueSum) {
    int nB = 3 * nM;
    for (int m = 0; m < nM; m++) {
        int h = 3 * m + 2, d = m * XB_STRIDE;
        for (int a = 0; a < 3; a++) {
            bodyForceSum[a * nB + h] = bodyForceSum[a * nB + h] + bondData[d + a];        /* float + float */
            bodyTorqueSum[a * nB + h] = bodyTorqueSum[a * nB + h] + bondData[d + 3 + a];
        }
    }
}

/* CrossBridgeSystem.registerForceDot, CrossBridgeSystem.java:166-187 */
void sbo_register_force_dot(int nM, const float *bondData, const int *boundSeg, float *forceDotFil, float *forceMag,
                            float *forceDotHist, int *forceDotPlace) {
    for (int m = 0; m < nM; m++) {
        if (boundSeg[m] >= 0) {
            int d = m * XB_STRIDE;
            float fd = bondData[d + 12];
            forceDotFil[m] = fd;
            float fx = bondData[d], fy = bondData[d + 1], fz = bondData[d + 2];
            forceMag[m] = (float)sqrt((double)(fx * fx + fy * fy + fz * fz));
            int p = forceDotPlace[m];
            forceDotHist[m * 10 + p] = fd;
            forceDotPlace[m] = (p + 1) % 10;
        } else {
            forceDotFil[m] = 0.0f;
            forceMag[m] = 0.0f;
            forceDotPlace[m] = 0;
            for (int k = 0; k < 10; k++) forceDotHist[m * 10 + k] = 0.0f;
        }
    }
}

/* CrossBridgeSystem.segGather, CrossBridgeSystem.java:302-320 */
void sbo_motor_seg_gather(int nSeg, const int *segMotorOffsets, const int *segMotorMyo, const float *bondData,
                          float *filForceSum, float *filTorqueSum) {
    for (int s = 0; s < nSeg; s++) {
        double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        for (int k = segMotorOffsets[s]; k < segMotorOffsets[s + 1]; k++) {
            int d = segMotorMyo[k] * XB_STRIDE;
            fx += bondData[d + 6]; fy += bondData[d + 7]; fz += bondData[d + 8];
            tx += bondData[d + 9]; ty += bondData[d + 10]; tz += bondData[d + 11];
        }
        filForceSum[s] = (float)(filForceSum[s] + fx);
        filForceSum[nSeg + s] = (float)(filForceSum[nSeg + s] + fy);
        filForceSum[2 * nSeg + s] = (float)(filForceSum[2 * nSeg + s] + fz);
        filTorqueSum[s] = (float)(filTorqueSum[s] + tx);
        filTorqueSum[nSeg + s] = (float)(filTorqueSum[nSeg + s] + ty);
        filTorqueSum[2 * nSeg + s] = (float)(filTorqueSum[2 * nSeg + s] + tz);
    }
}

/* MotorJointSystem.moveC, MotorJointSystem.java:61-70 */
static double move_c(double ux, double uy, double uz, double lx, double ly, double lz, double bTGx, double bTGy,
                     double bRGy, double lenUm) {
    double cosB = ux * lx + uy * ly + uz * lz;
    if (cosB > 1.0) cosB = 1.0;
    if (cosB < -1.0) cosB = -1.0;
    double cosB2 = cosB * cosB;
    double cosA2 = 1.0 - cosB2;
    double lSq = 1.0e-12 * lenUm * lenUm;
    return cosB2 / bTGx + cosA2 / bTGy + lSq * cosA2 / (4.0 * bRGy);
}

/* MotorJointSystem.joints, MotorJointSystem.java:72-198 */
void sbo_motor_joints(int nB, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                      const float *bRotGam, float *forceSum, float *torqueSum, const int *nucleotideState,
                      const float *jointParams) {
    double dt = jointParams[0];
    double j1FracMove = jointParams[1], j1FracR = jointParams[2];
    double j1FracMoveTorq = jointParams[3];
    double j2FracMove = jointParams[5], j2FracR = jointParams[6];
    double j2FracMoveTorq = jointParams[7], j2Rest = jointParams[8];
    double stallPN = jointParams[10];
    const double DEG2RAD = PI_D / 180.0, RAD2DEG = 180.0 / PI_D;
    for (int s = 0; s < nB; s++) {
        int m = s / 3;
        int role = s - 3 * m;
        double j1Rest = (nucleotideState[m] != NUC_ADPPI) ? 60.0 : 0.0;
        int rod = 3 * m, lever = 3 * m + 1, head = 3 * m + 2;
        double rcx = coord[rod], rcy = coord[nB + rod], rcz = coord[2 * nB + rod];
        double rux = uVec[rod], ruy = uVec[nB + rod], ruz = uVec[2 * nB + rod];
        double rlen = segLength[rod];
        double lcx = coord[lever], lcy = coord[nB + lever], lcz = coord[2 * nB + lever];
        double lux = uVec[lever], luy = uVec[nB + lever], luz = uVec[2 * nB + lever];
        double llen = segLength[lever];
        double hcx = coord[head], hcy = coord[nB + head], hcz = coord[2 * nB + head];
        double hux = uVec[head], huy = uVec[nB + head], huz = uVec[2 * nB + head];
        double hlen = segLength[head];
        double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        if (role == 0 || role == 1) {
            double aEx = rcx + 0.5 * rlen * rux, aEy = rcy + 0.5 * rlen * ruy, aEz = rcz + 0.5 * rlen * ruz;
            double bEx = lcx - 0.5 * llen * lux, bEy = lcy - 0.5 * llen * luy, bEz = lcz - 0.5 * llen * luz;
            double dx = aEx - bEx, dy = aEy - bEy, dz = aEz - bEz;
            double strain = sqrt(dx * dx + dy * dy + dz * dz);
            double inv = (strain > 0.0) ? 1.0 / strain : 0.0;
            double lx = dx * inv, ly = dy * inv, lz = dz * inv;
            double mcRod = move_c(rux, ruy, ruz, lx, ly, lz, bTransGam[rod], bTransGam[nB + rod], bRotGam[nB + rod], rlen);
            double mcLever = move_c(lux, luy, luz, lx, ly, lz, bTransGam[lever], bTransGam[nB + lever],
                                    bRotGam[nB + lever], llen);
            double denom = dt * (mcRod + mcLever);
            double forceMag = (denom > 0.0) ? (j2FracMove * 1.0e-6 * strain / denom) : 0.0;
            double tvx = ruy * luz - ruz * luy, tvy = ruz * lux - rux * luz, tvz = rux * luy - ruy * lux;
            double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
            double torsionMag = 0.0;
            if (tvm2 > 1.0e-30) {
                double im = 1.0 / sqrt(tvm2);
                tvx *= im; tvy *= im; tvz *= im;
                double dotV = rux * lux + ruy * luy + ruz * luz;
                if (dotV > 1.0) dotV = 1.0;
                if (dotV < -1.0) dotV = -1.0;
                double ang = sbo_accurate_acos(dotV) * RAD2DEG;
                double invBRG = 1.0 / bRotGam[nB + rod] + 1.0 / bRotGam[nB + lever];
                torsionMag = j2FracMoveTorq * DEG2RAD * (ang - j2Rest) / (invBRG * dt);
            }
            if (role == 0) {
                double Fx = -forceMag * lx, Fy = -forceMag * ly, Fz = -forceMag * lz;
                fx += Fx; fy += Fy; fz += Fz;
                double Rs = 0.5e-6 * rlen * j2FracR, Rx = Rs * rux, Ry = Rs * ruy, Rz = Rs * ruz;
                tx += Ry * Fz - Rz * Fy; ty += Rz * Fx - Rx * Fz; tz += Rx * Fy - Ry * Fx;
                tx += tvx * torsionMag; ty += tvy * torsionMag; tz += tvz * torsionMag;
            } else {
                double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
                fx += Fx; fy += Fy; fz += Fz;
                double Rs = -0.5e-6 * llen * j2FracR, Rx = Rs * lux, Ry = Rs * luy, Rz = Rs * luz;
                tx += Ry * Fz - Rz * Fy; ty += Rz * Fx - Rx * Fz; tz += Rx * Fy - Ry * Fx;
                tx -= tvx * torsionMag; ty -= tvy * torsionMag; tz -= tvz * torsionMag;
            }
        }
        if (role == 1 || role == 2) {
            double aEx = lcx + 0.5 * llen * lux, aEy = lcy + 0.5 * llen * luy, aEz = lcz + 0.5 * llen * luz;
            double bEx = hcx - 0.5 * hlen * hux, bEy = hcy - 0.5 * hlen * huy, bEz = hcz - 0.5 * hlen * huz;
            double dx = aEx - bEx, dy = aEy - bEy, dz = aEz - bEz;
            double strain = sqrt(dx * dx + dy * dy + dz * dz);
            double inv = (strain > 0.0) ? 1.0 / strain : 0.0;
            double lx = dx * inv, ly = dy * inv, lz = dz * inv;
            double mcLever = move_c(lux, luy, luz, lx, ly, lz, bTransGam[lever], bTransGam[nB + lever],
                                    bRotGam[nB + lever], llen);
            double mcHead = move_c(hux, huy, huz, lx, ly, lz, bTransGam[head], bTransGam[nB + head], bRotGam[nB + head],
                                   hlen);
            double denom = dt * (mcLever + mcHead);
            double forceMag = (denom > 0.0) ? (j1FracMove * 1.0e-6 * strain / denom) : 0.0;
            double tvx = luy * huz - luz * huy, tvy = luz * hux - lux * huz, tvz = lux * huy - luy * hux;
            double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
            double torsionMag = 0.0;
            if (tvm2 > 1.0e-30) {
                double im = 1.0 / sqrt(tvm2);
                tvx *= im; tvy *= im; tvz *= im;
                double dotV = lux * hux + luy * huy + luz * huz;
                if (dotV > 1.0) dotV = 1.0;
                if (dotV < -1.0) dotV = -1.0;
                double ang = sbo_accurate_acos(dotV) * RAD2DEG;
                double invBRG = 1.0 / bRotGam[nB + lever] + 1.0 / bRotGam[nB + head];
                torsionMag = j1FracMoveTorq * DEG2RAD * (ang - j1Rest) / (invBRG * dt);
                double maxMag = stallPN * 0.5 * hlen * 1.0e-18;
                if (torsionMag > maxMag) torsionMag = maxMag;
            }
            if (role == 1) {
                double Fx = -forceMag * lx, Fy = -forceMag * ly, Fz = -forceMag * lz;
                fx += Fx; fy += Fy; fz += Fz;
                double Rs = 0.5e-6 * llen * j1FracR, Rx = Rs * lux, Ry = Rs * luy, Rz = Rs * luz;
                tx += Ry * Fz - Rz * Fy; ty += Rz * Fx - Rx * Fz; tz += Rx * Fy - Ry * Fx;
                tx += tvx * torsionMag; ty += tvy * torsionMag; tz += tvz * torsionMag;
            } else {
                double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
                fx += Fx; fy += Fy; fz += Fz;
                double Rs = -0.5e-6 * hlen * j1FracR, Rx = Rs * hux, Ry = Rs * huy, Rz = Rs * huz;
                tx += Ry * Fz - Rz * Fy; ty += Rz * Fx - Rx * Fz; tz += Rx * Fy - Ry * Fx;
                tx -= tvx * torsionMag; ty -= tvy * torsionMag; tz -= tvz * torsionMag;
            }
        }
        forceSum[s] = (float)(forceSum[s] + fx);
        forceSum[nB + s] = (float)(forceSum[nB + s] + fy);
        forceSum[2 * nB + s] = (float)(forceSum[2 * nB + s] + fz);
        torqueSum[s] = (float)(torqueSum[s] + tx);
        torqueSum[nB + s] = (float)(torqueSum[nB + s] + ty);
        torqueSum[2 * nB + s] = (float)(torqueSum[2 * nB + s] + tz);
    }
}

/* DimerCouplingSystem.couple, DimerCouplingSystem.java:83-184 (moveC is MotorJointSystem's) */
void sbo_dimer_couple(int nB, int nD, const float *coord, const float *uVec, const float *segLength,
                      const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                      const int *motorA, const int *motorB, const int *parallel, const float *dimerParams,
                      const int *boundSeg) {
    double dt = dimerParams[0], fracMove = dimerParams[1], leverTorq = dimerParams[2];
    double leverDeg = dimerParams[3], rodLen = dimerParams[4];
    for (int d = 0; d < nD; d++) {
        int mA = motorA[d], mB = motorB[d];
        int rodA = 3 * mA, rodB = 3 * mB;
        int par = parallel[d] != 0;
        for (int variant = 0; variant < 2; variant++) {
            double endSelA = (variant == 0) ? -1.0 : 1.0;
            double endSelB = par ? endSelA : -endSelA;
            double aCx = coord[rodA], aCy = coord[nB + rodA], aCz = coord[2 * nB + rodA];
            double aUx = uVec[rodA], aUy = uVec[nB + rodA], aUz = uVec[2 * nB + rodA];
            double bCx = coord[rodB], bCy = coord[nB + rodB], bCz = coord[2 * nB + rodB];
            double bUx = uVec[rodB], bUy = uVec[nB + rodB], bUz = uVec[2 * nB + rodB];
            double aLen = segLength[rodA], bLen = segLength[rodB];
            double pAx = aCx + endSelA * 0.5 * aLen * aUx, pAy = aCy + endSelA * 0.5 * aLen * aUy,
                   pAz = aCz + endSelA * 0.5 * aLen * aUz;
            double pBx = bCx + endSelB * 0.5 * bLen * bUx, pBy = bCy + endSelB * 0.5 * bLen * bUy,
                   pBz = bCz + endSelB * 0.5 * bLen * bUz;
            double dx = pBx - pAx, dy = pBy - pAy, dz = pBz - pAz;
            double strain = sqrt(dx * dx + dy * dy + dz * dz);
            if (strain > 0.0) {
                double inv = 1.0 / strain;
                double lx = dx * inv, ly = dy * inv, lz = dz * inv;
                double mcA = move_c(aUx, aUy, aUz, lx, ly, lz, bTransGam[rodA], bTransGam[nB + rodA], bRotGam[nB + rodA], aLen);
                double mcB = move_c(bUx, bUy, bUz, lx, ly, lz, bTransGam[rodB], bTransGam[nB + rodB], bRotGam[nB + rodB], bLen);
                double denom = dt * (mcA + mcB);
                double forceMag = (denom > 0.0) ? (fracMove * 1.0e-6 * strain / denom) : 0.0;
                double FAx = forceMag * lx, FAy = forceMag * ly, FAz = forceMag * lz;
                double RsA = 0.5e-6 * rodLen * endSelA;
                double TAx = (RsA * aUy) * FAz - (RsA * aUz) * FAy;
                double TAy = (RsA * aUz) * FAx - (RsA * aUx) * FAz;
                double TAz = (RsA * aUx) * FAy - (RsA * aUy) * FAx;
                forceSum[rodA] = (float)(forceSum[rodA] + FAx);
                forceSum[nB + rodA] = (float)(forceSum[nB + rodA] + FAy);
                forceSum[2 * nB + rodA] = (float)(forceSum[2 * nB + rodA] + FAz);
                torqueSum[rodA] = (float)(torqueSum[rodA] + TAx);
                torqueSum[nB + rodA] = (float)(torqueSum[nB + rodA] + TAy);
                torqueSum[2 * nB + rodA] = (float)(torqueSum[2 * nB + rodA] + TAz);
                double RsB = 0.5e-6 * rodLen * endSelB;
                double TBx = (RsB * bUy) * (-FAz) - (RsB * bUz) * (-FAy);
                double TBy = (RsB * bUz) * (-FAx) - (RsB * bUx) * (-FAz);
                double TBz = (RsB * bUx) * (-FAy) - (RsB * bUy) * (-FAx);
                forceSum[rodB] = (float)(forceSum[rodB] - FAx);
                forceSum[nB + rodB] = (float)(forceSum[nB + rodB] - FAy);
                forceSum[2 * nB + rodB] = (float)(forceSum[2 * nB + rodB] - FAz);
                torqueSum[rodB] = (float)(torqueSum[rodB] + TBx);
                torqueSum[nB + rodB] = (float)(torqueSum[nB + rodB] + TBy);
                torqueSum[2 * nB + rodB] = (float)(torqueSum[2 * nB + rodB] + TBz);
            }
        }
        int bothBound = boundSeg[mA] >= 0 && boundSeg[mB] >= 0;
        if (par && !bothBound) {
            int leverA = 3 * mA + 1, leverB = 3 * mB + 1;
            double u1x = uVec[leverA], u1y = uVec[nB + leverA], u1z = uVec[2 * nB + leverA];
            double u2x = uVec[leverB], u2y = uVec[nB + leverB], u2z = uVec[2 * nB + leverB];
            double tvx = u1y * u2z - u1z * u2y, tvy = u1z * u2x - u1x * u2z, tvz = u1x * u2y - u1y * u2x;
            double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
            if (tvm2 > 1.0e-30) {
                double im = 1.0 / sqrt(tvm2);
                tvx *= im; tvy *= im; tvz *= im;
                double dot = u1x * u2x + u1y * u2y + u1z * u2z;
                if (dot > 1.0) dot = 1.0;
                if (dot < -1.0) dot = -1.0;
                double ang = sbo_accurate_acos(dot) * 180.0 / PI_D;
                double invBRG = 1.0 / bRotGam[nB + leverA] + 1.0 / bRotGam[nB + leverB];
                double torsionMag = leverTorq * (PI_D / 180.0) * (ang - leverDeg) / (invBRG * dt);
                torqueSum[leverA] = (float)(torqueSum[leverA] + tvx * torsionMag);
                torqueSum[nB + leverA] = (float)(torqueSum[nB + leverA] + tvy * torsionMag);
                torqueSum[2 * nB + leverA] = (float)(torqueSum[2 * nB + leverA] + tvz * torsionMag);
                torqueSum[leverB] = (float)(torqueSum[leverB] - tvx * torsionMag);
                torqueSum[nB + leverB] = (float)(torqueSum[nB + leverB] - tvy * torsionMag);
                torqueSum[2 * nB + leverB] = (float)(torqueSum[2 * nB + leverB] - tvz * torsionMag);
            }
        }
    }
}

/* MiniFilamentSystem.tether, MiniFilamentSystem.java:65-132 (its accurateAcos is the shared one) */
void sbo_mini_tether(int nMB, int nBb, int nD, const float *mCoord, const float *mUVec, const float *mSegLength,
                     const float *mBTransGam, const float *mBRotGam, float *mForceSum, float *mTorqueSum,
                     const float *bbCoord, const float *bbUVec, const float *bbInvDragY, const int *headBackboneSlot,
                     const int *motorA, const float *attachAxial, float *miniData, const float *miniParams) {
    double dt = miniParams[0], fracMove = miniParams[1], align = miniParams[2];
    for (int d = 0; d < nD; d++) {
        int bb = headBackboneSlot[d];
        int rodA = 3 * motorA[d];
        double ax = attachAxial[d];
        double bcx = bbCoord[bb], bcy = bbCoord[nBb + bb], bcz = bbCoord[2 * nBb + bb];
        double bux = bbUVec[bb], buy = bbUVec[nBb + bb], buz = bbUVec[2 * nBb + bb];
        double pax = bcx + ax * bux, pay = bcy + ax * buy, paz = bcz + ax * buz;
        double rcx = mCoord[rodA], rcy = mCoord[nMB + rodA], rcz = mCoord[2 * nMB + rodA];
        double rux = mUVec[rodA], ruy = mUVec[nMB + rodA], ruz = mUVec[2 * nMB + rodA];
        double rlen = mSegLength[rodA];
        double re1x = rcx - 0.5 * rlen * rux, re1y = rcy - 0.5 * rlen * ruy, re1z = rcz - 0.5 * rlen * ruz;
        double fbx = 0, fby = 0, fbz = 0, tbx = 0, tby = 0, tbz = 0;
        double dx = pax - re1x, dy = pay - re1y, dz = paz - re1z;
        double strain = sqrt(dx * dx + dy * dy + dz * dz);
        if (strain > 0.0) {
            double inv = 1.0 / strain;
            double lx = dx * inv, ly = dy * inv, lz = dz * inv;
            double denom = dt * (1.0 / mBTransGam[nMB + rodA] + bbInvDragY[bb]);
            double forceMag = (denom > 0.0) ? (fracMove * 1.0e-6 * strain / denom) : 0.0;
            double Frx = forceMag * lx, Fry = forceMag * ly, Frz = forceMag * lz;
            mForceSum[rodA] = (float)(mForceSum[rodA] + Frx);
            mForceSum[nMB + rodA] = (float)(mForceSum[nMB + rodA] + Fry);
            mForceSum[2 * nMB + rodA] = (float)(mForceSum[2 * nMB + rodA] + Frz);
            fbx = -Frx; fby = -Fry; fbz = -Frz;
        }
        double alignSign = (ax < 0.0) ? -1.0 : 1.0;
        double tux = alignSign * bux, tuy = alignSign * buy, tuz = alignSign * buz;
        double tvx = ruy * tuz - ruz * tuy, tvy = ruz * tux - rux * tuz, tvz = rux * tuy - ruy * tux;
        double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
        if (tvm2 > 1.0e-30) {
            double im = 1.0 / sqrt(tvm2);
            tvx *= im; tvy *= im; tvz *= im;
            double dot = rux * tux + ruy * tuy + ruz * tuz;
            if (dot > 1.0) dot = 1.0;
            if (dot < -1.0) dot = -1.0;
            double angDeg = sbo_accurate_acos(dot) * 180.0 / PI_D;
            double denomR = (1.0 / mBRotGam[nMB + rodA] + bbInvDragY[nBb + bb]) * dt;
            double tmag = align * (PI_D / 180.0) * angDeg / denomR;
            mTorqueSum[rodA] = (float)(mTorqueSum[rodA] + tvx * tmag);
            mTorqueSum[nMB + rodA] = (float)(mTorqueSum[nMB + rodA] + tvy * tmag);
            mTorqueSum[2 * nMB + rodA] = (float)(mTorqueSum[2 * nMB + rodA] + tvz * tmag);
            tbx = -tvx * tmag; tby = -tvy * tmag; tbz = -tvz * tmag;
        }
        int o = MINI_STRIDE * d;
        miniData[o] = (float)fbx; miniData[o + 1] = (float)fby; miniData[o + 2] = (float)fbz;
        miniData[o + 3] = (float)tbx; miniData[o + 4] = (float)tby; miniData[o + 5] = (float)tbz;
    }
}

/* MiniFilamentSystem.backboneGather, MiniFilamentSystem.java:137-155 */
void sbo_backbone_gather(int nBb, const int *bbDimerOffsets, const int *bbDimerList, const float *miniData,
                         float *bbForceSum, float *bbTorqueSum) {
    for (int s = 0; s < nBb; s++) {
        double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        for (int k = bbDimerOffsets[s]; k < bbDimerOffsets[s + 1]; k++) {
            int o = bbDimerList[k] * MINI_STRIDE;
            fx += miniData[o]; fy += miniData[o + 1]; fz += miniData[o + 2];
            tx += miniData[o + 3]; ty += miniData[o + 4]; tz += miniData[o + 5];
        }
        bbForceSum[s] = (float)(bbForceSum[s] + fx);
        bbForceSum[nBb + s] = (float)(bbForceSum[nBb + s] + fy);
        bbForceSum[2 * nBb + s] = (float)(bbForceSum[2 * nBb + s] + fz);
        bbTorqueSum[s] = (float)(bbTorqueSum[s] + tx);
        bbTorqueSum[nBb + s] = (float)(bbTorqueSum[nBb + s] + ty);
        bbTorqueSum[2 * nBb + s] = (float)(bbTorqueSum[2 * nBb + s] + tz);
    }
}

/* DenseContractileHarness.cpuStep (DenseContractileHarness.java:350-429): binding, motor structure,
 * filament block.  net->gridOn must be set; the grid covers filaments + motor heads.  net->center /
 * boundingRadius / bodyCell / gridCellContents are sized for cap = nRod + nM. */
void sbo_contractile_step(sbo_net *net, sbo_motors *mo, int32_t step) {
    int n = net->nRod, nM = mo->nM, nB = 3 * nM, nBb = mo->nBb;
    int cap = n + nM;
    int mc = net->maxCand;
    /* --- turnover at the biochem cadence (FullSystemDemoHarness.java:650-684), before binding --- */
    if (net->turn && net->turn->interval > 0 && step % net->turn->interval == 0)   /* GrowthStore.firesAt */
        sbo_turnover_cadence(net->turn, net, step);
    /* --- binding --- */
    sbo_publish_head_from_body(nM, mo->coord, mo->uVec, mo->segLength, mo->head, mo->hUVec, mo->rodUVec);
    sbo_publish_filaments(n, cap, net->coord, net->segLength, net->actinRadius, net->center, net->boundingRadius);
    for (int i = 0; i < n; i++) { mo->ownerStore[i] = 0; mo->ownerSlot[i] = i; }
    sbo_publish_motors(nM, cap, n, mo->head, mo->reach, net->center, net->boundingRadius, mo->ownerStore, mo->ownerSlot);
    sbo_body_cell(cap, cap, net->center, net->gridParams, net->gridDims, 0, net->bodyCell);
    sbo_grid_build_serial(cap, cap, net->bodyCell, net->gridDims, net->cellCount, net->gridCellOffsets,
                          net->gridCellContents);
    sbo_grid_reachable(nM, n, mo->head, mo->hUVec, mo->rodUVec, net->end1, net->end2, net->gridParams, net->gridDims,
                       net->gridCellOffsets, net->gridCellContents, mo->ownerStore, mo->ownerSlot, mc, mo->reachSeg,
                       mo->reachCount, mo->kinParams);
    sbo_catch_slip_release(nM, mo->boundSeg, mo->forceDotFil, mo->forceMag, mo->cooldown, mo->stats, mo->capStats,
                           mo->kinParams, step, mo->seedMotor, net->rng, net->stream);
    sbo_bind_nearest(nM, n, mo->head, mo->hUVec, mo->rodUVec, net->end1, net->end2, mo->reachSeg, mo->reachCount, mc,
                     mo->boundSeg, mo->bindArc, mo->kinParams);
    sbo_nucleotide_cycle(nM, mo->nucleotideState, mo->boundSeg, mo->forceDotHist, mo->nucParams, step, mo->seedMotor,
                         net->rng, net->stream);
    /* --- minifilament structure mechanics --- */
    sbo_zero_accumulators(nB, mo->forceSum, mo->torqueSum);
    sbo_zero_accumulators(nBb, mo->bbForceSum, mo->bbTorqueSum);
    sbo_brownian_force(nB, mo->randForce, mo->randTorque, mo->bTransGam, mo->bRotGam, mo->brownTransScale,
                       mo->brownRotScale, mo->bodyParams[1], step, mo->seedMotor, net->rng, net->stream, 0u);
    sbo_brownian_force(nBb, mo->bbRandForce, mo->bbRandTorque, mo->bbTransGam, mo->bbRotGam, mo->bbBts, mo->bbBrs,
                       mo->bbBodyParams[1], step, mo->seedBackbone, net->rng, net->stream, 0u);
    sbo_motor_joints(nB, mo->coord, mo->uVec, mo->segLength, mo->bTransGam, mo->bRotGam, mo->forceSum, mo->torqueSum,
                     mo->nucleotideState, mo->jointParams);
    sbo_dimer_couple(nB, mo->nD, mo->coord, mo->uVec, mo->segLength, mo->bTransGam, mo->bRotGam, mo->forceSum,
                     mo->torqueSum, mo->motorA, mo->motorB, mo->parallel, mo->dimerParams, mo->boundSeg);
    if (mo->nodeMode)   /* V2OneXHarness.java:399-400 */
        sbo_node_tether_keys(nB, nBb, mo->nMD, mo->coord, mo->uVec, mo->segLength, mo->bTransGam, mo->forceSum,
                             mo->torqueSum, mo->bbCoord, mo->bbUVec, mo->bbYVec, mo->bbInvDragY, mo->headBackboneSlot,
                             mo->tetherMotor, mo->radial, mo->attachCoeffK, mo->nodeMode == 2 ? mo->boundSeg : NULL,
                             mo->miniData, mo->miniParams);
    else
        sbo_mini_tether(nB, nBb, mo->nMD, mo->coord, mo->uVec, mo->segLength, mo->bTransGam, mo->bRotGam,
                        mo->forceSum, mo->torqueSum, mo->bbCoord, mo->bbUVec, mo->bbInvDragY, mo->headBackboneSlot,
                        mo->tetherMotor, mo->attachAxial, mo->miniData, mo->miniParams);
    sbo_csr_build(nBb, mo->nMD, mo->headBackboneSlot, mo->bbDimerCount, mo->bbDimerOffsets, mo->bbDimerList);
    sbo_backbone_gather(nBb, mo->bbDimerOffsets, mo->bbDimerList, mo->miniData, mo->bbForceSum, mo->bbTorqueSum);
    sbo_bond_forces(nM, n, mo->coord, mo->uVec, mo->yVec, mo->bRotGam, net->coord, net->uVec, net->yVec, net->bRotGam,
                    net->segLength, mo->boundSeg, mo->bindArc, mo->nucleotideState, mo->bondData, mo->xbParams);
    sbo_apply_head_force(nM, mo->bondData, mo->forceSum, mo->torqueSum);
    sbo_integrate(nB, mo->coord, mo->uVec, mo->yVec, mo->forceSum, mo->torqueSum, mo->randForce, mo->randTorque,
                  mo->bTransGam, mo->bRotGam, mo->bodyParams[0]);
    if (net->confineOn)
        sbo_confine(nBb, mo->bbCoord, mo->bbUVec, mo->bbSegLength, mo->bbTransGam, mo->bbForceSum, mo->bbTorqueSum,
                    net->boxParams, step);
    sbo_integrate(nBb, mo->bbCoord, mo->bbUVec, mo->bbYVec, mo->bbForceSum, mo->bbTorqueSum, mo->bbRandForce,
                  mo->bbRandTorque, mo->bbTransGam, mo->bbRotGam, mo->bbBodyParams[0]);
    sbo_derive(nB, mo->coord, mo->uVec, mo->yVec, mo->zVec, mo->end1, mo->end2, mo->segLength);
    sbo_derive(nBb, mo->bbCoord, mo->bbUVec, mo->bbYVec, mo->bbZVec, mo->bbEnd1, mo->bbEnd2, mo->bbSegLength);
    sbo_register_force_dot(nM, mo->bondData, mo->boundSeg, mo->forceDotFil, mo->forceMag, mo->forceDotHist,
                           mo->forceDotPlace);
    /* --- filament block: forces, motor -> segment CSR + gather, confine, integrate, derive --- */
    if (net->form && net->formInt > 0 && step % net->formInt == 0)   /* V2OneXHarness.java:379 */
        net->form->lastFormed = sbo_formation(net->form, step, net->seed, net->rng, net->stream);
    sbo_network_forces(net, step);
    sbo_csr_build(n, nM, mo->boundSeg, mo->segMotorCount, mo->segMotorOffsets, mo->segMotorMyo);
    sbo_motor_seg_gather(n, mo->segMotorOffsets, mo->segMotorMyo, mo->bondData, net->forceSum, net->torqueSum);
    sbo_network_finish(net, step);
}
