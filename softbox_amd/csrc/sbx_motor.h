// sbx_motor.h -- myosin force laws and kinetics as device functions (gfx950), shared by the
// stand-alone System entry points (sbx_motor.hip) and the contractile stage of the network step.
// Motor body: sub-body 3m rod, 3m+1 lever, 3m+2 head, planar stride nB = 3*nMotors
// (MotorStore.java:45-60); published head/axis arrays have stride nMotors.
#pragma once
#include "sbx_device.h"

namespace sbx {

constexpr int NUC_NONE = 0, NUC_ATP = 1, NUC_ADPPI = 2, NUC_ADP = 3;
constexpr int FREE_BINDABLE = -1, FREE_COOLDOWN = -2;
constexpr int XB_STRIDE = 13;
constexpr int MINI_STRIDE = 6;

// BindingDetectionSystem.reachTestDistSq (BindingDetectionSystem.java:59-85): conDist^2 or -1
__device__ __forceinline__ float reach_test(float mx, float my, float mz, float mux, float muy, float muz, float rux,
                                            float ruy, float ruz, float e1x, float e1y, float e1z, float e2x, float e2y,
                                            float e2z, float myoColTol, float alignTol) {
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

__device__ __forceinline__ int clamp_bin(float v, int n) {
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)n) return n - 1;
    int c = (int)v;
    return c >= n ? n - 1 : c;
}

struct ReachGrid {
    float xMin, yMin, zMin, invCell;
    int nX, nY, nZ;
};

// BindingDetectionSystem.gridReachable body for one motor (BindingDetectionSystem.java:157-227)
__device__ __forceinline__ int grid_reachable_one(int m, int nM, int nSeg, const float *__restrict__ head,
                                                  const float *__restrict__ hu, const float *__restrict__ ru,
                                                  const float *__restrict__ e1, const float *__restrict__ e2,
                                                  const ReachGrid &g, const int *__restrict__ offsets,
                                                  const int *__restrict__ contents, const int *__restrict__ ownerStore,
                                                  const int *__restrict__ ownerSlot, int *__restrict__ reachSeg,
                                                  int maxCand, float myoColTol, float alignTol) {
    float mx = head[m], my = head[nM + m], mz = head[2 * nM + m];
    float mux = hu[m], muy = hu[nM + m], muz = hu[2 * nM + m];
    float rux = ru[m], ruy = ru[nM + m], ruz = ru[2 * nM + m];
    int cx = clamp_bin((mx - g.xMin) * g.invCell, g.nX);
    int cy = clamp_bin((my - g.yMin) * g.invCell, g.nY);
    int cz = clamp_bin((mz - g.zMin) * g.invCell, g.nZ);
    int x0 = cx - 1 < 0 ? 0 : cx - 1, x1 = cx + 1 >= g.nX ? g.nX - 1 : cx + 1;
    int y0 = cy - 1 < 0 ? 0 : cy - 1, y1 = cy + 1 >= g.nY ? g.nY - 1 : cy + 1;
    int z0 = cz - 1 < 0 ? 0 : cz - 1, z1 = cz + 1 >= g.nZ ? g.nZ - 1 : cz + 1;
    int nXY = g.nX * g.nY;
    int out = 0;
    size_t slice = (size_t)m * maxCand;
    for (int zz = z0; zz <= z1; zz++)
        for (int yy = y0; yy <= y1; yy++)
            for (int xx = x0; xx <= x1; xx++) {
                int cc = xx + yy * g.nX + zz * nXY;
                int end = offsets[cc + 1];
                for (int idx = offsets[cc]; idx < end; idx++) {
                    int j = contents[idx];
                    if (ownerStore[j] != 0) continue;
                    int s = ownerSlot[j];
                    float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, e1[s], e1[nSeg + s], e1[2 * nSeg + s],
                                         e2[s], e2[nSeg + s], e2[2 * nSeg + s], myoColTol, alignTol);
                    if (d >= 0.0f) {
                        if (out < maxCand) reachSeg[slice + out] = s;
                        out++;
                    }
                }
            }
    return out;
}

struct KinParams {
    float kOff, aCatch, aSlip, xCatch, xSlip, kT, dt, myoColTol, alignTol;
    int refractorySteps;
    float breakForceN;
    int capOn;
    float blockProb;
};

// NucleotideCycleSystem.catchSlipRelease body for one motor (NucleotideCycleSystem.java:90-133)
__device__ __forceinline__ void catch_slip_one(int m, int &bs, int &cool, int *__restrict__ stats,
                                               int *__restrict__ capStats, float forceDotFil, float forceMag,
                                               const KinParams &k, int step, int seed, int rng, uint32_t stream) {
    if (bs >= 0) {
        stats[2 * m] = stats[2 * m] + 1;
        bool enterRefractory = k.refractorySteps > 0;
        if (k.blockProb < 1.0f) {
            float ru = uniform01(event_bits(rng, (uint32_t)m, step, seed, stream, 0x52465241u));
            enterRefractory = enterRefractory && (ru < k.blockProb);
        }
        if (k.capOn && forceMag > k.breakForceN) {
            capStats[m] = capStats[m] + 1;
            if (enterRefractory) { bs = FREE_COOLDOWN; cool = k.refractorySteps; }
            else bs = FREE_BINDABLE;
            return;
        }
        float F = forceDotFil;
        float rate = k.kOff * (k.aCatch * (float)exp((double)(-F * k.xCatch / k.kT)) +
                               k.aSlip * (float)exp((double)(F * k.xSlip / k.kT)));
        float u = uniform01(event_bits(rng, (uint32_t)m, step, seed, stream, 0x4D54u));
        if (u < rate * k.dt) {
            stats[2 * m + 1] = stats[2 * m + 1] + 1;
            if (enterRefractory) { bs = FREE_COOLDOWN; cool = k.refractorySteps; }
            else bs = FREE_BINDABLE;
        }
    } else if (bs == FREE_COOLDOWN) {
        int c = cool - 1;
        if (c <= 0) bs = FREE_BINDABLE;
        else cool = c;
    }
}

// The arc position of the head's foot along the segment (BindingDetectionSystem.java:380-384)
__device__ __forceinline__ float bind_arc(float mx, float my, float mz, float e1x, float e1y, float e1z, float e2x,
                                          float e2y, float e2z) {
    float r1x = e2x - e1x, r1y = e2y - e1y, r1z = e2z - e1z;
    float denom = r1x * r1x + r1y * r1y + r1z * r1z;
    float numer = (mx - e1x) * r1x + (my - e1y) * r1y + (mz - e1z) * r1z;
    return numer / (float)sqrt((double)denom);
}

// BindingDetectionSystem.bindNearest body (BindingDetectionSystem.java:367-387): strict < keeps the
// first candidate in list order on ties.
__device__ __forceinline__ void bind_nearest_one(int m, int nM, int nSeg, const float *__restrict__ head,
                                                 const float *__restrict__ hu, const float *__restrict__ ru,
                                                 const float *__restrict__ e1, const float *__restrict__ e2,
                                                 const int *__restrict__ reachSeg, int reachCount, int maxCand,
                                                 int &bs, float &arc, float myoColTol, float alignTol,
                                                 const int *__restrict__ seedNode = nullptr) {
    if (bs != FREE_BINDABLE) return;
    float mx = head[m], my = head[nM + m], mz = head[2 * nM + m];
    float mux = hu[m], muy = hu[nM + m], muz = hu[2 * nM + m];
    float rux = ru[m], ruy = ru[nM + m], ruz = ru[2 * nM + m];
    int cnt = reachCount > maxCand ? maxCand : reachCount;
    int bestSeg = -1;
    float bestD = 1.0e30f, bestArc = 0.0f;
    for (int k = 0; k < cnt; k++) {
        int s = reachSeg[(size_t)m * maxCand + k];
        if (seedNode && seedNode[s] >= 0) continue;   // node-held tip (bindNearestNodeAware)
        float e1x = e1[s], e1y = e1[nSeg + s], e1z = e1[2 * nSeg + s];
        float e2x = e2[s], e2y = e2[nSeg + s], e2z = e2[2 * nSeg + s];
        float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, e1x, e1y, e1z, e2x, e2y, e2z, myoColTol, alignTol);
        if (d >= 0.0f && d < bestD) {
            bestD = d;
            bestSeg = s;
            bestArc = bind_arc(mx, my, mz, e1x, e1y, e1z, e2x, e2y, e2z);
        }
    }
    if (bestSeg >= 0) {
        bs = bestSeg;
        arc = bestArc;
    }
}

struct NucParams {
    float dt, atpOn, onATP, offATP, onPi, offPi, onADP, offADP;
};

// NucleotideCycleSystem.cycle body (NucleotideCycleSystem.java:43-68)
__device__ __forceinline__ int nucleotide_one(int m, int state, bool bound, const float *__restrict__ forceDotHist,
                                              const NucParams &p, int step, int seed, int rng, uint32_t stream) {
    float u = uniform01(event_bits(rng, (uint32_t)m, step, seed, stream, 0x4E55u));
    if (state == NUC_NONE) {
        if (u < p.atpOn * p.dt) state = NUC_ATP;
    } else if (state == NUC_ATP) {
        float rate = bound ? p.onATP : p.offATP;
        if (u < rate * p.dt) state = NUC_ADPPI;
    } else if (state == NUC_ADPPI) {
        float rate = bound ? p.onPi : p.offPi;
        if (u < rate * p.dt) state = NUC_ADP;
    } else {
        float avg = 0.0f;
        for (int k = 0; k < 10; k++) avg += forceDotHist[m * 10 + k];
        avg *= 0.1f;
        if (avg <= 0.0f) {
            float rate = bound ? p.onADP : p.offADP;
            if (u < rate * p.dt) state = NUC_NONE;
        }
    }
    return state;
}

struct XbParams {
    double myoSpring, j1FMT, dt, headLen, xbias;
};

// CrossBridgeSystem.bondForces body for one bound motor (CrossBridgeSystem.java:78-140) -> 13 floats
__device__ __forceinline__ void bond_one(int m, int nM, int nSeg, int s, const float *__restrict__ mc,
                                         const float *__restrict__ mu, const float *__restrict__ my,
                                         const float *__restrict__ mbRG, const float *__restrict__ fc,
                                         const float *__restrict__ fu, const float *__restrict__ fy,
                                         const float *__restrict__ fbRG, const float *__restrict__ fLen, float bindArc,
                                         int nuc, const XbParams &xp, float out[13]) {
    const double DEG2RAD = SBX_PI_D / 180.0, RAD2DEG = 180.0 / SBX_PI_D;
    int nB = 3 * nM;
    int h = 3 * m + 2;
    double hcx = mc[h], hcy = mc[nB + h], hcz = mc[2 * nB + h];
    double hux = mu[h], huy = mu[nB + h], huz = mu[2 * nB + h];
    double hyx = my[h], hyy = my[nB + h], hyz = my[2 * nB + h];
    double hbRGx = mbRG[h], hbRGy = mbRG[nB + h];
    double htipx = hcx + 0.5 * xp.headLen * hux, htipy = hcy + 0.5 * xp.headLen * huy,
           htipz = hcz + 0.5 * xp.headLen * huz;
    double scx = fc[s], scy = fc[nSeg + s], scz = fc[2 * nSeg + s];
    double sux = fu[s], suy = fu[nSeg + s], suz = fu[2 * nSeg + s];
    double syx = fy[s], syy = fy[nSeg + s], syz = fy[2 * nSeg + s];
    double sbRGx = fbRG[s], sbRGy = fbRG[nSeg + s];
    double slen = fLen[s];
    double aOff = bindArc - 0.5 * slen;
    double apx = scx + aOff * sux, apy = scy + aOff * suy, apz = scz + aOff * suz;
    double dx = apx - htipx, dy = apy - htipy, dz = apz - htipz;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    double fmag = xp.myoSpring * dist;
    double Fx = 0, Fy = 0, Fz = 0;
    if (dist > 0.0) { double inv = fmag / dist; Fx = inv * dx; Fy = inv * dy; Fz = inv * dz; }
    double RHx = (htipx - hcx) * 1e-6, RHy = (htipy - hcy) * 1e-6, RHz = (htipz - hcz) * 1e-6;
    double THx = RHy * Fz - RHz * Fy, THy = RHz * Fx - RHx * Fz, THz = RHx * Fy - RHy * Fx;
    double RSx = (apx - scx) * 1e-6, RSy = (apy - scy) * 1e-6, RSz = (apz - scz) * 1e-6;
    double nFx = -Fx, nFy = -Fy, nFz = -Fz;
    double TSx = RSy * nFz - RSz * nFy, TSy = RSz * nFx - RSx * nFz, TSz = RSx * nFy - RSy * nFx;
    double restF9 = (nuc != NUC_ADPPI) ? 120.0 : 90.0;
    double t9x = suy * huz - suz * huy, t9y = suz * hux - sux * huz, t9z = sux * huy - suy * hux;
    double m9 = t9x * t9x + t9y * t9y + t9z * t9z;
    double T9x = 0, T9y = 0, T9z = 0;
    if (m9 > 1.0e-30) {
        double im = 1.0 / sqrt(m9);
        t9x *= im; t9y *= im; t9z *= im;
        double dot = sux * hux + suy * huy + suz * huz;
        if (dot > 1) dot = 1;
        if (dot < -1) dot = -1;
        double angD = accurate_acos(dot) * RAD2DEG - restF9;
        double tm = xp.j1FMT * DEG2RAD * angD / ((1.0 / hbRGy + 1.0 / sbRGy) * xp.dt);
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
        double ang = accurate_acos(dot) * RAD2DEG;
        double tm = xp.j1FMT * DEG2RAD * ang / ((1.0 / hbRGx + 1.0 / sbRGx) * xp.dt);
        T10x = tm * t10x; T10y = tm * t10y; T10z = tm * t10z;
    }
    out[0] = (float)Fx; out[1] = (float)Fy; out[2] = (float)Fz;
    out[3] = (float)(THx - T9x - T10x);
    out[4] = (float)(THy - T9y - T10y);
    out[5] = (float)(THz - T9z - T10z);
    out[6] = (float)(nFx - xp.xbias); out[7] = (float)nFy; out[8] = (float)nFz;
    out[9] = (float)(TSx + T9x + T10x);
    out[10] = (float)(TSy + T9y + T10y);
    out[11] = (float)(TSz + T9z + T10z);
    out[12] = (float)(Fx * sux + Fy * suy + Fz * suz);
}

// MotorJointSystem.moveC (MotorJointSystem.java:61-70)
__device__ __forceinline__ double move_c(double ux, double uy, double uz, double lx, double ly, double lz,
                                         double bTGx, double bTGy, double bRGy, double lenUm) {
    double cosB = ux * lx + uy * ly + uz * lz;
    if (cosB > 1.0) cosB = 1.0;
    if (cosB < -1.0) cosB = -1.0;
    double cosB2 = cosB * cosB;
    double cosA2 = 1.0 - cosB2;
    double lSq = 1.0e-12 * lenUm * lenUm;
    return cosB2 / bTGx + cosA2 / bTGy + lSq * cosA2 / (4.0 * bRGy);
}

struct JointParams {
    double dt, j1FracMove, j1FracR, j1FracMoveTorq, j2FracMove, j2FracR, j2FracMoveTorq, j2Rest, stallPN;
};

// MotorJointSystem.joints body for sub-body s (MotorJointSystem.java:86-196); adds into f, t (double)
__device__ __forceinline__ void joints_one(int s, int nB, const float *__restrict__ coord,
                                           const float *__restrict__ uVec, const float *__restrict__ segLength,
                                           const float *__restrict__ bTG, const float *__restrict__ bRG, int nuc,
                                           const JointParams &jp, double f[3], double t[3]) {
    const double DEG2RAD = SBX_PI_D / 180.0, RAD2DEG = 180.0 / SBX_PI_D;
    int m = s / 3;
    int role = s - 3 * m;
    double j1Rest = (nuc != NUC_ADPPI) ? 60.0 : 0.0;
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
    if (role == 0 || role == 1) {
        double aEx = rcx + 0.5 * rlen * rux, aEy = rcy + 0.5 * rlen * ruy, aEz = rcz + 0.5 * rlen * ruz;
        double bEx = lcx - 0.5 * llen * lux, bEy = lcy - 0.5 * llen * luy, bEz = lcz - 0.5 * llen * luz;
        double dx = aEx - bEx, dy = aEy - bEy, dz = aEz - bEz;
        double strain = sqrt(dx * dx + dy * dy + dz * dz);
        double inv = (strain > 0.0) ? 1.0 / strain : 0.0;
        double lx = dx * inv, ly = dy * inv, lz = dz * inv;
        double mcRod = move_c(rux, ruy, ruz, lx, ly, lz, bTG[rod], bTG[nB + rod], bRG[nB + rod], rlen);
        double mcLever = move_c(lux, luy, luz, lx, ly, lz, bTG[lever], bTG[nB + lever], bRG[nB + lever], llen);
        double denom = jp.dt * (mcRod + mcLever);
        double forceMag = (denom > 0.0) ? (jp.j2FracMove * 1.0e-6 * strain / denom) : 0.0;
        double tvx = ruy * luz - ruz * luy, tvy = ruz * lux - rux * luz, tvz = rux * luy - ruy * lux;
        double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
        double torsionMag = 0.0;
        if (tvm2 > 1.0e-30) {
            double im = 1.0 / sqrt(tvm2);
            tvx *= im; tvy *= im; tvz *= im;
            double dotV = rux * lux + ruy * luy + ruz * luz;
            if (dotV > 1.0) dotV = 1.0;
            if (dotV < -1.0) dotV = -1.0;
            double ang = accurate_acos(dotV) * RAD2DEG;
            double invBRG = 1.0 / bRG[nB + rod] + 1.0 / bRG[nB + lever];
            torsionMag = jp.j2FracMoveTorq * DEG2RAD * (ang - jp.j2Rest) / (invBRG * jp.dt);
        }
        if (role == 0) {
            double Fx = -forceMag * lx, Fy = -forceMag * ly, Fz = -forceMag * lz;
            f[0] += Fx; f[1] += Fy; f[2] += Fz;
            double Rs = 0.5e-6 * rlen * jp.j2FracR, Rx = Rs * rux, Ry = Rs * ruy, Rz = Rs * ruz;
            t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;
            t[0] += tvx * torsionMag; t[1] += tvy * torsionMag; t[2] += tvz * torsionMag;
        } else {
            double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
            f[0] += Fx; f[1] += Fy; f[2] += Fz;
            double Rs = -0.5e-6 * llen * jp.j2FracR, Rx = Rs * lux, Ry = Rs * luy, Rz = Rs * luz;
            t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;
            t[0] -= tvx * torsionMag; t[1] -= tvy * torsionMag; t[2] -= tvz * torsionMag;
        }
    }
    if (role == 1 || role == 2) {
        double aEx = lcx + 0.5 * llen * lux, aEy = lcy + 0.5 * llen * luy, aEz = lcz + 0.5 * llen * luz;
        double bEx = hcx - 0.5 * hlen * hux, bEy = hcy - 0.5 * hlen * huy, bEz = hcz - 0.5 * hlen * huz;
        double dx = aEx - bEx, dy = aEy - bEy, dz = aEz - bEz;
        double strain = sqrt(dx * dx + dy * dy + dz * dz);
        double inv = (strain > 0.0) ? 1.0 / strain : 0.0;
        double lx = dx * inv, ly = dy * inv, lz = dz * inv;
        double mcLever = move_c(lux, luy, luz, lx, ly, lz, bTG[lever], bTG[nB + lever], bRG[nB + lever], llen);
        double mcHead = move_c(hux, huy, huz, lx, ly, lz, bTG[head], bTG[nB + head], bRG[nB + head], hlen);
        double denom = jp.dt * (mcLever + mcHead);
        double forceMag = (denom > 0.0) ? (jp.j1FracMove * 1.0e-6 * strain / denom) : 0.0;
        double tvx = luy * huz - luz * huy, tvy = luz * hux - lux * huz, tvz = lux * huy - luy * hux;
        double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
        double torsionMag = 0.0;
        if (tvm2 > 1.0e-30) {
            double im = 1.0 / sqrt(tvm2);
            tvx *= im; tvy *= im; tvz *= im;
            double dotV = lux * hux + luy * huy + luz * huz;
            if (dotV > 1.0) dotV = 1.0;
            if (dotV < -1.0) dotV = -1.0;
            double ang = accurate_acos(dotV) * RAD2DEG;
            double invBRG = 1.0 / bRG[nB + lever] + 1.0 / bRG[nB + head];
            torsionMag = jp.j1FracMoveTorq * DEG2RAD * (ang - j1Rest) / (invBRG * jp.dt);
            double maxMag = jp.stallPN * 0.5 * hlen * 1.0e-18;
            if (torsionMag > maxMag) torsionMag = maxMag;
        }
        if (role == 1) {
            double Fx = -forceMag * lx, Fy = -forceMag * ly, Fz = -forceMag * lz;
            f[0] += Fx; f[1] += Fy; f[2] += Fz;
            double Rs = 0.5e-6 * llen * jp.j1FracR, Rx = Rs * lux, Ry = Rs * luy, Rz = Rs * luz;
            t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;
            t[0] += tvx * torsionMag; t[1] += tvy * torsionMag; t[2] += tvz * torsionMag;
        } else {
            double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
            f[0] += Fx; f[1] += Fy; f[2] += Fz;
            double Rs = -0.5e-6 * hlen * jp.j1FracR, Rx = Rs * hux, Ry = Rs * huy, Rz = Rs * huz;
            t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;
            t[0] -= tvx * torsionMag; t[1] -= tvy * torsionMag; t[2] -= tvz * torsionMag;
        }
    }
}

struct DimerParams {
    double dt, fracMove, leverTorq, leverDeg, rodLen;
};

// DimerCouplingSystem.couple, one end pairing (DimerCouplingSystem.java:94-160): variant 0 couples rod A's
// end1 side (endSel -1), variant 1 its end2 side; rod B's end follows the parallel flag.  FA is the force on
// rod A (rod B takes -FA), TA and TB the torques on rods A and B; false when the two ends coincide (no force).
__device__ __forceinline__ bool dimer_variant(int variant, int mA, int mB, int par, int nB,
                                              const float *__restrict__ coord, const float *__restrict__ uVec,
                                              const float *__restrict__ segLength, const float *__restrict__ bTG,
                                              const float *__restrict__ bRG, const DimerParams &dp, double FA[3],
                                              double TA[3], double TB[3]) {
    int rodA = 3 * mA, rodB = 3 * mB;
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
    if (!(strain > 0.0)) return false;
    double inv = 1.0 / strain;
    double lx = dx * inv, ly = dy * inv, lz = dz * inv;
    double mcA = move_c(aUx, aUy, aUz, lx, ly, lz, bTG[rodA], bTG[nB + rodA], bRG[nB + rodA], aLen);
    double mcB = move_c(bUx, bUy, bUz, lx, ly, lz, bTG[rodB], bTG[nB + rodB], bRG[nB + rodB], bLen);
    double denom = dp.dt * (mcA + mcB);
    double forceMag = (denom > 0.0) ? (dp.fracMove * 1.0e-6 * strain / denom) : 0.0;
    double FAx = forceMag * lx, FAy = forceMag * ly, FAz = forceMag * lz;
    double RsA = 0.5e-6 * dp.rodLen * endSelA;
    TA[0] = (RsA * aUy) * FAz - (RsA * aUz) * FAy;
    TA[1] = (RsA * aUz) * FAx - (RsA * aUx) * FAz;
    TA[2] = (RsA * aUx) * FAy - (RsA * aUy) * FAx;
    double RsB = 0.5e-6 * dp.rodLen * endSelB;
    TB[0] = (RsB * bUy) * (-FAz) - (RsB * bUz) * (-FAy);
    TB[1] = (RsB * bUz) * (-FAx) - (RsB * bUx) * (-FAz);
    TB[2] = (RsB * bUx) * (-FAy) - (RsB * bUy) * (-FAx);
    FA[0] = FAx;
    FA[1] = FAy;
    FA[2] = FAz;
    return true;
}

// The levers' torsion of a parallel dimer whose heads are not both bound (DimerCouplingSystem.java:162-182):
// the torque on lever A (lever B takes its negative); false when the levers are parallel.
__device__ __forceinline__ bool dimer_lever(int mA, int mB, int nB, const float *__restrict__ uVec,
                                            const float *__restrict__ bRG, const DimerParams &dp, double TL[3]) {
    int leverA = 3 * mA + 1, leverB = 3 * mB + 1;
    double u1x = uVec[leverA], u1y = uVec[nB + leverA], u1z = uVec[2 * nB + leverA];
    double u2x = uVec[leverB], u2y = uVec[nB + leverB], u2z = uVec[2 * nB + leverB];
    double tvx = u1y * u2z - u1z * u2y, tvy = u1z * u2x - u1x * u2z, tvz = u1x * u2y - u1y * u2x;
    double tvm2 = tvx * tvx + tvy * tvy + tvz * tvz;
    if (!(tvm2 > 1.0e-30)) return false;
    double im = 1.0 / sqrt(tvm2);
    tvx *= im; tvy *= im; tvz *= im;
    double dot = u1x * u2x + u1y * u2y + u1z * u2z;
    if (dot > 1.0) dot = 1.0;
    if (dot < -1.0) dot = -1.0;
    double ang = accurate_acos(dot) * 180.0 / SBX_PI_D;
    double invBRG = 1.0 / bRG[nB + leverA] + 1.0 / bRG[nB + leverB];
    double torsionMag = dp.leverTorq * (SBX_PI_D / 180.0) * (ang - dp.leverDeg) / (invBRG * dp.dt);
    TL[0] = tvx * torsionMag;
    TL[1] = tvy * torsionMag;
    TL[2] = tvz * torsionMag;
    return true;
}

// DimerCouplingSystem.couple body for one dimer (DimerCouplingSystem.java:94-182): read-modify-write
// of the two rods' and two levers' accumulators (the dimer owns its two motors).
__device__ __forceinline__ void dimer_one(int mA, int mB, int par, bool bothBound, int nB,
                                          const float *__restrict__ coord, const float *__restrict__ uVec,
                                          const float *__restrict__ segLength, const float *__restrict__ bTG,
                                          const float *__restrict__ bRG, float *__restrict__ F, float *__restrict__ T,
                                          const DimerParams &dp) {
    int rodA = 3 * mA, rodB = 3 * mB;
    for (int variant = 0; variant < 2; variant++) {
        double FA[3], TA[3], TB[3];
        if (!dimer_variant(variant, mA, mB, par, nB, coord, uVec, segLength, bTG, bRG, dp, FA, TA, TB)) continue;
        for (int a = 0; a < 3; a++) {
            F[a * nB + rodA] = (float)(F[a * nB + rodA] + FA[a]);
            T[a * nB + rodA] = (float)(T[a * nB + rodA] + TA[a]);
        }
        for (int a = 0; a < 3; a++) {
            F[a * nB + rodB] = (float)(F[a * nB + rodB] - FA[a]);
            T[a * nB + rodB] = (float)(T[a * nB + rodB] + TB[a]);
        }
    }
    double TL[3];
    if (par && !bothBound && dimer_lever(mA, mB, nB, uVec, bRG, dp, TL)) {
        int leverA = 3 * mA + 1, leverB = 3 * mB + 1;
        for (int a = 0; a < 3; a++) T[a * nB + leverA] = (float)(T[a * nB + leverA] + TL[a]);
        for (int a = 0; a < 3; a++) T[a * nB + leverB] = (float)(T[a * nB + leverB] - TL[a]);
    }
}

struct MiniParams {
    double dt, fracMove, align;
};

// NodeSystem.tether body for one attachment (NodeSystem.java:66-126; tetherBoundStiffen :185-230 differs only
// in the coefficient the caller passes): the rod's end1 pulled toward the node-frame surface point
// c + ru u + ry y + rz (u x y) by a damping-limited spring; a dimer (atEnd1) also takes the end1 torque on the
// rod and the torque of the reaction at the body-frame offset (ru, ry, rz) on the node.
__device__ __forceinline__ void node_tether_one(int a, int nMB, int nN, int nA, int k, int rod, bool atEnd1,
                                                double coeff, double dt, const float *__restrict__ mC,
                                                const float *__restrict__ mU, const float *__restrict__ mLen,
                                                const float *__restrict__ mBTG, float *__restrict__ mF,
                                                float *__restrict__ mT, const float *__restrict__ nC,
                                                const float *__restrict__ nU, const float *__restrict__ nY,
                                                const float *__restrict__ nInv, const float *__restrict__ radial,
                                                float *__restrict__ nodeData) {
    double ncx = nC[k], ncy = nC[nN + k], ncz = nC[2 * nN + k];
    double nux = nU[k], nuy = nU[nN + k], nuz = nU[2 * nN + k];
    double nyx = nY[k], nyy = nY[nN + k], nyz = nY[2 * nN + k];
    double nzx = nuy * nyz - nuz * nyy, nzy = nuz * nyx - nux * nyz, nzz = nux * nyy - nuy * nyx;
    double ru = radial[a], ry = radial[nA + a], rz = radial[2 * nA + a];
    double pax = ncx + ru * nux + ry * nyx + rz * nzx;
    double pay = ncy + ru * nuy + ry * nyy + rz * nzy;
    double paz = ncz + ru * nuz + ry * nyz + rz * nzz;
    double rcx = mC[rod], rcy = mC[nMB + rod], rcz = mC[2 * nMB + rod];
    double rux = mU[rod], ruy = mU[nMB + rod], ruz = mU[2 * nMB + rod];
    double rlen = mLen[rod];
    double re1x = rcx - 0.5 * rlen * rux, re1y = rcy - 0.5 * rlen * ruy, re1z = rcz - 0.5 * rlen * ruz;
    double fnx = 0, fny = 0, fnz = 0, tnx = 0, tny = 0, tnz = 0;
    double dx = pax - re1x, dy = pay - re1y, dz = paz - re1z;
    double strain = sqrt(dx * dx + dy * dy + dz * dz);
    if (strain > 0.0) {
        double inv = 1.0 / strain;
        double lx = dx * inv, ly = dy * inv, lz = dz * inv;
        double denom = dt * (1.0 / mBTG[nMB + rod] + nInv[k]);
        double forceMag = (denom > 0.0) ? (coeff * 1.0e-6 * strain / denom) : 0.0;
        double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
        mF[rod] = (float)(mF[rod] + Fx);
        mF[nMB + rod] = (float)(mF[nMB + rod] + Fy);
        mF[2 * nMB + rod] = (float)(mF[2 * nMB + rod] + Fz);
        if (atEnd1) {
            double Rx = -0.5 * rlen * rux * 1.0e-6, Ry = -0.5 * rlen * ruy * 1.0e-6, Rz = -0.5 * rlen * ruz * 1.0e-6;
            mT[rod] = (float)(mT[rod] + (Ry * Fz - Rz * Fy));
            mT[nMB + rod] = (float)(mT[nMB + rod] + (Rz * Fx - Rx * Fz));
            mT[2 * nMB + rod] = (float)(mT[2 * nMB + rod] + (Rx * Fy - Ry * Fx));
        }
        fnx = -Fx; fny = -Fy; fnz = -Fz;
        if (atEnd1) {
            double Rnx = ru * 1.0e-6, Rny = ry * 1.0e-6, Rnz = rz * 1.0e-6;
            tnx = Rny * fnz - Rnz * fny;
            tny = Rnz * fnx - Rnx * fnz;
            tnz = Rnx * fny - Rny * fnx;
        }
    }
    float *o = nodeData + (size_t)MINI_STRIDE * a;
    o[0] = (float)fnx; o[1] = (float)fny; o[2] = (float)fnz;
    o[3] = (float)tnx; o[4] = (float)tny; o[5] = (float)tnz;
}

// the coefficient of one attachment: |attachCoeffK| (sign = atEnd1), or boundCoeff for a bound motor under
// tetherBoundStiffen (boundSeg non-null; NodeSystem.java:188-191)
__device__ __forceinline__ double node_coeff(float cs, int m, const int *__restrict__ boundSeg, double boundCoeff,
                                             bool &atEnd1) {
    atEnd1 = cs > 0.0f;
    double coeff = atEnd1 ? (double)cs : -(double)cs;
    if (boundSeg && boundSeg[m] >= 0) coeff = boundCoeff;
    return coeff;
}

// MiniFilamentSystem.tether body for one tethered dimer (MiniFilamentSystem.java:76-130)
__device__ __forceinline__ void tether_one(int d, int nMB, int nBb, int bb, int rodA, double ax,
                                           const float *__restrict__ mC, const float *__restrict__ mU,
                                           const float *__restrict__ mLen, const float *__restrict__ mBTG,
                                           const float *__restrict__ mBRG, float *__restrict__ mF,
                                           float *__restrict__ mT, const float *__restrict__ bbC,
                                           const float *__restrict__ bbU, const float *__restrict__ bbInv,
                                           float *__restrict__ miniData, const MiniParams &mp) {
    double bcx = bbC[bb], bcy = bbC[nBb + bb], bcz = bbC[2 * nBb + bb];
    double bux = bbU[bb], buy = bbU[nBb + bb], buz = bbU[2 * nBb + bb];
    double pax = bcx + ax * bux, pay = bcy + ax * buy, paz = bcz + ax * buz;
    double rcx = mC[rodA], rcy = mC[nMB + rodA], rcz = mC[2 * nMB + rodA];
    double rux = mU[rodA], ruy = mU[nMB + rodA], ruz = mU[2 * nMB + rodA];
    double rlen = mLen[rodA];
    double re1x = rcx - 0.5 * rlen * rux, re1y = rcy - 0.5 * rlen * ruy, re1z = rcz - 0.5 * rlen * ruz;
    double fbx = 0, fby = 0, fbz = 0, tbx = 0, tby = 0, tbz = 0;
    double dx = pax - re1x, dy = pay - re1y, dz = paz - re1z;
    double strain = sqrt(dx * dx + dy * dy + dz * dz);
    if (strain > 0.0) {
        double inv = 1.0 / strain;
        double lx = dx * inv, ly = dy * inv, lz = dz * inv;
        double denom = mp.dt * (1.0 / mBTG[nMB + rodA] + bbInv[bb]);
        double forceMag = (denom > 0.0) ? (mp.fracMove * 1.0e-6 * strain / denom) : 0.0;
        double Frx = forceMag * lx, Fry = forceMag * ly, Frz = forceMag * lz;
        mF[rodA] = (float)(mF[rodA] + Frx);
        mF[nMB + rodA] = (float)(mF[nMB + rodA] + Fry);
        mF[2 * nMB + rodA] = (float)(mF[2 * nMB + rodA] + Frz);
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
        double angDeg = accurate_acos(dot) * 180.0 / SBX_PI_D;
        double denomR = (1.0 / mBRG[nMB + rodA] + bbInv[nBb + bb]) * mp.dt;
        double tmag = mp.align * (SBX_PI_D / 180.0) * angDeg / denomR;
        mT[rodA] = (float)(mT[rodA] + tvx * tmag);
        mT[nMB + rodA] = (float)(mT[nMB + rodA] + tvy * tmag);
        mT[2 * nMB + rodA] = (float)(mT[2 * nMB + rodA] + tvz * tmag);
        tbx = -tvx * tmag; tby = -tvy * tmag; tbz = -tvz * tmag;
    }
    float *o = miniData + (size_t)MINI_STRIDE * d;
    o[0] = (float)fbx; o[1] = (float)fby; o[2] = (float)fbz;
    o[3] = (float)tbx; o[4] = (float)tby; o[5] = (float)tbz;
}

// CrossBridgeSystem.segGather / MiniFilamentSystem.backboneGather range sum (FP64, CSR order)
__device__ __forceinline__ void csr_payload_sum(int s, const int *__restrict__ offsets, const int *__restrict__ list,
                                                const float *__restrict__ data, int stride, int off, double f[3],
                                                double t[3]) {
    int end = offsets[s + 1];
    for (int k = offsets[s]; k < end; k++) {
        const float *d = data + (size_t)list[k] * stride + off;
        f[0] += d[0]; f[1] += d[1]; f[2] += d[2];
        t[0] += d[3]; t[1] += d[4]; t[2] += d[5];
    }
}

KinParams kin_params(const float *kp);
NucParams nuc_params(const float *np);
XbParams xb_params(const float *xp);
JointParams joint_params(const float *jp);
DimerParams dimer_params(const float *dp);
MiniParams mini_params(const float *mp);

}  // namespace sbx
