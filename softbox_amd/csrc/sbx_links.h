// sbx_links.h -- crosslink force laws (CrosslinkerSystem.java) as device functions shared by
// the stand-alone entry points and the fused network step.
#pragma once
#include "sbx_device.h"

namespace sbx {

constexpr int XL_STRIDE = 12;
constexpr int STRAIN_WIN = 10;
constexpr int TORQUE_WIN = 5;

struct LinkParams {
    double restLength, fracMoveBase, dt;                 // xlParams
    double offConst, offCoeff, offExp, offDt, offRest;  // offParams
    double torsionSpring;
    int torsionOn;
};

// What the link laws read of one segment: centre, axis, end1 and the parallel drag.
struct LinkEnd {
    float c[3], u[3], e1[3], gpar;
};

// planar store arrays (the reference layout, stride nSeg)
__device__ __forceinline__ LinkEnd link_end_planar(int s, int nSeg, const float *__restrict__ coord,
                                                   const float *__restrict__ uVec, const float *__restrict__ end1,
                                                   const float *__restrict__ bTransGam) {
    LinkEnd e;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        e.c[k] = coord ? coord[k * nSeg + s] : 0.0f;
        e.u[k] = uVec[k * nSeg + s];
        e.e1[k] = end1[k * nSeg + s];
    }
    e.gpar = bTransGam ? bTransGam[s] : 0.0f;
    return e;
}

// packed mirror written by the fused step: 3 float4 per segment {c, gpar} {u, 0} {end1, 0}, so one
// random endpoint gather touches 48 contiguous bytes instead of 10 planar streams
constexpr int LINK_PACK = 3;
__device__ __forceinline__ LinkEnd link_end_packed(int s, const float4 *__restrict__ pack) {
    float4 a = pack[LINK_PACK * s], b = pack[LINK_PACK * s + 1], c = pack[LINK_PACK * s + 2];
    LinkEnd e;
    e.c[0] = a.x; e.c[1] = a.y; e.c[2] = a.z; e.gpar = a.w;
    e.u[0] = b.x; e.u[1] = b.y; e.u[2] = b.z;
    e.e1[0] = c.x; e.e1[1] = c.y; e.e1[2] = c.z;
    return e;
}

// attachment points p1 = end1(A) + loc1*u(A), p2 = end1(B) + loc2*u(B) (CrosslinkerSystem.java:72-83)
__device__ __forceinline__ void link_points(const LinkEnd &A, const LinkEnd &B, float l1f, float l2f, double p1[3],
                                            double p2[3]) {
    double l1 = l1f, l2 = l2f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p1[k] = A.e1[k] + l1 * (double)A.u[k];
        p2[k] = B.e1[k] + l2 * (double)B.u[k];
    }
}

// CrosslinkerSystem.unbind body (CrosslinkerSystem.java:245-279) for one ACTIVE link; returns the
// new linkState (unchanged or -1) and updates the strain ring in place.  k indexes the rings, key is
// the link index the draw is keyed by (the same, except in a batch: the replica's own link index).
__device__ __forceinline__ int link_unbind_one(int k, uint32_t key, const LinkEnd &A, const LinkEnd &B, float l1,
                                               float l2,
                                               float *__restrict__ strainHist, int *__restrict__ strainPlace,
                                               const LinkParams &lp, int step, int seed, int rng, uint32_t stream,
                                               int state) {
    double p1[3], p2[3];
    link_points(A, B, l1, l2, p1, p2);
    double lvx = p2[0] - p1[0], lvy = p2[1] - p1[1], lvz = p2[2] - p1[2];
    double linkLength = sqrt(lvx * lvx + lvy * lvy + lvz * lvz);
    double stretch = linkLength - lp.offRest;
    if (stretch < 0.0) stretch = 0.0;
    double strain = stretch / lp.offRest;
    int base = k * STRAIN_WIN;
    int p = strainPlace[k];
    strainHist[base + p] = (float)strain;
    strainPlace[k] = (p + 1) % STRAIN_WIN;
    double ave = 0.0;
#pragma unroll
    for (int j = 0; j < STRAIN_WIN; j++) ave += strainHist[base + j];
    ave /= STRAIN_WIN;
    double kOff = lp.offConst + lp.offCoeff * exp(ave * lp.offExp);
    double pBreak = kOff * lp.offDt;
    uint32_t h = event_bits(rng, key, step, seed, stream, 0x584C4B42u);
    float u = uniform01(h);
    return ((double)u < pBreak) ? -1 : state;
}

// CrosslinkerSystem.linkForces body (CrosslinkerSystem.java:72-122) -> 12-float payload
__device__ __forceinline__ void link_force_one(const LinkEnd &A, const LinkEnd &B, float l1, float l2, int ctA,
                                               int ctB, const LinkParams &lp, float out[12]) {
    double p1[3], p2[3];
    link_points(A, B, l1, l2, p1, p2);
    double lvx = p2[0] - p1[0], lvy = p2[1] - p1[1], lvz = p2[2] - p1[2];
    double linkLength = sqrt(lvx * lvx + lvy * lvy + lvz * lvz);
    double stretch = linkLength - lp.restLength;
    if (stretch < 0.0) stretch = 0.0;
    int maxLinks = ctA > ctB ? ctA : ctB;
    if (maxLinks < 1) maxLinks = 1;
    double fracMove = lp.fracMoveBase / maxLinks;
    double g1 = A.gpar;
    double g2 = B.gpar;
    double curForceMag = (fracMove * 1.0e-6 * stretch / lp.dt) / (1.0 / g1 + 1.0 / g2);
    double Fx = curForceMag * lvx, Fy = curForceMag * lvy, Fz = curForceMag * lvz;
    double rax = (p1[0] - A.c[0]) * 1e-6;
    double ray = (p1[1] - A.c[1]) * 1e-6;
    double raz = (p1[2] - A.c[2]) * 1e-6;
    double TAx = ray * Fz - raz * Fy;
    double TAy = raz * Fx - rax * Fz;
    double TAz = rax * Fy - ray * Fx;
    double nFx = -Fx, nFy = -Fy, nFz = -Fz;
    double rbx = (p2[0] - B.c[0]) * 1e-6;
    double rby = (p2[1] - B.c[1]) * 1e-6;
    double rbz = (p2[2] - B.c[2]) * 1e-6;
    double TBx = rby * nFz - rbz * nFy;
    double TBy = rbz * nFx - rbx * nFz;
    double TBz = rbx * nFy - rby * nFx;
    out[0] = (float)Fx; out[1] = (float)Fy; out[2] = (float)Fz;
    out[3] = (float)TAx; out[4] = (float)TAy; out[5] = (float)TAz;
    out[6] = (float)nFx; out[7] = (float)nFy; out[8] = (float)nFz;
    out[9] = (float)TBx; out[10] = (float)TBy; out[11] = (float)TBz;
}

// CrosslinkerSystem.linkTorsion body (CrosslinkerSystem.java:97-127) for one ACTIVE link with valid
// keys; adds into the payload torque slots exactly as (float)(payload + T).
__device__ __forceinline__ void link_torsion_one(int k, const LinkEnd &A, const LinkEnd &B, int same,
                                                 float *__restrict__ torqueMagHist, int *__restrict__ torqueMagPlace,
                                                 const LinkParams &lp, float pay[12]) {
    double uax = A.u[0], uay = A.u[1], uaz = A.u[2];
    double ubx = B.u[0], uby = B.u[1], ubz = B.u[2];
    double vbx = same ? ubx : -ubx, vby = same ? uby : -uby, vbz = same ? ubz : -ubz;
    double tx = uay * vbz - uaz * vby, ty = uaz * vbx - uax * vbz, tz = uax * vby - uay * vbx;
    double m2 = tx * tx + ty * ty + tz * tz;
    if (!(m2 > 1.0e-30)) return;
    double im = 1.0 / sqrt(m2);
    tx *= im; ty *= im; tz *= im;
    double dot = uax * ubx + uay * uby + uaz * ubz;
    if (dot > 1.0) dot = 1.0;
    if (dot < -1.0) dot = -1.0;
    double angT = same ? fast_acos(dot) : fabs(fast_acos(dot) - SBX_PI_D);
    double curMag = lp.torsionSpring * angT;
    int base = k * TORQUE_WIN;
    int p = torqueMagPlace[k];
    torqueMagHist[base + p] = (float)curMag;
    torqueMagPlace[k] = (p + 1) % TORQUE_WIN;
    double avg = 0.0;
#pragma unroll
    for (int j = 0; j < TORQUE_WIN; j++) avg += torqueMagHist[base + j];
    avg /= TORQUE_WIN;
    double Tx = avg * tx, Ty = avg * ty, Tz = avg * tz;
    pay[3] = (float)(pay[3] + Tx);
    pay[4] = (float)(pay[4] + Ty);
    pay[5] = (float)(pay[5] + Tz);
    pay[9] = (float)(pay[9] - Tx);
    pay[10] = (float)(pay[10] - Ty);
    pay[11] = (float)(pay[11] - Tz);
}

// One side of segGatherA/B (CrosslinkerSystem.java:132-147): FP64 sum in CSR order, skip dead links.
__device__ __forceinline__ void gather_range(int start, int end, const int *__restrict__ idx,
                                             const float *__restrict__ xlinkData, const int *__restrict__ linkState,
                                             int off, double f[3], double t[3], int &active) {
    for (int k = start; k < end; k++) {
        int li = idx[k];
        if (linkState[li] < 0) continue;
        active++;
        const float *d = xlinkData + (size_t)li * XL_STRIDE + off;
        f[0] += d[0]; f[1] += d[1]; f[2] += d[2];
        t[0] += d[3]; t[1] += d[4]; t[2] += d[5];
    }
}

// segGatherA/B of segment s over its CSR range
__device__ __forceinline__ void gather_side(int s, const int *__restrict__ offsets, const int *__restrict__ idx,
                                            const float *__restrict__ xlinkData, const int *__restrict__ linkState,
                                            int off, double f[3], double t[3], int &active) {
    gather_range(offsets[s], offsets[s + 1], idx, xlinkData, linkState, off, f, t, active);
}

// CrosslinkerSystem.countActiveLinks for one segment (CrosslinkerSystem.java:694-706): ACTIVE links
// keyed to s on either side, read from the two static CSR ranges.
__device__ __forceinline__ int active_links(int s, const int *__restrict__ offA, const int *__restrict__ idxA,
                                            const int *__restrict__ offB, const int *__restrict__ idxB,
                                            const int *__restrict__ linkState) {
    int cnt = 0;
    for (int k = offA[s]; k < offA[s + 1]; k++) cnt += linkState[idxA[k]] >= 0 ? 1 : 0;
    for (int k = offB[s]; k < offB[s + 1]; k++) cnt += linkState[idxB[k]] >= 0 ? 1 : 0;
    return cnt;
}

LinkParams link_params(const float *xlParams, const float *offParams, const float *torsionParams);

}  // namespace sbx
