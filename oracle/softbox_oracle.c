/*
 * softbox_oracle.c -- CPU restatement of SoftBox's rigid-rod mechanics hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See softbox_oracle.h for the contract,
 * the list of parity pins and the rule that the product path never uses this file.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off -fno-fast-math (oracle/Makefile).
 * -ffp-contract=off keeps every a*b+c as two IEEE roundings, as Java does.
 */
#include "softbox_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793

/* ------------------------------------------------------------------ constants */
/* Constants.java:25-51 */
/* Built with -fopenmp (liboracle_softbox_omp.so, the all-cores CPU baseline) the per-item loops of
 * the reference's @Parallel Systems run on all host threads; every such loop writes only its own
 * item's slots (the reference's one-writer-per-slot rule), so the results do not change.  Without
 * OpenMP (liboracle_softbox.so, the checker) they are plain loops. */
#ifdef _OPENMP
#define SBO_PAR_FOR _Pragma("omp parallel for schedule(static)")
#define SBO_PAR_FOR_DYNAMIC _Pragma("omp parallel for schedule(dynamic, 256)")
#else
#define SBO_PAR_FOR
#define SBO_PAR_FOR_DYNAMIC
#endif

static const double K_BOLTZ = 1.380662e-23;
static const double K_TEMPK = 298.15;
static const double K_AETA = 0.1;
static const double K_ACTIN_MONO_RADIUS = 0.0054 / 2.0;
static const double K_RADIUS = 0.007 / 2.0;
static const double K_APARALLEL = -0.20;
static const double K_AORTHOG = 0.84;
static const double K_ATURNING = -0.662;
static const int K_STD_SEG_LENGTH = 32;

double sbo_const(int which) {
    switch (which) {
    case 0: return K_BOLTZ * K_TEMPK;
    case 1: return K_AETA;
    case 2: return K_ACTIN_MONO_RADIUS;
    case 3: return K_RADIUS;
    case 4: return K_APARALLEL;
    case 5: return K_AORTHOG;
    case 6: return K_ATURNING;
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------ drag */
/* DragTensorSystem.rodDragSI, DragTensorSystem.java:33-45.  aeta is an argument (the
 * reference reads Constants.aeta = 0.1 Pa s; the workloads here use water, 1e-3). */
void sbo_rod_drag_si(double lengthUm, double radiusUm, double aeta, double g[6]) {
    double LM = 1.0e-6 * lengthUm;
    double RM = radiusUm * 1.0e-6;
    double logT = log(LM / (2 * RM));
    g[0] = (2 * PI_D * aeta * LM) / (logT + K_APARALLEL);
    g[1] = (4 * PI_D * aeta * LM) / (logT + K_AORTHOG);
    g[2] = g[1];
    g[3] = 4 * PI_D * aeta * RM * RM * LM;
    g[4] = (PI_D * aeta * (LM * LM * LM)) / (3 * (logT + K_ATURNING));
    g[5] = g[4];
}

/* DragTensorSystem.sphereDragSI, DragTensorSystem.java:52-58 */
void sbo_sphere_drag_si(double radiusUm, double aeta, double g[6]) {
    double RM = radiusUm * 1.0e-6;
    double bTG = 6 * PI_D * aeta * RM;
    double bRG = 8 * PI_D * aeta * (RM * RM * RM);
    g[0] = g[1] = g[2] = bTG;
    g[3] = g[4] = g[5] = bRG;
}

/* DragTensorSystem.storeDrag, DragTensorSystem.java:61-69 */
static void store_drag(int n, int i, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff,
                       const double g[6], double kT) {
    for (int a = 0; a < 3; a++) {
        bTransGam[a * n + i] = (float)g[a];
        bRotGam[a * n + i] = (float)g[3 + a];
        if (bTransDiff) bTransDiff[a * n + i] = (float)(kT / g[a]);
        if (bRotDiff) bRotDiff[a * n + i] = (float)(kT / g[3 + a]);
    }
}

/* DragTensorSystem.run(FilamentStore), DragTensorSystem.java:71-97 */
void sbo_drag_run_filament(int n, const int *monomerCount, const int *end1NbrSlot, const int *end2NbrSlot,
                           float *segLength, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff,
                           double aeta, double kT) {
    const double halfmono = K_ACTIN_MONO_RADIUS;
    const int minMonomerCt = 30;
    for (int i = 0; i < n; i++) {
        double length = (monomerCount[i] + 1) * K_ACTIN_MONO_RADIUS;
        segLength[i] = (float)length;
        double minLength;
        if (end1NbrSlot[i] == -1 || end2NbrSlot[i] == -1) minLength = K_STD_SEG_LENGTH * halfmono;
        else minLength = minMonomerCt * halfmono;
        double asIfLength = length;
        if (asIfLength < minLength) asIfLength = minLength;
        double g[6];
        sbo_rod_drag_si(asIfLength, K_RADIUS, aeta, g);
        store_drag(n, i, bTransGam, bRotGam, bTransDiff, bRotDiff, g, kT);
    }
}

/* Drag for rods of a given length (the workloads set segLength directly, no monomer clamp;
 * same formula, DragTensorSystem.java:33-45 + :61-69). */
void sbo_drag_fixed_length(int n, double lengthUm, double radiusUm, double aeta, double kT,
                           float *segLength, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff) {
    double g[6];
    sbo_rod_drag_si(lengthUm, radiusUm, aeta, g);
    for (int i = 0; i < n; i++) {
        segLength[i] = (float)lengthUm;
        store_drag(n, i, bTransGam, bRotGam, bTransDiff, bRotDiff, g, kT);
    }
}

/* ------------------------------------------------------------------ random streams */
/* BrownianForceSystem.wangHash, BrownianForceSystem.java:35-42 (Java int arithmetic = uint32 wrap) */
uint32_t sbo_wang_hash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key bumps 0x9E3779B9 / 0xBB67AE85. */
void sbo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Reference keying (BrownianForceSystem.java:64): (i*1000003) ^ (step*999983) ^ (seed*7919), 32-bit wrap */
static uint32_t wang_base(uint32_t i, int32_t step, int32_t seed) {
    return (i * 1000003u) ^ ((uint32_t)step * 999983u) ^ ((uint32_t)seed * 7919u);
}

/* The six per-body-step draws.  WANG: salts of BrownianForceSystem.java:65-70.
 * PHILOX: counter (body, step, block, 0), key (seed, stream); block 0 gives h1..h4, block 1 gives h5,h6. */
void sbo_brownian_bits(int rng, uint32_t body, int32_t step, int32_t seed, uint32_t stream, uint32_t h[6]) {
    if (rng == SBO_RNG_WANG) {
        uint32_t base = wang_base(body, step, seed);
        h[0] = sbo_wang_hash(base);
        h[1] = sbo_wang_hash(base ^ 0x9e3779b9u);
        h[2] = sbo_wang_hash(base ^ 0x85ebca6bu);
        h[3] = sbo_wang_hash(base ^ 0xc2b2ae35u);
        h[4] = sbo_wang_hash(base ^ 0x517cc1b7u);
        h[5] = sbo_wang_hash(base ^ 0x1f0a7ed5u);
    } else {
        uint32_t key[2] = {(uint32_t)seed, stream};
        uint32_t c0[4] = {body, (uint32_t)step, 0u, 0u};
        uint32_t c1[4] = {body, (uint32_t)step, 1u, 0u};
        uint32_t o0[4], o1[4];
        sbo_philox4x32_10(c0, key, o0);
        sbo_philox4x32_10(c1, key, o1);
        h[0] = o0[0]; h[1] = o0[1]; h[2] = o0[2]; h[3] = o0[3];
        h[4] = o1[0]; h[5] = o1[1];
    }
}

/* Event draw.  WANG: CrosslinkerSystem.java:277 form (base ^ salt).  PHILOX: counter (item, step, salt, 2). */
uint32_t sbo_event_bits(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream, uint32_t salt) {
    if (rng == SBO_RNG_WANG) return sbo_wang_hash(wang_base(item, step, seed) ^ salt);
    uint32_t key[2] = {(uint32_t)seed, stream};
    uint32_t c[4] = {item, (uint32_t)step, salt, 2u};
    uint32_t o[4];
    sbo_philox4x32_10(c, key, o);
    return o[0];
}

/* BrownianForceSystem.java:72-77: (h >>> 1) / 2147483647.0f (int -> float rounds to nearest) */
float sbo_uniform(uint32_t h) {
    return (float)(int32_t)(h >> 1) / 2147483647.0f;
}

static float fmaxf_java(float a, float b) { return a >= b ? a : b; }

/* ------------------------------------------------------------------ per-body systems */
/* ChainBendingForceSystem.zeroAccumulators, ChainBendingForceSystem.java:38-45 */
void sbo_zero_accumulators(int n, float *forceSum, float *torqueSum) {
    memset(forceSum, 0, sizeof(float) * 3 * (size_t)n);
    memset(torqueSum, 0, sizeof(float) * 3 * (size_t)n);
}

/* BrownianForceSystem.brownianForce, BrownianForceSystem.java:44-106.
 * bodyBase offsets the RNG body index (replicas share one counter space). */
void sbo_brownian_force(int n, float *randForce, float *randTorque, const float *bTransGam, const float *bRotGam,
                        const float *brownTransScale, const float *brownRotScale, float brownianForceMag,
                        int32_t step, int32_t seed, int rng, uint32_t stream, uint32_t bodyBase) {
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        int iy = n + i, iz = 2 * n + i;
        uint32_t h[6];
        sbo_brownian_bits(rng, bodyBase + (uint32_t)i, step, seed, stream, h);
        float u1 = fmaxf_java(1.0e-7f, sbo_uniform(h[0]));
        float u2 = sbo_uniform(h[1]);
        float u3 = fmaxf_java(1.0e-7f, sbo_uniform(h[2]));
        float u4 = sbo_uniform(h[3]);
        float u5 = fmaxf_java(1.0e-7f, sbo_uniform(h[4]));
        float u6 = sbo_uniform(h[5]);

        float r1 = (float)sqrt((double)(-2.0f * (float)log((double)u1)));
        float th1 = 2.0f * 3.14159265f * u2;
        float gfx = r1 * (float)cos((double)th1);
        float gtx = r1 * (float)sin((double)th1);

        float r2 = (float)sqrt((double)(-2.0f * (float)log((double)u3)));
        float th2 = 2.0f * 3.14159265f * u4;
        float gfy = r2 * (float)cos((double)th2);
        float gty = r2 * (float)sin((double)th2);

        float r3 = (float)sqrt((double)(-2.0f * (float)log((double)u5)));
        float th3 = 2.0f * 3.14159265f * u6;
        float gfz = r3 * (float)cos((double)th3);
        float gtz = r3 * (float)sin((double)th3);

        float tS = brownTransScale[i];
        float rS = brownRotScale[i];
        randForce[i] = tS * brownianForceMag * (float)sqrt((double)bTransGam[i]) * gfx;
        randForce[iy] = tS * brownianForceMag * (float)sqrt((double)bTransGam[iy]) * gfy;
        randForce[iz] = tS * brownianForceMag * (float)sqrt((double)bTransGam[iz]) * gfz;
        randTorque[i] = rS * brownianForceMag * (float)sqrt((double)bRotGam[i]) * gtx;
        randTorque[iy] = rS * brownianForceMag * (float)sqrt((double)bRotGam[iy]) * gty;
        randTorque[iz] = rS * brownianForceMag * (float)sqrt((double)bRotGam[iz]) * gtz;
    }
}

/* RigidRodLangevinIntegrationSystem.integrate, RigidRodLangevinIntegrationSystem.java:33-112 */
void sbo_integrate(int n, float *coord, float *uVec, float *yVec, const float *forceSum, const float *torqueSum,
                   const float *randForce, const float *randTorque, const float *bTransGam, const float *bRotGam,
                   float dt) {
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        int iy = n + i, iz = 2 * n + i;
        float ux = uVec[i], uyc = uVec[iy], uzc = uVec[iz];
        float yx = yVec[i], yyc = yVec[iy], yzc = yVec[iz];
        float zx = uyc * yzc - uzc * yyc;
        float zy = uzc * yx - ux * yzc;
        float zz = ux * yyc - uyc * yx;
        float zlen = 1.0f / (float)sqrt((double)(zx * zx + zy * zy + zz * zz));
        zx *= zlen; zy *= zlen; zz *= zlen;

        float fx = forceSum[i], fy = forceSum[iy], fz = forceSum[iz];
        float tx = torqueSum[i], ty = torqueSum[iy], tz = torqueSum[iz];
        float bfx = ux * fx + uyc * fy + uzc * fz + randForce[i];
        float bfy = yx * fx + yyc * fy + yzc * fz + randForce[iy];
        float bfz = zx * fx + zy * fy + zz * fz + randForce[iz];
        float btx = ux * tx + uyc * ty + uzc * tz + randTorque[i];
        float bty = yx * tx + yyc * ty + yzc * tz + randTorque[iy];
        float btz = zx * tx + zy * ty + zz * tz + randTorque[iz];

        float bvx = 1.0e6f * bfx / bTransGam[i];
        float bvy = 1.0e6f * bfy / bTransGam[iy];
        float bvz = 1.0e6f * bfz / bTransGam[iz];
        float bwx = btx / bRotGam[i];
        float bwy = bty / bRotGam[iy];
        float bwz = btz / bRotGam[iz];

        float vx = ux * bvx + yx * bvy + zx * bvz;
        float vy = uyc * bvx + yyc * bvy + zy * bvz;
        float vz = uzc * bvx + yzc * bvy + zz * bvz;
        coord[i] = coord[i] + dt * vx;
        coord[iy] = coord[iy] + dt * vy;
        coord[iz] = coord[iz] + dt * vz;

        float uTransInZ = -bwy * dt;
        float uTransInY = bwz * dt;
        float nuX = ux + yx * uTransInY + zx * uTransInZ;
        float nuY = uyc + yyc * uTransInY + zy * uTransInZ;
        float nuZ = uzc + yzc * uTransInY + zz * uTransInZ;
        float nuInv = 1.0f / (float)sqrt((double)(nuX * nuX + nuY * nuY + nuZ * nuZ));
        uVec[i] = nuX * nuInv;
        uVec[iy] = nuY * nuInv;
        uVec[iz] = nuZ * nuInv;

        float yTransInX = -uTransInY;
        float yTransInZ = bwx * dt;
        float nyX = ux * yTransInX + yx + zx * yTransInZ;
        float nyY = uyc * yTransInX + yyc + zy * yTransInZ;
        float nyZ = uzc * yTransInX + yzc + zz * yTransInZ;
        float nyInv = 1.0f / (float)sqrt((double)(nyX * nyX + nyY * nyY + nyZ * nyZ));
        yVec[i] = nyX * nyInv;
        yVec[iy] = nyY * nyInv;
        yVec[iz] = nyZ * nyInv;
    }
}

/* DerivedGeometrySystem.derive, DerivedGeometrySystem.java:27-73 */
void sbo_derive(int n, const float *coord, const float *uVec, float *yVec, float *zVec, float *end1, float *end2,
                const float *segLength) {
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        int iy = n + i, iz = 2 * n + i;
        float ux = uVec[i], uyc = uVec[iy], uzc = uVec[iz];
        float yx = yVec[i], yyc = yVec[iy], yzc = yVec[iz];
        float zx = uyc * yzc - uzc * yyc;
        float zy = uzc * yx - ux * yzc;
        float zz = ux * yyc - uyc * yx;
        float zmag2 = zx * zx + zy * zy + zz * zz;
        if (zmag2 > 0.0f) {
            float inv = 1.0f / (float)sqrt((double)zmag2);
            zx *= inv; zy *= inv; zz *= inv;
        }
        if (zVec) { zVec[i] = zx; zVec[iy] = zy; zVec[iz] = zz; }
        float nyx = zy * uzc - zz * uyc;
        float nyy = zz * ux - zx * uzc;
        float nyz = zx * uyc - zy * ux;
        yVec[i] = nyx; yVec[iy] = nyy; yVec[iz] = nyz;
        float cx = coord[i], cy = coord[iy], cz = coord[iz];
        float halfLen = segLength[i] * 0.5f;
        end1[i] = cx - halfLen * ux;
        end1[iy] = cy - halfLen * uyc;
        end1[iz] = cz - halfLen * uzc;
        end2[i] = cx + halfLen * ux;
        end2[iy] = cy + halfLen * uyc;
        end2[iz] = cz + halfLen * uzc;
    }
}

/* ContainmentSystem.wallPush, ContainmentSystem.java:130-136 */
static double wall_push(double d, double halfDim) {
    double s = (d > 0.0) ? 1.0 : ((d < 0.0) ? -1.0 : 0.0);
    double push = s * halfDim - d;
    double ps = (push > 0.0) ? 1.0 : ((push < 0.0) ? -1.0 : 0.0);
    if (ps == s) return 0.0;
    return push;
}

/* ContainmentSystem.confine, ContainmentSystem.java:58-126 */
void sbo_confine(int n, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                 float *forceSum, float *torqueSum, const float *boxParams, int32_t step) {
    int checkInt = (int)boxParams[6];
    if (checkInt > 1 && (step % checkInt) != 0) return;
    double tau = boxParams[0];
    double halfX = 0.5 * boxParams[1] - boxParams[4];
    double halfY = 0.5 * boxParams[2] - boxParams[4];
    double halfZ = 0.5 * boxParams[3] - boxParams[4];
    double coeff = boxParams[5];
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        int iy = n + i, iz = 2 * n + i;
        double cx = coord[i], cy = coord[iy], cz = coord[iz];
        double ux = uVec[i], uy = uVec[iy], uz = uVec[iz];
        double half = 0.5 * (double)segLength[i];
        double bTGx = bTransGam[i];
        double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
        int touched = 0;
        for (int e = 0; e < 2; e++) {
            double sgn = (e == 0) ? -1.0 : 1.0;
            double dx = cx + sgn * half * ux;
            double dy = cy + sgn * half * uy;
            double dz = cz + sgn * half * uz;
            double fux = wall_push(dx, halfX);
            double fuy = wall_push(dy, halfY);
            double fuz = wall_push(dz, halfZ);
            double delta2 = fux * fux + fuy * fuy + fuz * fuz;
            if (delta2 > 0.0) {
                double delta = sqrt(delta2);
                double inv = 1.0 / delta;
                double lux = fux * inv, luy = fuy * inv, luz = fuz * inv;
                double mag = coeff * 1.0e-6 * delta * bTGx / tau;
                double Fx = mag * lux, Fy = mag * luy, Fz = mag * luz;
                double rx = sgn * half * ux * 1.0e-6;
                double ry = sgn * half * uy * 1.0e-6;
                double rz = sgn * half * uz * 1.0e-6;
                fx += Fx; fy += Fy; fz += Fz;
                tx += ry * Fz - rz * Fy;
                ty += rz * Fx - rx * Fz;
                tz += rx * Fy - ry * Fx;
                touched = 1;
            }
        }
        if (touched) {
            forceSum[i] = (float)((double)forceSum[i] + fx);
            forceSum[iy] = (float)((double)forceSum[iy] + fy);
            forceSum[iz] = (float)((double)forceSum[iz] + fz);
            torqueSum[i] = (float)((double)torqueSum[i] + tx);
            torqueSum[iy] = (float)((double)torqueSum[iy] + ty);
            torqueSum[iz] = (float)((double)torqueSum[iz] + tz);
        }
    }
}

/* ------------------------------------------------------------------ chain */
/* ChainBendingForceSystem.accurateAcos, ChainBendingForceSystem.java:48-72 */
double sbo_accurate_acos(double x) {
    if (x > 1.0) x = 1.0;
    if (x < -1.0) x = -1.0;
    double y;
    if (x > 0.95) {
        double t = 1.0 - x;
        if (t < 0.0) t = 0.0;
        y = sqrt(2.0 * t);
    } else if (x < -0.95) {
        double t = 1.0 + x;
        if (t < 0.0) t = 0.0;
        y = 3.141592653589793 - sqrt(2.0 * t);
    } else {
        double ax = (x < 0.0) ? -x : x;
        double p = (-0.0187293 * ax + 0.0742610) * ax - 0.2121144;
        p = (p * ax + 1.5707963);
        p = p * sqrt(1.0 - ax);
        y = (x < 0.0) ? (3.141592653589793 - p) : p;
    }
    double s = sin(y);
    if (s > 1.0e-12 || s < -1.0e-12) y = y + (cos(y) - x) / s;
    s = sin(y);
    if (s > 1.0e-12 || s < -1.0e-12) y = y + (cos(y) - x) / s;
    return y;
}

/* ChainBendingForceSystem.accurateAsin, ChainBendingForceSystem.java:76-84 */
double sbo_accurate_asin(double s) {
    double s2 = s * s;
    double y = s * (1.0 + s2 * (0.16666666666666666 + s2 * (0.075 + s2 * 0.044642857142857)));
    double cy = cos(y);
    if (cy > 1.0e-12) y = y + (s - sin(y)) / cy;
    cy = cos(y);
    if (cy > 1.0e-12) y = y + (s - sin(y)) / cy;
    return y;
}

/* ChainBendingForceSystem.angleFromSinCos, ChainBendingForceSystem.java:95-103 */
double sbo_angle_from_sin_cos(double sin2, double c) {
    double s = sqrt(sin2);
    double ac = (c < 0.0) ? -c : c;
    if (s <= ac) {
        double base = sbo_accurate_asin(s);
        return (c >= 0.0) ? base : (PI_D - base);
    }
    return sbo_accurate_acos(c);
}

/* The chain mobility's sin^2(acos(c)).  The reference squares Math.sin(accurateAcos(c))
 * (ChainBendingForceSystem.java:196-197, :202-203, :281-282, :287-288); by default it is taken in the
 * equal form (1 - c)(1 + c), which the device uses too.  sbo_set_chain_reference_mobility(1) switches
 * the oracle to the reference's own evaluation, so a test can bound what the identity changes. */
static int g_chain_reference_mobility = 0;
void sbo_set_chain_reference_mobility(int on) { g_chain_reference_mobility = on; }
static double sin2_of_acos(double c) {
    if (g_chain_reference_mobility) {
        double s = sin(sbo_accurate_acos(c));
        return s * s;
    }
    return (1.0 - c) * (1.0 + c);
}

/* One joint of ChainBendingForceSystem.chainForces (end2 side :152-235 with endSign=+1,
 * end1 side :237-321 with endSign=-1).  The two reference blocks differ only in the sign of
 * the endpoint offset, the self link-point offset, the lever arm and the torsion self axis.
 * The reference squares Math.sin(accurateAcos(cosB)) (:196-197, :202-203, :281-282, :287-288);
 * that is sin^2(acos(cosB)) = (1 - cosB)(1 + cosB) for the clamped cosB, taken here in that form
 * (a few double ulps from the reference's evaluation; the mobility it enters is dominated by the
 * cosB^2 term). */
static void chain_joint(int n, int i, int nbr, int side, double endSign,
                        const float *coord, const float *uVec, const float *segLength,
                        const float *bTransGam, const float *bRotGam,
                        double cx, double cy, double cz, double ux, double uy, double uz, double len,
                        double halfLen_um, double lSqSelf, double rBTGx, double rBTGy, double rBRGy,
                        double dt, double fracMove, double fracR, double fracMoveTorq,
                        double filTorqSpringActive, double filTorqSpring, double actinMonoRadius,
                        double *f, double *t) {
    const double DEG2RAD = PI_D / 180.0;
    const double RAD2DEG = 180.0 / PI_D;
    int jx = nbr, jy = n + nbr, jz = 2 * n + nbr;
    double ncx = coord[jx], ncy = coord[jy], ncz = coord[jz];
    double nux = uVec[jx], nuy = uVec[jy], nuz = uVec[jz];
    double nlen = segLength[nbr];
    double nHalfLen_um = 0.5 * nlen;
    double lSqN = 1.0e-12 * nlen * nlen;
    double nBTGx = bTransGam[jx];
    double nBTGy = bTransGam[jy];
    double nBRGy = bRotGam[jy];

    double ex, ey, ez;
    if (endSign > 0) { ex = cx + halfLen_um * ux; ey = cy + halfLen_um * uy; ez = cz + halfLen_um * uz; }
    else { ex = cx - halfLen_um * ux; ey = cy - halfLen_um * uy; ez = cz - halfLen_um * uz; }
    double selfLpx, selfLpy, selfLpz;
    if (endSign > 0) { selfLpx = ex - actinMonoRadius * ux; selfLpy = ey - actinMonoRadius * uy; selfLpz = ez - actinMonoRadius * uz; }
    else { selfLpx = ex + actinMonoRadius * ux; selfLpy = ey + actinMonoRadius * uy; selfLpz = ez + actinMonoRadius * uz; }
    double nbrTipx, nbrTipy, nbrTipz;
    if (side == 0) { nbrTipx = ncx - nHalfLen_um * nux; nbrTipy = ncy - nHalfLen_um * nuy; nbrTipz = ncz - nHalfLen_um * nuz; }
    else { nbrTipx = ncx + nHalfLen_um * nux; nbrTipy = ncy + nHalfLen_um * nuy; nbrTipz = ncz + nHalfLen_um * nuz; }
    double nbrLpx, nbrLpy, nbrLpz;
    if (side == 0) { nbrLpx = nbrTipx + actinMonoRadius * nux; nbrLpy = nbrTipy + actinMonoRadius * nuy; nbrLpz = nbrTipz + actinMonoRadius * nuz; }
    else { nbrLpx = nbrTipx - actinMonoRadius * nux; nbrLpy = nbrTipy - actinMonoRadius * nuy; nbrLpz = nbrTipz - actinMonoRadius * nuz; }
    double linkPtX, linkPtY, linkPtZ, ptAtEndX, ptAtEndY, ptAtEndZ;
    if (i < nbr) {
        linkPtX = selfLpx; linkPtY = selfLpy; linkPtZ = selfLpz;
        ptAtEndX = nbrTipx; ptAtEndY = nbrTipy; ptAtEndZ = nbrTipz;
    } else {
        linkPtX = nbrLpx; linkPtY = nbrLpy; linkPtZ = nbrLpz;
        ptAtEndX = ex; ptAtEndY = ey; ptAtEndZ = ez;
    }
    double dx = ptAtEndX - linkPtX, dy = ptAtEndY - linkPtY, dz = ptAtEndZ - linkPtZ;
    double strainDist = sqrt(dx * dx + dy * dy + dz * dz);
    double invStrain = (strainDist > 0.0) ? (1.0 / strainDist) : 0.0;
    double luX = dx * invStrain, luY = dy * invStrain, luZ = dz * invStrain;

    double cosB1 = ux * luX + uy * luY + uz * luZ;
    if (cosB1 > 1.0) cosB1 = 1.0;
    if (cosB1 < -1.0) cosB1 = -1.0;
    double cosA1_2 = sin2_of_acos(cosB1);
    double moveC1 = cosB1 * cosB1 / rBTGx + cosA1_2 / rBTGy + lSqSelf * cosA1_2 / (4.0 * rBRGy);

    double cosB2 = nux * luX + nuy * luY + nuz * luZ;
    if (cosB2 > 1.0) cosB2 = 1.0;
    if (cosB2 < -1.0) cosB2 = -1.0;
    double cosA2_2 = sin2_of_acos(cosB2);
    double moveC2 = cosB2 * cosB2 / nBTGx + cosA2_2 / nBTGy + lSqN * cosA2_2 / (4.0 * nBRGy);

    double denom = dt * (moveC1 + moveC2);
    double forceMag = (denom > 0.0) ? (fracMove * 1.0e-6 * strainDist / denom) : 0.0;
    double fSign = (i < nbr) ? 1.0 : -1.0;
    double Fx = fSign * forceMag * luX, Fy = fSign * forceMag * luY, Fz = fSign * forceMag * luZ;
    f[0] += Fx; f[1] += Fy; f[2] += Fz;
    double Rscale = (endSign > 0) ? (0.5e-6 * len * fracR) : (-0.5e-6 * len * fracR);
    double Rx = Rscale * ux, Ry = Rscale * uy, Rz = Rscale * uz;
    t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;

    double nuxE, nuyE, nuzE;
    if (side == 0) { nuxE = nux; nuyE = nuy; nuzE = nuz; } else { nuxE = -nux; nuyE = -nuy; nuzE = -nuz; }
    double mux, muy, muz;
    if (endSign > 0) { mux = ux; muy = uy; muz = uz; } else { mux = -ux; muy = -uy; muz = -uz; }
    double tvx = muy * nuzE - muz * nuyE, tvy = muz * nuxE - mux * nuzE, tvz = mux * nuyE - muy * nuxE;
    double tvMag2 = tvx * tvx + tvy * tvy + tvz * tvz;
    if (tvMag2 > 1.0e-30) {
        double invMag = 1.0 / sqrt(tvMag2);
        tvx *= invMag; tvy *= invMag; tvz *= invMag;
        double dotV = mux * nuxE + muy * nuyE + muz * nuzE;
        if (dotV > 1.0) dotV = 1.0;
        if (dotV < -1.0) dotV = -1.0;
        double angTween = sbo_angle_from_sin_cos(tvMag2, dotV) * RAD2DEG;
        double torsionMag;
        if (filTorqSpringActive > 0.5) {
            torsionMag = fracMoveTorq * filTorqSpring * angTween;
        } else {
            double invBRG = 1.0 / rBRGy + 1.0 / nBRGy;
            torsionMag = fracMoveTorq * DEG2RAD * angTween / (invBRG * dt);
        }
        t[0] += tvx * torsionMag; t[1] += tvy * torsionMag; t[2] += tvz * torsionMag;
    }
}

/* ChainBendingForceSystem.chainForces, ChainBendingForceSystem.java:105-331 */
void sbo_chain_forces(int n, const float *coord, const float *uVec, const float *segLength,
                      const int *end2NbrSlot, const int *end2NbrSide, const int *end1NbrSlot, const int *end1NbrSide,
                      const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                      const float *chainParams) {
    double dt = chainParams[0], fracMove = chainParams[1], fracR = chainParams[2];
    double fracMoveTorq = chainParams[3], filTorqSpringActive = chainParams[4];
    double filTorqSpring = chainParams[5], actinMonoRadius = chainParams[6];
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        int iy = n + i, iz = 2 * n + i;
        double cx = coord[i], cy = coord[iy], cz = coord[iz];
        double ux = uVec[i], uy = uVec[iy], uz = uVec[iz];
        double len = segLength[i];
        double halfLen_um = 0.5 * len;
        double lSqSelf = 1.0e-12 * len * len;
        double rBTGx = bTransGam[i], rBTGy = bTransGam[iy], rBRGy = bRotGam[iy];
        double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
        int e2Slot = end2NbrSlot[i];
        int e1Slot = end1NbrSlot[i];
        if (e2Slot >= 0)
            chain_joint(n, i, e2Slot, end2NbrSide[i], +1.0, coord, uVec, segLength, bTransGam, bRotGam, cx, cy, cz,
                        ux, uy, uz, len, halfLen_um, lSqSelf, rBTGx, rBTGy, rBRGy, dt, fracMove, fracR, fracMoveTorq,
                        filTorqSpringActive, filTorqSpring, actinMonoRadius, f, t);
        if (e1Slot >= 0)
            chain_joint(n, i, e1Slot, end1NbrSide[i], -1.0, coord, uVec, segLength, bTransGam, bRotGam, cx, cy, cz,
                        ux, uy, uz, len, halfLen_um, lSqSelf, rBTGx, rBTGy, rBRGy, dt, fracMove, fracR, fracMoveTorq,
                        filTorqSpringActive, filTorqSpring, actinMonoRadius, f, t);
        forceSum[i] = (float)(forceSum[i] + f[0]);
        forceSum[iy] = (float)(forceSum[iy] + f[1]);
        forceSum[iz] = (float)(forceSum[iz] + f[2]);
        torqueSum[i] = (float)(torqueSum[i] + t[0]);
        torqueSum[iy] = (float)(torqueSum[iy] + t[1]);
        torqueSum[iz] = (float)(torqueSum[iz] + t[2]);
    }
}

/* ------------------------------------------------------------------ spatial grid */
/* FilamentStore.publishToBodyView, FilamentStore.java:275-295 */
void sbo_publish_filaments(int n, int cap, const float *coord, const float *segLength, float actinRadius,
                           float *center, float *boundingRadius) {
    SBO_PAR_FOR
    for (int i = 0; i < n; i++) {
        center[i] = coord[i];
        center[cap + i] = coord[n + i];
        center[2 * cap + i] = coord[2 * n + i];
        boundingRadius[i] = 0.5f * segLength[i] + actinRadius;
    }
}

/* Java (int) of a float followed by the reference clamp to [0, nBins-1]: any value below 0
 * (including NaN, which Java converts to 0) lands on 0. */
static int clamp_cell(float v, int nBins) {
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)nBins) return nBins - 1;
    int c = (int)v;
    return c >= nBins ? nBins - 1 : c;
}

/* periodic extension: floor, then wrap into [0, nBins) */
static int wrap_cell(float v, int nBins) {
    if (!(v == v)) return 0;
    float fl = floorf(v);
    double m = fmod((double)fl, (double)nBins);
    if (m < 0) m += nBins;
    int c = (int)m;
    return c >= nBins ? 0 : c;
}

/* SpatialGrid.bodyCell, SpatialGrid.java:51-75 (periodic: build extension) */
void sbo_body_cell(int cap, int S, const float *center, const float *gridParams, const int *gridDims, int periodic,
                   int *bodyCell) {
    float xMin = gridParams[0], yMin = gridParams[1], zMin = gridParams[2];
    float invCell = gridParams[4];
    int nX = gridDims[0], nY = gridDims[1], nZ = gridDims[2];
    int nXY = nX * nY;
    SBO_PAR_FOR
    for (int i = 0; i < cap; i++) {
        if (i >= S) { bodyCell[i] = -1; continue; }
        float fx = (center[i] - xMin) * invCell;
        float fy = (center[cap + i] - yMin) * invCell;
        float fz = (center[2 * cap + i] - zMin) * invCell;
        int ix, iy, iz;
        if (periodic) { ix = wrap_cell(fx, nX); iy = wrap_cell(fy, nY); iz = wrap_cell(fz, nZ); }
        else { ix = clamp_cell(fx, nX); iy = clamp_cell(fy, nY); iz = clamp_cell(fz, nZ); }
        bodyCell[i] = ix + iy * nX + iz * nXY;
    }
}

/* Serial stable counting sort: SpatialGrid.gridZero/gridHistogram/gridScanLocal/gridScanChunks/gridScanAdd/gridScatter,
 * SpatialGrid.java:78-182 (the two-level scan yields the same exclusive prefix as a serial one). */
void sbo_grid_build_serial(int S, int contentsCap, const int *bodyCell, const int *gridDims,
                           int *cellCount, int *gridCellOffsets, int *gridCellContents) {
    int totalCells = gridDims[3];
    for (int c = 0; c < totalCells; c++) cellCount[c] = 0;
    for (int s = 0; s < S; s++) {
        int c = bodyCell[s];
        if (c >= 0) cellCount[c]++;
    }
    int acc = 0;
    for (int c = 0; c < totalCells; c++) {
        gridCellOffsets[c] = acc;
        acc += cellCount[c];
        cellCount[c] = 0;
    }
    gridCellOffsets[totalCells] = acc < contentsCap ? acc : contentsCap;
    for (int s = 0; s < S; s++) {
        int c = bodyCell[s];
        if (c < 0) continue;
        int writePos = gridCellOffsets[c] + cellCount[c];
        if (writePos < contentsCap) gridCellContents[writePos] = s;
        cellCount[c]++;
    }
}

static float min_image(float d, float L) {
    float inv = 1.0f / L;
    return d - L * rintf(d * inv);
}

/* SpatialGrid.broadPhase, SpatialGrid.java:340-406.  Box mode is the reference verbatim.
 * Periodic mode (build extension): the 27 stencil cells wrap and the distance is the minimum
 * image; requires nBins >= 3 per axis so no cell is visited twice. */
void sbo_broad_phase(int cap, int S, const float *center, const float *boundingRadius, const int *bodyCell,
                     const int *gridCellOffsets, const int *gridCellContents, const int *gridDims,
                     const float *gridParams, int periodic, const float *boxLen, int maxCand,
                     int *candPartner, int *candCount) {
    int nX = gridDims[0], nY = gridDims[1], nZ = gridDims[2];
    int nXY = nX * nY;
    float cutoff = gridParams[5];
    SBO_PAR_FOR_DYNAMIC
    for (int i = 0; i < cap; i++) {
        candCount[i] = 0;
        if (i >= S) continue;
        int c = bodyCell[i];
        if (c < 0) continue;
        int cz = c / nXY;
        int rem = c - cz * nXY;
        int cy = rem / nX;
        int cx = rem - cy * nX;
        float icx = center[i], icy = center[cap + i], icz = center[2 * cap + i];
        float ri = boundingRadius[i];
        int x0 = cx - 1, x1 = cx + 1, y0 = cy - 1, y1 = cy + 1, z0 = cz - 1, z1 = cz + 1;
        if (!periodic) {
            if (x0 < 0) x0 = 0;
            if (x1 >= nX) x1 = nX - 1;
            if (y0 < 0) y0 = 0;
            if (y1 >= nY) y1 = nY - 1;
            if (z0 < 0) z0 = 0;
            if (z1 >= nZ) z1 = nZ - 1;
        }
        int cnt = 0;
        for (int zz = z0; zz <= z1; zz++) {
            int wz = periodic ? (zz + nZ) % nZ : zz;
            for (int yy = y0; yy <= y1; yy++) {
                int wy = periodic ? (yy + nY) % nY : yy;
                for (int xx = x0; xx <= x1; xx++) {
                    int wx = periodic ? (xx + nX) % nX : xx;
                    int cc = wx + wy * nX + wz * nXY;
                    int start = gridCellOffsets[cc];
                    int end = gridCellOffsets[cc + 1];
                    for (int idx = start; idx < end; idx++) {
                        int j = gridCellContents[idx];
                        if (j <= i) continue;
                        float dx = icx - center[j];
                        float dy = icy - center[cap + j];
                        float dz = icz - center[2 * cap + j];
                        if (periodic) {
                            dx = min_image(dx, boxLen[0]);
                            dy = min_image(dy, boxLen[1]);
                            dz = min_image(dz, boxLen[2]);
                        }
                        float distSq = dx * dx + dy * dy + dz * dz;
                        float reach = ri + boundingRadius[j] + cutoff;
                        if (distSq <= reach * reach) {
                            if (cnt < maxCand) candPartner[(size_t)i * maxCand + cnt] = j;
                            cnt++;
                        }
                    }
                }
            }
        }
        candCount[i] = cnt;
    }
}

/* SpatialGrid.bruteForce, SpatialGrid.java:413-445 (periodic: minimum image) */
void sbo_brute_force(int cap, int S, const float *center, const float *boundingRadius, const float *gridParams,
                     int periodic, const float *boxLen, int maxCand, int *brutePartner, int *bruteCount) {
    float cutoff = gridParams[5];
    for (int i = 0; i < cap; i++) {
        bruteCount[i] = 0;
        if (i >= S) continue;
        float icx = center[i], icy = center[cap + i], icz = center[2 * cap + i];
        float ri = boundingRadius[i];
        int cnt = 0;
        for (int j = i + 1; j < S; j++) {
            float dx = icx - center[j];
            float dy = icy - center[cap + j];
            float dz = icz - center[2 * cap + j];
            if (periodic) {
                dx = min_image(dx, boxLen[0]);
                dy = min_image(dy, boxLen[1]);
                dz = min_image(dz, boxLen[2]);
            }
            float distSq = dx * dx + dy * dy + dz * dz;
            float reach = ri + boundingRadius[j] + cutoff;
            if (distSq <= reach * reach) {
                if (cnt < maxCand) brutePartner[(size_t)i * maxCand + cnt] = j;
                cnt++;
            }
        }
        bruteCount[i] = cnt;
    }
}

/* ------------------------------------------------------------------ CSR inverse */
/* CrossBridgeSystem.csrHistogram/csrScan/csrScatter, CrossBridgeSystem.java:190-217:
 * stable counting sort of items by key (key < 0 skipped); count is left as the per-key totals. */
void sbo_csr_build(int nKeys, int nItems, const int *key, int *count, int *offsets, int *list) {
    for (int s = 0; s < nKeys; s++) count[s] = 0;
    for (int m = 0; m < nItems; m++) {
        int s = key[m];
        if (s >= 0) count[s]++;
    }
    int acc = 0;
    for (int s = 0; s < nKeys; s++) {
        offsets[s] = acc;
        acc += count[s];
        count[s] = 0;
    }
    offsets[nKeys] = acc;
    for (int m = 0; m < nItems; m++) {
        int s = key[m];
        if (s < 0) continue;
        list[offsets[s] + count[s]] = m;
        count[s]++;
    }
}

/* ------------------------------------------------------------------ crosslinks */
/* CrosslinkerSystem.countActiveLinks, CrosslinkerSystem.java:694-706 */
void sbo_count_active_links(int C, int nSeg, const int *linkState, const int *linkFilA, const int *linkFilB,
                            int *activeLinkCount) {
    for (int s = 0; s < nSeg; s++) activeLinkCount[s] = 0;
    for (int L = 0; L < C; L++) {
        if (linkState[L] < 0) continue;
        activeLinkCount[linkFilA[L]]++;
        activeLinkCount[linkFilB[L]]++;
    }
}

#define XL_STRIDE 12
#define STRAIN_WIN 10
#define TORQUE_WIN 5

/* CrosslinkerSystem.unbind, CrosslinkerSystem.java:231-281 */
void sbo_link_unbind(int nSeg, int nLinks, const float *filUVec, const float *filEnd1,
                     const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2,
                     int *linkState, float *strainHist, int *strainPlace, const float *offParams,
                     int32_t step, int32_t seed, int rng, uint32_t stream) {
    double offConst = offParams[0], offCoeff = offParams[1], offExp = offParams[2];
    double dt = offParams[3], restLength = offParams[4];
    const int W = STRAIN_WIN;
    SBO_PAR_FOR
    for (int k = 0; k < nLinks; k++) {
        if (linkState[k] < 0) continue;
        int a = linkFilA[k], b = linkFilB[k];
        double l1 = loc1[k];
        double uax = filUVec[a], uay = filUVec[nSeg + a], uaz = filUVec[2 * nSeg + a];
        double p1x = filEnd1[a] + l1 * uax;
        double p1y = filEnd1[nSeg + a] + l1 * uay;
        double p1z = filEnd1[2 * nSeg + a] + l1 * uaz;
        double l2 = loc2[k];
        double ubx = filUVec[b], uby = filUVec[nSeg + b], ubz = filUVec[2 * nSeg + b];
        double p2x = filEnd1[b] + l2 * ubx;
        double p2y = filEnd1[nSeg + b] + l2 * uby;
        double p2z = filEnd1[2 * nSeg + b] + l2 * ubz;
        double lvx = p2x - p1x, lvy = p2y - p1y, lvz = p2z - p1z;
        double linkLength = sqrt(lvx * lvx + lvy * lvy + lvz * lvz);
        double stretch = linkLength - restLength;
        if (stretch < 0.0) stretch = 0.0;
        double strain = stretch / restLength;
        int base = k * W;
        int p = strainPlace[k];
        strainHist[base + p] = (float)strain;
        strainPlace[k] = (p + 1) % W;
        double ave = 0.0;
        for (int j = 0; j < W; j++) ave += strainHist[base + j];
        ave /= W;
        double kOff = offConst + offCoeff * exp(ave * offExp);
        double pBreak = kOff * dt;
        uint32_t h = sbo_event_bits(rng, (uint32_t)k, step, seed, stream, 0x584C4B42u);
        float u = sbo_uniform(h);
        if ((double)u < pBreak) linkState[k] = -1;
    }
}

/* CrosslinkerSystem.linkForces, CrosslinkerSystem.java:51-124 */
void sbo_link_forces(int nSeg, int nLinks, const float *filCoord, const float *filUVec, const float *filEnd1,
                     const float *filBTransGam, const int *linkFilA, const int *linkFilB, const float *loc1,
                     const float *loc2, const int *filLinkCt, float *xlinkData, const float *xlParams) {
    double restLength = xlParams[0];
    double fracMoveBase = xlParams[1];
    double dt = xlParams[2];
    SBO_PAR_FOR
    for (int k = 0; k < nLinks; k++) {
        int d = k * XL_STRIDE;
        for (int j = 0; j < XL_STRIDE; j++) xlinkData[d + j] = 0.0f;
        int a = linkFilA[k], b = linkFilB[k];
        if (a < 0 || b < 0) continue;
        double l1 = loc1[k];
        double uax = filUVec[a], uay = filUVec[nSeg + a], uaz = filUVec[2 * nSeg + a];
        double p1x = filEnd1[a] + l1 * uax;
        double p1y = filEnd1[nSeg + a] + l1 * uay;
        double p1z = filEnd1[2 * nSeg + a] + l1 * uaz;
        double l2 = loc2[k];
        double ubx = filUVec[b], uby = filUVec[nSeg + b], ubz = filUVec[2 * nSeg + b];
        double p2x = filEnd1[b] + l2 * ubx;
        double p2y = filEnd1[nSeg + b] + l2 * uby;
        double p2z = filEnd1[2 * nSeg + b] + l2 * ubz;
        double lvx = p2x - p1x, lvy = p2y - p1y, lvz = p2z - p1z;
        double linkLength = sqrt(lvx * lvx + lvy * lvy + lvz * lvz);
        double stretch = linkLength - restLength;
        if (stretch < 0.0) stretch = 0.0;
        int ctA = filLinkCt[a], ctB = filLinkCt[b];
        int maxLinks = ctA > ctB ? ctA : ctB;
        if (maxLinks < 1) maxLinks = 1;
        double fracMove = fracMoveBase / maxLinks;
        double g1 = filBTransGam[a];
        double g2 = filBTransGam[b];
        double curForceMag = (fracMove * 1.0e-6 * stretch / dt) / (1.0 / g1 + 1.0 / g2);
        double Fx = curForceMag * lvx, Fy = curForceMag * lvy, Fz = curForceMag * lvz;
        double rax = (p1x - filCoord[a]) * 1e-6;
        double ray = (p1y - filCoord[nSeg + a]) * 1e-6;
        double raz = (p1z - filCoord[2 * nSeg + a]) * 1e-6;
        double TAx = ray * Fz - raz * Fy;
        double TAy = raz * Fx - rax * Fz;
        double TAz = rax * Fy - ray * Fx;
        double nFx = -Fx, nFy = -Fy, nFz = -Fz;
        double rbx = (p2x - filCoord[b]) * 1e-6;
        double rby = (p2y - filCoord[nSeg + b]) * 1e-6;
        double rbz = (p2z - filCoord[2 * nSeg + b]) * 1e-6;
        double TBx = rby * nFz - rbz * nFy;
        double TBy = rbz * nFx - rbx * nFz;
        double TBz = rbx * nFy - rby * nFx;
        xlinkData[d] = (float)Fx; xlinkData[d + 1] = (float)Fy; xlinkData[d + 2] = (float)Fz;
        xlinkData[d + 3] = (float)TAx; xlinkData[d + 4] = (float)TAy; xlinkData[d + 5] = (float)TAz;
        xlinkData[d + 6] = (float)nFx; xlinkData[d + 7] = (float)nFy; xlinkData[d + 8] = (float)nFz;
        xlinkData[d + 9] = (float)TBx; xlinkData[d + 10] = (float)TBy; xlinkData[d + 11] = (float)TBz;
    }
}

/* CrosslinkerSystem.fastAcos, CrosslinkerSystem.java:372-376 (its accurateAcos :382-400 is the
 * same polynomial + Newton refinement as ChainBendingForceSystem.accurateAcos) */
static double fast_acos(double dot) {
    if (dot > 0.95) { double t = 1.0 - dot; if (t < 0) t = 0; return sqrt(2.0 * t); }
    if (dot < -0.95) { double t = 1.0 + dot; if (t < 0) t = 0; return PI_D - sqrt(2.0 * t); }
    return sbo_accurate_acos(dot);
}

/* CrosslinkerSystem.linkTorsion, CrosslinkerSystem.java:767-808 */
void sbo_link_torsion(int nSeg, int nLinks, const float *filUVec, const int *linkFilA, const int *linkFilB,
                      const int *linkState, const int *linkOrientSame, float *torqueMagHist, int *torqueMagPlace,
                      float *xlinkData, const float *torsionParams) {
    if (torsionParams[1] < 0.5f) return;
    double spring = torsionParams[0];
    const int W = TORQUE_WIN;
    SBO_PAR_FOR
    for (int k = 0; k < nLinks; k++) {
        if (linkState[k] < 0) continue;
        int a = linkFilA[k], b = linkFilB[k];
        if (a < 0 || b < 0) continue;
        double uax = filUVec[a], uay = filUVec[nSeg + a], uaz = filUVec[2 * nSeg + a];
        double ubx = filUVec[b], uby = filUVec[nSeg + b], ubz = filUVec[2 * nSeg + b];
        int same = linkOrientSame[k] != 0;
        double vbx = same ? ubx : -ubx, vby = same ? uby : -uby, vbz = same ? ubz : -ubz;
        double tx = uay * vbz - uaz * vby, ty = uaz * vbx - uax * vbz, tz = uax * vby - uay * vbx;
        double m2 = tx * tx + ty * ty + tz * tz;
        if (!(m2 > 1.0e-30)) continue;
        double im = 1.0 / sqrt(m2);
        tx *= im; ty *= im; tz *= im;
        double dot = uax * ubx + uay * uby + uaz * ubz;
        if (dot > 1.0) dot = 1.0;
        if (dot < -1.0) dot = -1.0;
        double angT = same ? fast_acos(dot) : fabs(fast_acos(dot) - PI_D);
        double curMag = spring * angT;
        int base = k * W;
        int p = torqueMagPlace[k];
        torqueMagHist[base + p] = (float)curMag;
        torqueMagPlace[k] = (p + 1) % W;
        double avg = 0.0;
        for (int j = 0; j < W; j++) avg += torqueMagHist[base + j];
        avg /= W;
        double Tx = avg * tx, Ty = avg * ty, Tz = avg * tz;
        int d = k * XL_STRIDE;
        xlinkData[d + 3] = (float)(xlinkData[d + 3] + Tx);
        xlinkData[d + 4] = (float)(xlinkData[d + 4] + Ty);
        xlinkData[d + 5] = (float)(xlinkData[d + 5] + Tz);
        xlinkData[d + 9] = (float)(xlinkData[d + 9] - Tx);
        xlinkData[d + 10] = (float)(xlinkData[d + 10] - Ty);
        xlinkData[d + 11] = (float)(xlinkData[d + 11] - Tz);
    }
}

static void gather_side(int nSeg, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                        const int *linkState, float *filForceSum, float *filTorqueSum, int off) {
    SBO_PAR_FOR
    for (int s = 0; s < nSeg; s++) {
        double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        int start = segOffsets[s], end = segOffsets[s + 1];
        for (int k = start; k < end; k++) {
            int li = segLinkIdx[k];
            if (linkState[li] < 0) continue;
            int d = li * XL_STRIDE + off;
            fx += xlinkData[d]; fy += xlinkData[d + 1]; fz += xlinkData[d + 2];
            tx += xlinkData[d + 3]; ty += xlinkData[d + 4]; tz += xlinkData[d + 5];
        }
        filForceSum[s] = (float)(filForceSum[s] + fx);
        filForceSum[nSeg + s] = (float)(filForceSum[nSeg + s] + fy);
        filForceSum[2 * nSeg + s] = (float)(filForceSum[2 * nSeg + s] + fz);
        filTorqueSum[s] = (float)(filTorqueSum[s] + tx);
        filTorqueSum[nSeg + s] = (float)(filTorqueSum[nSeg + s] + ty);
        filTorqueSum[2 * nSeg + s] = (float)(filTorqueSum[2 * nSeg + s] + tz);
    }
}

/* CrosslinkerSystem.segGatherA, CrosslinkerSystem.java:129-149 */
void sbo_seg_gather_a(int nSeg, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                      const int *linkState, float *filForceSum, float *filTorqueSum) {
    gather_side(nSeg, segOffsets, segLinkIdx, xlinkData, linkState, filForceSum, filTorqueSum, 0);
}

/* CrosslinkerSystem.segGatherB, CrosslinkerSystem.java:153-173 */
void sbo_seg_gather_b(int nSeg, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                      const int *linkState, float *filForceSum, float *filTorqueSum) {
    gather_side(nSeg, segOffsets, segLinkIdx, xlinkData, linkState, filForceSum, filTorqueSum, 6);
}

/* CrosslinkerSystem.bruteGather, CrosslinkerSystem.java:178-198 */
void sbo_link_brute_gather(int nSeg, int nLinks, const int *linkFilA, const int *linkFilB, const float *xlinkData,
                           const int *linkState, float *bForceSum, float *bTorqueSum) {
    for (int s = 0; s < nSeg; s++) {
        double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
        for (int k = 0; k < nLinks; k++) {
            if (linkState[k] < 0) continue;
            int d = k * XL_STRIDE;
            if (linkFilA[k] == s) {
                fx += xlinkData[d]; fy += xlinkData[d + 1]; fz += xlinkData[d + 2];
                tx += xlinkData[d + 3]; ty += xlinkData[d + 4]; tz += xlinkData[d + 5];
            }
            if (linkFilB[k] == s) {
                fx += xlinkData[d + 6]; fy += xlinkData[d + 7]; fz += xlinkData[d + 8];
                tx += xlinkData[d + 9]; ty += xlinkData[d + 10]; tz += xlinkData[d + 11];
            }
        }
        bForceSum[s] = (float)fx; bForceSum[nSeg + s] = (float)fy; bForceSum[2 * nSeg + s] = (float)fz;
        bTorqueSum[s] = (float)tx; bTorqueSum[nSeg + s] = (float)ty; bTorqueSum[2 * nSeg + s] = (float)tz;
    }
}

/* ------------------------------------------------------------------ whole step */
void sbo_network_grid(sbo_net *net) {
    int n = net->nRod;
    sbo_publish_filaments(n, n, net->coord, net->segLength, net->actinRadius, net->center, net->boundingRadius);
    sbo_body_cell(n, n, net->center, net->gridParams, net->gridDims, net->periodic, net->bodyCell);
    sbo_grid_build_serial(n, n, net->bodyCell, net->gridDims, net->cellCount, net->gridCellOffsets,
                          net->gridCellContents);
    sbo_broad_phase(n, n, net->center, net->boundingRadius, net->bodyCell, net->gridCellOffsets,
                    net->gridCellContents, net->gridDims, net->gridParams, net->periodic, net->boxLen, net->maxCand,
                    net->candPartner, net->candCount);
}

/* zero + Brownian + chain + crosslink unbind/count/force/torsion/gathers (accumulates forceSum) */
void sbo_network_forces(sbo_net *net, int32_t step) {
    int n = net->nRod;
    sbo_zero_accumulators(n, net->forceSum, net->torqueSum);
    sbo_brownian_force(n, net->randForce, net->randTorque, net->bTransGam, net->bRotGam, net->brownTransScale,
                       net->brownRotScale, net->params[1], step, net->seed, net->rng, net->stream, 0u);
    if (net->chainOn)
        sbo_chain_forces(n, net->coord, net->uVec, net->segLength, net->end2NbrSlot, net->end2NbrSide,
                         net->end1NbrSlot, net->end1NbrSide, net->bTransGam, net->bRotGam, net->forceSum,
                         net->torqueSum, net->chainParams);
    if (net->nLinks > 0) {
        int L = net->nLinks;
        if (net->unbindOn)
            sbo_link_unbind(n, L, net->uVec, net->end1, net->linkFilA, net->linkFilB, net->loc1, net->loc2,
                            net->linkState, net->strainHist, net->strainPlace, net->offParams, step, net->seed,
                            net->rng, net->stream);
        sbo_count_active_links(L, n, net->linkState, net->linkFilA, net->linkFilB, net->activeLinkCount);
        sbo_link_forces(n, L, net->coord, net->uVec, net->end1, net->bTransGam, net->linkFilA, net->linkFilB,
                        net->loc1, net->loc2, net->activeLinkCount, net->xlinkData, net->xlParams);
        sbo_link_torsion(n, L, net->uVec, net->linkFilA, net->linkFilB, net->linkState, net->linkOrientSame,
                         net->torqueMagHist, net->torqueMagPlace, net->xlinkData, net->torsionParams);
        sbo_csr_build(n, L, net->linkFilA, net->segCountA, net->segOffsetsA, net->segIdxA);
        sbo_seg_gather_a(n, net->segOffsetsA, net->segIdxA, net->xlinkData, net->linkState, net->forceSum,
                         net->torqueSum);
        sbo_csr_build(n, L, net->linkFilB, net->segCountB, net->segOffsetsB, net->segIdxB);
        sbo_seg_gather_b(n, net->segOffsetsB, net->segIdxB, net->xlinkData, net->linkState, net->forceSum,
                         net->torqueSum);
    }
}

/* confine + integrate + derive */
void sbo_network_finish(sbo_net *net, int32_t step) {
    int n = net->nRod;
    if (net->confineOn)
        sbo_confine(n, net->coord, net->uVec, net->segLength, net->bTransGam, net->forceSum, net->torqueSum,
                    net->boxParams, step);
    sbo_integrate(n, net->coord, net->uVec, net->yVec, net->forceSum, net->torqueSum, net->randForce,
                  net->randTorque, net->bTransGam, net->bRotGam, net->params[0]);
    sbo_derive(n, net->coord, net->uVec, net->yVec, net->zVec, net->end1, net->end2, net->segLength);
}

/* Filament block of DenseContractileHarness.cpuStep (DenseContractileHarness.java:350-429) with no
 * motors: [publish + grid + broad phase] -> zero + Brownian -> chain -> [unbind -> countActiveLinks ->
 * linkForces -> linkTorsion -> CSR A -> gatherA -> CSR B -> gatherB] -> confine -> integrate -> derive. */
void sbo_network_step(sbo_net *net, int32_t step) {
    if (net->turn && net->turn->interval > 0 && step % net->turn->interval == 0)   /* GrowthStore.firesAt */
        sbo_turnover_cadence(net->turn, net, step);
    if (net->gridOn) sbo_network_grid(net);
    if (net->form && net->formInt > 0 && step % net->formInt == 0)   /* V2OneXHarness.java:379 */
        net->form->lastFormed = sbo_formation(net->form, step, net->seed, net->rng, net->stream);
    sbo_network_forces(net, step);
    sbo_network_finish(net, step);
}
