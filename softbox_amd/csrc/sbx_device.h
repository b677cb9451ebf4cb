// sbx_device.h -- device-side arithmetic shared by every SoftBox mechanics kernel (gfx950).
//
// Every helper restates one piece of a reference System method with the same float/double
// types and the same operation order; the library is compiled with -ffp-contract=off and
// IEEE division/sqrt so float results match the reference's strict Java arithmetic.
// Planar SoA throughout: component a of body i lives at [a*n + i] (RigidRodBody.java:26-57).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SBX_PI_D 3.141592653589793

namespace sbx {

enum { RNG_WANG = 0, RNG_PHILOX = 1 };

// Thomas Wang's 32-bit integer mixer as used by BrownianForceSystem.wangHash
// (BrownianForceSystem.java:35-42); Java int wrap-around == uint32 arithmetic.
__device__ __forceinline__ uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

// Philox4x32-10 (Salmon et al., SC'11).  mul_hi/mul_lo map to v_mul_hi_u32 / v_mul_lo_u32.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Reference key (BrownianForceSystem.java:64): (i*1000003) ^ (step*999983) ^ (seed*7919)
__device__ __forceinline__ uint32_t wang_base(uint32_t i, int32_t step, int32_t seed) {
    return (i * 1000003u) ^ ((uint32_t)step * 999983u) ^ ((uint32_t)seed * 7919u);
}

// The six raw draws of one body-step.  WANG = BrownianForceSystem.java:65-70 salts;
// PHILOX = counter (body, step, block, 0), key (seed, stream).
__device__ __forceinline__ void brownian_bits(int rng, uint32_t body, int32_t step, int32_t seed, uint32_t stream,
                                              uint32_t h[6]) {
    if (rng == RNG_WANG) {
        uint32_t base = wang_base(body, step, seed);
        h[0] = wang_hash(base);
        h[1] = wang_hash(base ^ 0x9e3779b9u);
        h[2] = wang_hash(base ^ 0x85ebca6bu);
        h[3] = wang_hash(base ^ 0xc2b2ae35u);
        h[4] = wang_hash(base ^ 0x517cc1b7u);
        h[5] = wang_hash(base ^ 0x1f0a7ed5u);
    } else {
        uint32_t o0[4], o1[4];
        philox4x32_10(body, (uint32_t)step, 0u, 0u, (uint32_t)seed, stream, o0);
        philox4x32_10(body, (uint32_t)step, 1u, 0u, (uint32_t)seed, stream, o1);
        h[0] = o0[0]; h[1] = o0[1]; h[2] = o0[2]; h[3] = o0[3];
        h[4] = o1[0]; h[5] = o1[1];
    }
}

// Event draw keyed by (item, step, seed, salt): WANG = CrosslinkerSystem.java:277 form.
__device__ __forceinline__ uint32_t event_bits(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream,
                                               uint32_t salt) {
    if (rng == RNG_WANG) return wang_hash(wang_base(item, step, seed) ^ salt);
    uint32_t o[4];
    philox4x32_10(item, (uint32_t)step, salt, 2u, (uint32_t)seed, stream, o);
    return o[0];
}

// (h >>> 1) / 2147483647f  (BrownianForceSystem.java:72-77)
__device__ __forceinline__ float uniform01(uint32_t h) { return (float)(int32_t)(h >> 1) / 2147483647.0f; }

__device__ __forceinline__ float fmax_java(float a, float b) { return a >= b ? a : b; }

// Box-Muller noise of one body-step: gf = cos branch (force), gt = sin branch (torque),
// BrownianForceSystem.java:72-93.  log/cos/sin are evaluated in double and cast, as Java's
// (float) Math.log((double) u) does.
__device__ __forceinline__ void brownian_noise(int rng, uint32_t body, int32_t step, int32_t seed, uint32_t stream,
                                               float gf[3], float gt[3]) {
    uint32_t h[6];
    brownian_bits(rng, body, step, seed, stream, h);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ua = fmax_java(1.0e-7f, uniform01(h[2 * a]));
        float ub = uniform01(h[2 * a + 1]);
        float r = sqrtf(-2.0f * (float)log((double)ua));
        float th = 2.0f * 3.14159265f * ub;
        double s, c;
        sincos((double)th, &s, &c);
        gf[a] = r * (float)c;
        gt[a] = r * (float)s;
    }
}

// randForce/randTorque amplitude (BrownianForceSystem.java:95-104): ((tS*bfm)*sqrt(g))*gauss.
// The reference's (float)Math.sqrt((double)x) is sqrtf(x) for every float x: a square root rounded to
// double and then to float is correctly rounded (53 >= 2 * 24 + 2), and sqrtf is correctly rounded
// here (-fhip-fp32-correctly-rounded-divide-sqrt).
__device__ __forceinline__ float brown_amp(float scale, float bfm, float gam, float g) {
    return scale * bfm * sqrtf(gam) * g;
}

// ---------------------------------------------------------------------------------------
// Integrate + derive of one body (RigidRodLangevinIntegrationSystem.java:49-111 followed by
// DerivedGeometrySystem.java:43-71; the fused order of MechanicsFusion.integDerive,
// MechanicsFusion.java:129-200).  In: pose, lab-frame F/T, body-frame random F/T, drag.
// Out: new coord, u, y (re-orthogonalised), z, end1, end2.
struct Pose {
    float c[3], u[3], y[3];
};
struct Derived {
    float z[3], e1[3], e2[3];
};

__device__ __forceinline__ void integrate_derive(Pose &p, const float f[3], const float t[3], const float rf[3],
                                                 const float rt[3], const float gT[3], const float gR[3], float dt,
                                                 float segLen, Derived &d) {
    float ux = p.u[0], uyc = p.u[1], uzc = p.u[2];
    float yx = p.y[0], yyc = p.y[1], yzc = p.y[2];
    float zx = uyc * yzc - uzc * yyc;
    float zy = uzc * yx - ux * yzc;
    float zz = ux * yyc - uyc * yx;
    float zlen = 1.0f / sqrtf(zx * zx + zy * zy + zz * zz);
    zx *= zlen; zy *= zlen; zz *= zlen;
    float fx = f[0], fy = f[1], fz = f[2], tx = t[0], ty = t[1], tz = t[2];
    float bfx = ux * fx + uyc * fy + uzc * fz + rf[0];
    float bfy = yx * fx + yyc * fy + yzc * fz + rf[1];
    float bfz = zx * fx + zy * fy + zz * fz + rf[2];
    float btx = ux * tx + uyc * ty + uzc * tz + rt[0];
    float bty = yx * tx + yyc * ty + yzc * tz + rt[1];
    float btz = zx * tx + zy * ty + zz * tz + rt[2];
    float bvx = 1.0e6f * bfx / gT[0];
    float bvy = 1.0e6f * bfy / gT[1];
    float bvz = 1.0e6f * bfz / gT[2];
    float bwx = btx / gR[0];
    float bwy = bty / gR[1];
    float bwz = btz / gR[2];
    float vx = ux * bvx + yx * bvy + zx * bvz;
    float vy = uyc * bvx + yyc * bvy + zy * bvz;
    float vz = uzc * bvx + yzc * bvy + zz * bvz;
    p.c[0] = p.c[0] + dt * vx;
    p.c[1] = p.c[1] + dt * vy;
    p.c[2] = p.c[2] + dt * vz;
    float uTransInZ = -bwy * dt;
    float uTransInY = bwz * dt;
    float nuX = ux + yx * uTransInY + zx * uTransInZ;
    float nuY = uyc + yyc * uTransInY + zy * uTransInZ;
    float nuZ = uzc + yzc * uTransInY + zz * uTransInZ;
    float nuInv = 1.0f / sqrtf(nuX * nuX + nuY * nuY + nuZ * nuZ);
    float nux = nuX * nuInv, nuy = nuY * nuInv, nuz = nuZ * nuInv;
    float yTransInX = -uTransInY;
    float yTransInZ = bwx * dt;
    float nyX = ux * yTransInX + yx + zx * yTransInZ;
    float nyY = uyc * yTransInX + yyc + zy * yTransInZ;
    float nyZ = uzc * yTransInX + yzc + zz * yTransInZ;
    float nyInv = 1.0f / sqrtf(nyX * nyX + nyY * nyY + nyZ * nyZ);
    float ny0 = nyX * nyInv, ny1 = nyY * nyInv, ny2 = nyZ * nyInv;
    // derive
    float dzx = nuy * ny2 - nuz * ny1;
    float dzy = nuz * ny0 - nux * ny2;
    float dzz = nux * ny1 - nuy * ny0;
    float zmag2 = dzx * dzx + dzy * dzy + dzz * dzz;
    if (zmag2 > 0.0f) {
        float inv = 1.0f / sqrtf(zmag2);
        dzx *= inv; dzy *= inv; dzz *= inv;
    }
    d.z[0] = dzx; d.z[1] = dzy; d.z[2] = dzz;
    p.u[0] = nux; p.u[1] = nuy; p.u[2] = nuz;
    p.y[0] = dzy * nuz - dzz * nuy;
    p.y[1] = dzz * nux - dzx * nuz;
    p.y[2] = dzx * nuy - dzy * nux;
    float halfLen = segLen * 0.5f;
    d.e1[0] = p.c[0] - halfLen * nux;
    d.e1[1] = p.c[1] - halfLen * nuy;
    d.e1[2] = p.c[2] - halfLen * nuz;
    d.e2[0] = p.c[0] + halfLen * nux;
    d.e2[1] = p.c[1] + halfLen * nuy;
    d.e2[2] = p.c[2] + halfLen * nuz;
}

// ContainmentSystem.wallPush (ContainmentSystem.java:130-136)
__device__ __forceinline__ double wall_push(double d, double halfDim) {
    double s = (d > 0.0) ? 1.0 : ((d < 0.0) ? -1.0 : 0.0);
    double push = s * halfDim - d;
    double ps = (push > 0.0) ? 1.0 : ((push < 0.0) ? -1.0 : 0.0);
    if (ps == s) return 0.0;
    return push;
}

// Confine one body (ContainmentSystem.java:76-124); returns true when it touched a wall and
// adds into f/t as (float)((double)f + df).  half* are the inset half-extents.
__device__ __forceinline__ bool confine_body(const float c[3], const float u[3], float segLen, float gTx, double tau,
                                             double halfX, double halfY, double halfZ, double coeff, float f[3],
                                             float t[3]) {
    double cx = c[0], cy = c[1], cz = c[2];
    double ux = u[0], uy = u[1], uz = u[2];
    double half = 0.5 * (double)segLen;
    double bTGx = gTx;
    double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
    bool touched = false;
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
            touched = true;
        }
    }
    if (touched) {
        f[0] = (float)((double)f[0] + fx);
        f[1] = (float)((double)f[1] + fy);
        f[2] = (float)((double)f[2] + fz);
        t[0] = (float)((double)t[0] + tx);
        t[1] = (float)((double)t[1] + ty);
        t[2] = (float)((double)t[2] + tz);
    }
    return touched;
}

// ---------------------------------------------------------------------------------------
// Chain (ChainBendingForceSystem.java:47-331)
__device__ __forceinline__ double accurate_acos(double x) {
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
    double s, c;
    sincos(y, &s, &c);
    if (s > 1.0e-12 || s < -1.0e-12) y = y + (c - x) / s;
    sincos(y, &s, &c);
    if (s > 1.0e-12 || s < -1.0e-12) y = y + (c - x) / s;
    return y;
}

__device__ __forceinline__ double accurate_asin(double s) {
    double s2 = s * s;
    double y = s * (1.0 + s2 * (0.16666666666666666 + s2 * (0.075 + s2 * 0.044642857142857)));
    double sy, cy;
    sincos(y, &sy, &cy);
    if (cy > 1.0e-12) y = y + (s - sy) / cy;
    sincos(y, &sy, &cy);
    if (cy > 1.0e-12) y = y + (s - sy) / cy;
    return y;
}

__device__ __forceinline__ double angle_from_sin_cos(double sin2, double c) {
    double s = sqrt(sin2);
    double ac = (c < 0.0) ? -c : c;
    if (s <= ac) {
        double base = accurate_asin(s);
        return (c >= 0.0) ? base : (SBX_PI_D - base);
    }
    return accurate_acos(c);
}

struct ChainParams {
    double dt, fracMove, fracR, fracMoveTorq, springActive, spring, monoR;
};

// The terms of one joint of chainForces that the two rods of the joint share: the bond force F on
// this rod (the damping-limited link, :182-209) and the torsion torque Tq (:215-234; `torsion` is 0
// when |mu x nu|^2 <= 1e-30 and nothing is added).  endSign +1 = end2 block (:152-235), -1 = end1
// block (:237-321).  Evaluated from the neighbour's side, the same joint gives exactly -F and -Tq:
// the link and tip points, the strain, moveC1 + moveC2 (the two terms swap), forceMag, |tv|^2, the
// dot product and the torsion magnitude are the same floating-point operations on the same
// values, fSign flips, and tv = mu x nuE flips with every product of it unchanged.
// The reference squares Math.sin(accurateAcos(cosB)) (ChainBendingForceSystem.java:196-197,
// :202-203); sin^2(acos(cosB)) = (1 - cosB)(1 + cosB) on the clamped range, evaluated in that form:
// two roundings instead of six transcendental evaluations, a few double ulps from the reference's.
struct JointTerms {
    double F[3], Tq[3];
    int torsion;
};

__device__ __forceinline__ JointTerms chain_joint_terms(int i, int nbr, int side, double endSign, double cx, double cy,
                                                        double cz, double ux, double uy, double uz, double len,
                                                        double rBTGx, double rBTGy, double rBRGy, double ncx,
                                                        double ncy, double ncz, double nux, double nuy, double nuz,
                                                        double nlen, double nBTGx, double nBTGy, double nBRGy,
                                                        const ChainParams &cp) {
    const double DEG2RAD = SBX_PI_D / 180.0;
    const double RAD2DEG = 180.0 / SBX_PI_D;
    JointTerms J;
    double halfLen_um = 0.5 * len;
    double lSqSelf = 1.0e-12 * len * len;
    double nHalfLen_um = 0.5 * nlen;
    double lSqN = 1.0e-12 * nlen * nlen;
    double ex, ey, ez, selfLpx, selfLpy, selfLpz;
    if (endSign > 0) {
        ex = cx + halfLen_um * ux; ey = cy + halfLen_um * uy; ez = cz + halfLen_um * uz;
        selfLpx = ex - cp.monoR * ux; selfLpy = ey - cp.monoR * uy; selfLpz = ez - cp.monoR * uz;
    } else {
        ex = cx - halfLen_um * ux; ey = cy - halfLen_um * uy; ez = cz - halfLen_um * uz;
        selfLpx = ex + cp.monoR * ux; selfLpy = ey + cp.monoR * uy; selfLpz = ez + cp.monoR * uz;
    }
    double nbrTipx, nbrTipy, nbrTipz, nbrLpx, nbrLpy, nbrLpz;
    if (side == 0) {
        nbrTipx = ncx - nHalfLen_um * nux; nbrTipy = ncy - nHalfLen_um * nuy; nbrTipz = ncz - nHalfLen_um * nuz;
        nbrLpx = nbrTipx + cp.monoR * nux; nbrLpy = nbrTipy + cp.monoR * nuy; nbrLpz = nbrTipz + cp.monoR * nuz;
    } else {
        nbrTipx = ncx + nHalfLen_um * nux; nbrTipy = ncy + nHalfLen_um * nuy; nbrTipz = ncz + nHalfLen_um * nuz;
        nbrLpx = nbrTipx - cp.monoR * nux; nbrLpy = nbrTipy - cp.monoR * nuy; nbrLpz = nbrTipz - cp.monoR * nuz;
    }
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
    double cosA1_2 = (1.0 - cosB1) * (1.0 + cosB1);  // sin^2(acos(cosB1)), an identity on [-1, 1]
    double moveC1 = cosB1 * cosB1 / rBTGx + cosA1_2 / rBTGy + lSqSelf * cosA1_2 / (4.0 * rBRGy);
    double cosB2 = nux * luX + nuy * luY + nuz * luZ;
    if (cosB2 > 1.0) cosB2 = 1.0;
    if (cosB2 < -1.0) cosB2 = -1.0;
    double cosA2_2 = (1.0 - cosB2) * (1.0 + cosB2);  // sin^2(acos(cosB2)), an identity on [-1, 1]
    double moveC2 = cosB2 * cosB2 / nBTGx + cosA2_2 / nBTGy + lSqN * cosA2_2 / (4.0 * nBRGy);
    double denom = cp.dt * (moveC1 + moveC2);
    double forceMag = (denom > 0.0) ? (cp.fracMove * 1.0e-6 * strainDist / denom) : 0.0;
    double fSign = (i < nbr) ? 1.0 : -1.0;
    J.F[0] = fSign * forceMag * luX; J.F[1] = fSign * forceMag * luY; J.F[2] = fSign * forceMag * luZ;
    double nuxE, nuyE, nuzE;
    if (side == 0) { nuxE = nux; nuyE = nuy; nuzE = nuz; } else { nuxE = -nux; nuyE = -nuy; nuzE = -nuz; }
    double mux, muy, muz;
    if (endSign > 0) { mux = ux; muy = uy; muz = uz; } else { mux = -ux; muy = -uy; muz = -uz; }
    double tvx = muy * nuzE - muz * nuyE, tvy = muz * nuxE - mux * nuzE, tvz = mux * nuyE - muy * nuxE;
    double tvMag2 = tvx * tvx + tvy * tvy + tvz * tvz;
    J.torsion = tvMag2 > 1.0e-30;
    J.Tq[0] = J.Tq[1] = J.Tq[2] = 0.0;
    if (J.torsion) {
        double invMag = 1.0 / sqrt(tvMag2);
        tvx *= invMag; tvy *= invMag; tvz *= invMag;
        double dotV = mux * nuxE + muy * nuyE + muz * nuzE;
        if (dotV > 1.0) dotV = 1.0;
        if (dotV < -1.0) dotV = -1.0;
        double angTween = angle_from_sin_cos(tvMag2, dotV) * RAD2DEG;
        double torsionMag;
        if (cp.springActive > 0.5) {
            torsionMag = cp.fracMoveTorq * cp.spring * angTween;
        } else {
            double invBRG = 1.0 / rBRGy + 1.0 / nBRGy;
            torsionMag = cp.fracMoveTorq * DEG2RAD * angTween / (invBRG * cp.dt);
        }
        J.Tq[0] = tvx * torsionMag; J.Tq[1] = tvy * torsionMag; J.Tq[2] = tvz * torsionMag;
    }
    return J;
}

// The joint's contribution to this rod, in the reference's order: F, the lever-arm torque R x F
// (R = +-1/2 L fracR u), then the torsion torque.
__device__ __forceinline__ void chain_joint_apply(double endSign, double len, double ux, double uy, double uz,
                                                  const ChainParams &cp, const JointTerms &J, double f[3],
                                                  double t[3]) {
    const double Fx = J.F[0], Fy = J.F[1], Fz = J.F[2];
    f[0] += Fx; f[1] += Fy; f[2] += Fz;
    double Rscale = (endSign > 0) ? (0.5e-6 * len * cp.fracR) : (-0.5e-6 * len * cp.fracR);
    double Rx = Rscale * ux, Ry = Rscale * uy, Rz = Rscale * uz;
    t[0] += Ry * Fz - Rz * Fy; t[1] += Rz * Fx - Rx * Fz; t[2] += Rx * Fy - Ry * Fx;
    if (J.torsion) { t[0] += J.Tq[0]; t[1] += J.Tq[1]; t[2] += J.Tq[2]; }
}

// One joint of chainForces evaluated and applied by this rod.
__device__ __forceinline__ void chain_joint(int i, int nbr, int side, double endSign, double cx, double cy, double cz,
                                            double ux, double uy, double uz, double len, double rBTGx, double rBTGy,
                                            double rBRGy, double ncx, double ncy, double ncz, double nux, double nuy,
                                            double nuz, double nlen, double nBTGx, double nBTGy, double nBRGy,
                                            const ChainParams &cp, double f[3], double t[3]) {
    const JointTerms J = chain_joint_terms(i, nbr, side, endSign, cx, cy, cz, ux, uy, uz, len, rBTGx, rBTGy, rBRGy,
                                           ncx, ncy, ncz, nux, nuy, nuz, nlen, nBTGx, nBTGy, nBRGy, cp);
    chain_joint_apply(endSign, len, ux, uy, uz, cp, J, f, t);
}

// Whole chain contribution of body i reading planar arrays (stride n).
__device__ __forceinline__ void chain_body(int n, int i, const float *__restrict__ coord, const float *__restrict__ uVec,
                                           const float *__restrict__ segLength, int e2Slot, int e2Side, int e1Slot,
                                           int e1Side, const float *__restrict__ bTransGam,
                                           const float *__restrict__ bRotGam, const ChainParams &cp, double f[3],
                                           double t[3]) {
    double cx = coord[i], cy = coord[n + i], cz = coord[2 * n + i];
    double ux = uVec[i], uy = uVec[n + i], uz = uVec[2 * n + i];
    double len = segLength[i];
    double rBTGx = bTransGam[i], rBTGy = bTransGam[n + i], rBRGy = bRotGam[n + i];
    if (e2Slot >= 0) {
        int j = e2Slot;
        chain_joint(i, j, e2Side, +1.0, cx, cy, cz, ux, uy, uz, len, rBTGx, rBTGy, rBRGy, coord[j], coord[n + j],
                    coord[2 * n + j], uVec[j], uVec[n + j], uVec[2 * n + j], segLength[j], bTransGam[j],
                    bTransGam[n + j], bRotGam[n + j], cp, f, t);
    }
    if (e1Slot >= 0) {
        int j = e1Slot;
        chain_joint(i, j, e1Side, -1.0, cx, cy, cz, ux, uy, uz, len, rBTGx, rBTGy, rBRGy, coord[j], coord[n + j],
                    coord[2 * n + j], uVec[j], uVec[n + j], uVec[2 * n + j], segLength[j], bTransGam[j],
                    bTransGam[n + j], bRotGam[n + j], cp, f, t);
    }
}

// Rods finished per wave by chain_body_wave: lane 0 of a wave holds the previous wave's last rod
// and only evaluates its end2 joint for lane 1.
constexpr int CHAIN_SPAN = 63;

// chain_body for the rods of one wave laid out as rod = 63 * wave + lane - 1.  Every lane evaluates
// the joint at its rod's end2; a rod whose end1 is joined to the end2 of the rod in the lane below
// takes that joint's terms negated (exactly its own evaluation, see JointTerms), any other end1
// joint is evaluated by the rod itself.  One joint evaluation per rod instead of two.  Called by
// every lane of the wave (the shuffles read all lanes); `live` = the lane holds a rod, `own` = the
// lane finishes it (f, t are only meaningful there).  c, u, len, gT, gR: the rod's own data.
__device__ __forceinline__ void chain_body_wave(int n, int i, bool live, bool own, int lane, const float c[3],
                                                const float u[3], float len, const float gT[3], const float gR[3],
                                                const float *__restrict__ coord, const float *__restrict__ uVec,
                                                const float *__restrict__ segLength,
                                                const float *__restrict__ bTransGam,
                                                const float *__restrict__ bRotGam, int e2Slot, int e2Side,
                                                int e1Slot, int e1Side, const ChainParams &cp, double f[3],
                                                double t[3]) {
    const int j2 = e2Slot >= 0 ? e2Slot : i, j1 = e1Slot >= 0 ? e1Slot : i;
    // both neighbours' data, issued before either joint is evaluated
    float n2[9], n1[9];
    n2[0] = coord[j2]; n2[1] = coord[n + j2]; n2[2] = coord[2 * n + j2];
    n2[3] = uVec[j2]; n2[4] = uVec[n + j2]; n2[5] = uVec[2 * n + j2];
    n2[6] = segLength[j2]; n2[7] = bTransGam[j2]; n2[8] = bTransGam[n + j2];
    n1[0] = coord[j1]; n1[1] = coord[n + j1]; n1[2] = coord[2 * n + j1];
    n1[3] = uVec[j1]; n1[4] = uVec[n + j1]; n1[5] = uVec[2 * n + j1];
    n1[6] = segLength[j1]; n1[7] = bTransGam[j1]; n1[8] = bTransGam[n + j1];
    const float nR2 = bRotGam[n + j2], nR1 = bRotGam[n + j1];
    const double cx = c[0], cy = c[1], cz = c[2], ux = u[0], uy = u[1], uz = u[2], L = len;
    const double rBTGx = gT[0], rBTGy = gT[1], rBRGy = gR[1];
    const bool has2 = live && e2Slot >= 0;
    JointTerms J2;
    J2.F[0] = J2.F[1] = J2.F[2] = 0.0;
    J2.Tq[0] = J2.Tq[1] = J2.Tq[2] = 0.0;
    J2.torsion = 0;
    if (has2)
        J2 = chain_joint_terms(i, j2, e2Side, +1.0, cx, cy, cz, ux, uy, uz, L, rBTGx, rBTGy, rBRGy, n2[0], n2[1],
                               n2[2], n2[3], n2[4], n2[5], n2[6], n2[7], n2[8], nR2, cp);
    const int src = lane > 0 ? lane - 1 : 0;
    const int pSlot = __shfl(has2 ? e2Slot : -1, src, 64);
    const int pSide = __shfl(e2Side, src, 64);
    JointTerms P;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        P.F[k] = -__shfl(J2.F[k], src, 64);
        P.Tq[k] = -__shfl(J2.Tq[k], src, 64);
    }
    P.torsion = __shfl(J2.torsion, src, 64);
    if (!own) return;
    if (has2) chain_joint_apply(+1.0, L, ux, uy, uz, cp, J2, f, t);
    if (e1Slot >= 0) {
        if (lane > 0 && e1Slot == i - 1 && e1Side == 1 && pSlot == i && pSide == 0) {
            chain_joint_apply(-1.0, L, ux, uy, uz, cp, P, f, t);
        } else {
            const JointTerms J1 = chain_joint_terms(i, j1, e1Side, -1.0, cx, cy, cz, ux, uy, uz, L, rBTGx, rBTGy,
                                                    rBRGy, n1[0], n1[1], n1[2], n1[3], n1[4], n1[5], n1[6], n1[7],
                                                    n1[8], nR1, cp);
            chain_joint_apply(-1.0, L, ux, uy, uz, cp, J1, f, t);
        }
    }
}

// CrosslinkerSystem.fastAcos (CrosslinkerSystem.java:372-376)
__device__ __forceinline__ double fast_acos(double dot) {
    if (dot > 0.95) { double t = 1.0 - dot; if (t < 0) t = 0; return sqrt(2.0 * t); }
    if (dot < -0.95) { double t = 1.0 + dot; if (t < 0) t = 0; return SBX_PI_D - sqrt(2.0 * t); }
    return accurate_acos(dot);
}

}  // namespace sbx
