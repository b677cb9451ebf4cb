// sbx_mechanics.hip -- per-body rigid-rod Systems as gfx950 kernels behind the C ABI.
//
// One thread per body, planar SoA (coalesced x/y/z plane streams), wave64.  Each kernel is the
// drop-in for one reference System method; the arithmetic lives in sbx_device.h so the fused
// network step (sbx_network.hip) runs the identical code.
#include "sbx_device.h"
#include "sbx_internal.h"

using namespace sbx;

namespace {

constexpr int T = 256;

__global__ void __launch_bounds__(T) k_zero_accumulators(float *__restrict__ forceSum, float *__restrict__ torqueSum,
                                                         int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        forceSum[a * n + i] = 0.0f;
        torqueSum[a * n + i] = 0.0f;
    }
}

__global__ void __launch_bounds__(T) k_brownian(float *__restrict__ randForce, float *__restrict__ randTorque,
                                                const float *__restrict__ bTransGam, const float *__restrict__ bRotGam,
                                                const float *__restrict__ bts, const float *__restrict__ brs,
                                                float bfm, int step, int seed, int rng, uint32_t stream, int n,
                                                float *__restrict__ forceSum, float *__restrict__ torqueSum) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    if (forceSum) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            forceSum[a * n + i] = 0.0f;
            torqueSum[a * n + i] = 0.0f;
        }
    }
    float gf[3], gt[3];
    brownian_noise(rng, (uint32_t)i, step, seed, stream, gf, gt);
    float tS = bts[i], rS = brs[i];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        randForce[a * n + i] = brown_amp(tS, bfm, bTransGam[a * n + i], gf[a]);
        randTorque[a * n + i] = brown_amp(rS, bfm, bRotGam[a * n + i], gt[a]);
    }
}

// Stand-alone integrate: RigidRodLangevinIntegrationSystem.java:49-111 (y is the integrate
// output, not yet re-orthogonalised; derive does that).
__global__ void __launch_bounds__(T) k_integrate(float *__restrict__ coord, float *__restrict__ uVec,
                                                 float *__restrict__ yVec, const float *__restrict__ forceSum,
                                                 const float *__restrict__ torqueSum,
                                                 const float *__restrict__ randForce,
                                                 const float *__restrict__ randTorque,
                                                 const float *__restrict__ bTransGam,
                                                 const float *__restrict__ bRotGam, float dt, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
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

// DerivedGeometrySystem.java:43-71
__global__ void __launch_bounds__(T) k_derive(const float *__restrict__ coord, const float *__restrict__ uVec,
                                              float *__restrict__ yVec, float *__restrict__ zVec,
                                              float *__restrict__ end1, float *__restrict__ end2,
                                              const float *__restrict__ segLength, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
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
    zVec[i] = zx; zVec[iy] = zy; zVec[iz] = zz;
    yVec[i] = zy * uzc - zz * uyc;
    yVec[iy] = zz * ux - zx * uzc;
    yVec[iz] = zx * uyc - zy * ux;
    float cx = coord[i], cy = coord[iy], cz = coord[iz];
    float halfLen = segLength[i] * 0.5f;
    end1[i] = cx - halfLen * ux;
    end1[iy] = cy - halfLen * uyc;
    end1[iz] = cz - halfLen * uzc;
    end2[i] = cx + halfLen * ux;
    end2[iy] = cy + halfLen * uyc;
    end2[iz] = cz + halfLen * uzc;
}

struct BoxArgs {
    double tau, halfX, halfY, halfZ, coeff;
};

static BoxArgs box_args(const float *boxParams) {
    BoxArgs b;
    b.tau = boxParams[0];
    b.halfX = 0.5 * boxParams[1] - boxParams[4];
    b.halfY = 0.5 * boxParams[2] - boxParams[4];
    b.halfZ = 0.5 * boxParams[3] - boxParams[4];
    b.coeff = boxParams[5];
    return b;
}

static bool confine_fires(const float *boxParams, int step) {
    int checkInt = (int)boxParams[6];
    return !(checkInt > 1 && (step % checkInt) != 0);
}

// ContainmentSystem.java:75-125
__global__ void __launch_bounds__(T) k_confine(const float *__restrict__ coord, const float *__restrict__ uVec,
                                               const float *__restrict__ segLength,
                                               const float *__restrict__ bTransGam, float *__restrict__ forceSum,
                                               float *__restrict__ torqueSum, BoxArgs b, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    float c[3] = {coord[i], coord[n + i], coord[2 * n + i]};
    float u[3] = {uVec[i], uVec[n + i], uVec[2 * n + i]};
    float f[3] = {forceSum[i], forceSum[n + i], forceSum[2 * n + i]};
    float t[3] = {torqueSum[i], torqueSum[n + i], torqueSum[2 * n + i]};
    if (confine_body(c, u, segLength[i], bTransGam[i], b.tau, b.halfX, b.halfY, b.halfZ, b.coeff, f, t)) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            forceSum[a * n + i] = f[a];
            torqueSum[a * n + i] = t[a];
        }
    }
}

// MechanicsFusion.integDerive / integDeriveConfined (MechanicsFusion.java:121-338)
__global__ void __launch_bounds__(T) k_integ_derive(float *__restrict__ coord, float *__restrict__ uVec,
                                                    float *__restrict__ yVec, float *__restrict__ zVec,
                                                    const float *__restrict__ segLength, float *__restrict__ end1,
                                                    float *__restrict__ end2, float *__restrict__ forceSum,
                                                    float *__restrict__ torqueSum,
                                                    const float *__restrict__ randForce,
                                                    const float *__restrict__ randTorque,
                                                    const float *__restrict__ bTransGam,
                                                    const float *__restrict__ bRotGam, float dt, int doConfine,
                                                    BoxArgs b, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    Pose p;
    float f[3], t[3], rf[3], rt[3], gT[3], gR[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        p.c[a] = coord[a * n + i];
        p.u[a] = uVec[a * n + i];
        p.y[a] = yVec[a * n + i];
        f[a] = forceSum[a * n + i];
        t[a] = torqueSum[a * n + i];
        rf[a] = randForce[a * n + i];
        rt[a] = randTorque[a * n + i];
        gT[a] = bTransGam[a * n + i];
        gR[a] = bRotGam[a * n + i];
    }
    float segLen = segLength[i];
    if (doConfine) {
        if (confine_body(p.c, p.u, segLen, gT[0], b.tau, b.halfX, b.halfY, b.halfZ, b.coeff, f, t)) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                forceSum[a * n + i] = f[a];
                torqueSum[a * n + i] = t[a];
            }
        }
    }
    Derived d;
    integrate_derive(p, f, t, rf, rt, gT, gR, dt, segLen, d);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        coord[a * n + i] = p.c[a];
        uVec[a * n + i] = p.u[a];
        yVec[a * n + i] = p.y[a];
        zVec[a * n + i] = d.z[a];
        end1[a * n + i] = d.e1[a];
        end2[a * n + i] = d.e2[a];
    }
}

// ChainBendingForceSystem.chainForces (ChainBendingForceSystem.java:133-330)
__global__ void __launch_bounds__(T) k_chain(const float *__restrict__ coord, const float *__restrict__ uVec,
                                             const float *__restrict__ segLength,
                                             const int *__restrict__ end2NbrSlot, const int *__restrict__ end2NbrSide,
                                             const int *__restrict__ end1NbrSlot, const int *__restrict__ end1NbrSide,
                                             const float *__restrict__ bTransGam, const float *__restrict__ bRotGam,
                                             float *__restrict__ forceSum, float *__restrict__ torqueSum,
                                             ChainParams cp, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
    chain_body(n, i, coord, uVec, segLength, end2NbrSlot[i], end2NbrSide[i], end1NbrSlot[i], end1NbrSide[i],
               bTransGam, bRotGam, cp, f, t);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        forceSum[a * n + i] = (float)(forceSum[a * n + i] + f[a]);
        torqueSum[a * n + i] = (float)(torqueSum[a * n + i] + t[a]);
    }
}

}  // namespace

namespace sbx {
ChainParams chain_params(const float *cp) {
    ChainParams c;
    c.dt = cp[0]; c.fracMove = cp[1]; c.fracR = cp[2]; c.fracMoveTorq = cp[3];
    c.springActive = cp[4]; c.spring = cp[5]; c.monoR = cp[6];
    return c;
}
}  // namespace sbx

extern "C" {

int sbx_zero_accumulators(sbx_ctx *c, float *forceSum, float *torqueSum, int n) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0, "zero_accumulators: n < 0");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, forceSum && torqueSum, "zero_accumulators: null array");
    hipLaunchKernelGGL(k_zero_accumulators, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, forceSum, torqueSum, n);
    return check_launch(c, "k_zero_accumulators");
}

int sbx_brownian_force(sbx_ctx *c, float *randForce, float *randTorque, const float *bTransGam, const float *bRotGam,
                       const float *brownTransScale, const float *brownRotScale, const float *params,
                       const int *counts, int n, int rng, uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && params && counts, "brownian_force: bad n or null params/counts");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "brownian_force: unknown rng %d", rng);
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, randForce && randTorque && bTransGam && bRotGam && brownTransScale && brownRotScale,
                "brownian_force: null array");
    hipLaunchKernelGGL(k_brownian, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, randForce, randTorque, bTransGam,
                       bRotGam, brownTransScale, brownRotScale, params[1], counts[1], counts[2], rng, stream, n,
                       (float *)nullptr, (float *)nullptr);
    return check_launch(c, "k_brownian");
}

int sbx_force_build(sbx_ctx *c, float *forceSum, float *torqueSum, float *randForce, float *randTorque,
                    const float *bTransGam, const float *bRotGam, const float *brownTransScale,
                    const float *brownRotScale, const float *params, const int *counts, int n, int rng,
                    uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && params && counts, "force_build: bad n or null params/counts");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "force_build: unknown rng %d", rng);
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, forceSum && torqueSum && randForce && randTorque && bTransGam && bRotGam && brownTransScale &&
                       brownRotScale,
                "force_build: null array");
    hipLaunchKernelGGL(k_brownian, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, randForce, randTorque, bTransGam,
                       bRotGam, brownTransScale, brownRotScale, params[1], counts[1], counts[2], rng, stream, n,
                       forceSum, torqueSum);
    return check_launch(c, "k_brownian(force_build)");
}

int sbx_integrate(sbx_ctx *c, float *coord, float *uVec, float *yVec, const float *forceSum, const float *torqueSum,
                  const float *randForce, const float *randTorque, const float *bTransGam, const float *bRotGam,
                  const float *params, const int *counts, int n) {
    (void)counts;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && params, "integrate: bad n or null params");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && yVec && forceSum && torqueSum && randForce && randTorque && bTransGam && bRotGam,
                "integrate: null array");
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, yVec, forceSum,
                       torqueSum, randForce, randTorque, bTransGam, bRotGam, params[0], n);
    return check_launch(c, "k_integrate");
}

int sbx_derive(sbx_ctx *c, const float *coord, const float *uVec, float *yVec, float *zVec, float *end1, float *end2,
               const float *segLength, const int *counts, int n) {
    (void)counts;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0, "derive: n < 0");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && yVec && zVec && end1 && end2 && segLength, "derive: null array");
    hipLaunchKernelGGL(k_derive, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, yVec, zVec, end1, end2,
                       segLength, n);
    return check_launch(c, "k_derive");
}

int sbx_integ_derive(sbx_ctx *c, float *coord, float *uVec, float *yVec, float *zVec, const float *segLength,
                     float *end1, float *end2, const float *forceSum, const float *torqueSum, const float *randForce,
                     const float *randTorque, const float *bTransGam, const float *bRotGam, const float *params,
                     int n) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && params, "integ_derive: bad n or null params");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && yVec && zVec && segLength && end1 && end2 && forceSum && torqueSum && randForce &&
                       randTorque && bTransGam && bRotGam,
                "integ_derive: null array");
    BoxArgs b = {1.0, 0, 0, 0, 0};
    hipLaunchKernelGGL(k_integ_derive, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, yVec, zVec,
                       segLength, end1, end2, (float *)forceSum, (float *)torqueSum, randForce, randTorque, bTransGam,
                       bRotGam, params[0], 0, b, n);
    return check_launch(c, "k_integ_derive");
}

int sbx_integ_derive_confined(sbx_ctx *c, float *coord, float *uVec, float *yVec, float *zVec,
                              const float *segLength, float *end1, float *end2, float *forceSum, float *torqueSum,
                              const float *randForce, const float *randTorque, const float *bTransGam,
                              const float *bRotGam, const float *boxParams, const int *counts, int n) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && boxParams && counts, "integ_derive_confined: bad n or null params");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && yVec && zVec && segLength && end1 && end2 && forceSum && torqueSum && randForce &&
                       randTorque && bTransGam && bRotGam,
                "integ_derive_confined: null array");
    SBX_REQUIRE(c, boxParams[0] != 0.0f, "integ_derive_confined: tau must be non-zero");
    BoxArgs b = box_args(boxParams);
    hipLaunchKernelGGL(k_integ_derive, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, yVec, zVec,
                       segLength, end1, end2, forceSum, torqueSum, randForce, randTorque, bTransGam, bRotGam,
                       boxParams[7], confine_fires(boxParams, counts[1]) ? 1 : 0, b, n);
    return check_launch(c, "k_integ_derive(confined)");
}

int sbx_confine(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                float *forceSum, float *torqueSum, const float *boxParams, const int *counts, int n) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && boxParams && counts, "confine: bad n or null params");
    if (n == 0 || !confine_fires(boxParams, counts[1])) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && segLength && bTransGam && forceSum && torqueSum, "confine: null array");
    SBX_REQUIRE(c, boxParams[0] != 0.0f, "confine: tau must be non-zero");
    hipLaunchKernelGGL(k_confine, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, segLength, bTransGam,
                       forceSum, torqueSum, box_args(boxParams), n);
    return check_launch(c, "k_confine");
}

int sbx_chain_forces(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const int *end2NbrSlot, const int *end2NbrSide, const int *end1NbrSlot, const int *end1NbrSide,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const float *chainParams, const int *counts, int n) {
    (void)counts;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, n >= 0 && chainParams, "chain_forces: bad n or null chainParams");
    if (n == 0) return SBX_OK;
    SBX_REQUIRE(c, coord && uVec && segLength && end2NbrSlot && end2NbrSide && end1NbrSlot && end1NbrSide &&
                       bTransGam && bRotGam && forceSum && torqueSum,
                "chain_forces: null array");
    hipLaunchKernelGGL(k_chain, dim3(blocks_for(n, T)), dim3(T), 0, c->stream, coord, uVec, segLength, end2NbrSlot,
                       end2NbrSide, end1NbrSlot, end1NbrSide, bTransGam, bRotGam, forceSum, torqueSum,
                       chain_params(chainParams), n);
    return check_launch(c, "k_chain");
}

}  // extern "C"
