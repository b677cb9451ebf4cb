// sbx_motor.hip -- myosin Systems behind the C ABI: binding detection, catch-slip release, the
// nucleotide cycle, the cross-bridge bond and its segment gather, motor joints, dimer coupling and
// the minifilament tether + backbone gather.  One thread per motor / sub-body / dimer / segment,
// each writing only the slots it owns (the reference's one-writer-per-slot discipline, no atomics).
#include "sbx_internal.h"
#include "sbx_motor.h"

namespace sbx {

KinParams kin_params(const float *kp) {
    KinParams k;
    k.kOff = kp[0]; k.aCatch = kp[1]; k.aSlip = kp[2]; k.xCatch = kp[3]; k.xSlip = kp[4];
    k.kT = kp[5]; k.dt = kp[6]; k.myoColTol = kp[7]; k.alignTol = kp[8];
    k.refractorySteps = (int)kp[10];
    k.breakForceN = kp[11];
    k.capOn = kp[12] > 0.5f ? 1 : 0;
    k.blockProb = kp[13];
    return k;
}
NucParams nuc_params(const float *p) {
    NucParams n = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    return n;
}
XbParams xb_params(const float *p) {
    XbParams x;
    x.myoSpring = p[0]; x.j1FMT = p[2]; x.dt = p[3]; x.headLen = p[4]; x.xbias = p[5];
    return x;
}
JointParams joint_params(const float *p) {
    JointParams j;
    j.dt = p[0]; j.j1FracMove = p[1]; j.j1FracR = p[2]; j.j1FracMoveTorq = p[3];
    j.j2FracMove = p[5]; j.j2FracR = p[6]; j.j2FracMoveTorq = p[7]; j.j2Rest = p[8]; j.stallPN = p[10];
    return j;
}
DimerParams dimer_params(const float *p) {
    DimerParams d = {p[0], p[1], p[2], p[3], p[4]};
    return d;
}
MiniParams mini_params(const float *p) {
    MiniParams m = {p[0], p[1], p[2]};
    return m;
}

}  // namespace sbx

using namespace sbx;

namespace {

constexpr int T = 256;

__global__ void __launch_bounds__(T) k_publish_head(const float *__restrict__ bc, const float *__restrict__ bu,
                                                    const float *__restrict__ bl, float *__restrict__ head,
                                                    float *__restrict__ hu, float *__restrict__ ru, int nM) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    int nB = 3 * nM, h = 3 * m + 2, r = 3 * m;
    float hl = bl[h];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        head[a * nM + m] = bc[a * nB + h] + 0.5f * hl * bu[a * nB + h];  // MotorStore.java:309-311
        hu[a * nM + m] = bu[a * nB + h];
        ru[a * nM + m] = bu[a * nB + r];
    }
}

__global__ void __launch_bounds__(T) k_publish_motors(const float *__restrict__ head, const float *__restrict__ reach,
                                                      float *__restrict__ center, float *__restrict__ bound,
                                                      int *__restrict__ ownerStore, int *__restrict__ ownerSlot,
                                                      int nM, int cap, int base) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    int b = base + m;
    center[b] = head[m];
    center[cap + b] = head[nM + m];
    center[2 * cap + b] = head[2 * nM + m];
    bound[b] = reach[m];
    if (ownerStore) ownerStore[b] = 1;
    if (ownerSlot) ownerSlot[b] = m;
}

__global__ void __launch_bounds__(T) k_grid_reachable(const float *__restrict__ head, const float *__restrict__ hu,
                                                      const float *__restrict__ ru, const float *__restrict__ e1,
                                                      const float *__restrict__ e2, ReachGrid g,
                                                      const int *__restrict__ off, const int *__restrict__ cont,
                                                      const int *__restrict__ ownerStore,
                                                      const int *__restrict__ ownerSlot, int *__restrict__ reachSeg,
                                                      int *__restrict__ reachCount, int maxCand, float tol,
                                                      float align, int nM, int nSeg) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    reachCount[m] = grid_reachable_one(m, nM, nSeg, head, hu, ru, e1, e2, g, off, cont, ownerStore, ownerSlot,
                                       reachSeg, maxCand, tol, align);
}

__global__ void __launch_bounds__(T) k_bind_nearest(const float *__restrict__ head, const float *__restrict__ hu,
                                                    const float *__restrict__ ru, const float *__restrict__ e1,
                                                    const float *__restrict__ e2, const int *__restrict__ reachSeg,
                                                    const int *__restrict__ reachCount, int maxCand,
                                                    int *__restrict__ boundSeg, float *__restrict__ bindArc, float tol,
                                                    float align, int nM, int nSeg,
                                                    const int *__restrict__ seedNode) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    int bs = boundSeg[m];
    if (bs != FREE_BINDABLE) return;
    float arc = bindArc[m];
    bind_nearest_one(m, nM, nSeg, head, hu, ru, e1, e2, reachSeg, reachCount[m], maxCand, bs, arc, tol, align,
                     seedNode);
    boundSeg[m] = bs;
    bindArc[m] = arc;
}

// BindingDetectionSystem.bruteReachable / bruteReachableNodeAware (BindingDetectionSystem.java:260-283,
// 410-436): every segment in index order against one head; the segment loads are uniform across the wavefront
__global__ void __launch_bounds__(T) k_brute_reachable(const float *__restrict__ head, const float *__restrict__ hu,
                                                       const float *__restrict__ ru, const float *__restrict__ e1,
                                                       const float *__restrict__ e2, int *__restrict__ reachSeg,
                                                       int *__restrict__ reachCount, const int *__restrict__ seedNode,
                                                       float tol, float align, int nM, int nSeg, int maxCand) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    float mx = head[m], my = head[nM + m], mz = head[2 * nM + m];
    float mux = hu[m], muy = hu[nM + m], muz = hu[2 * nM + m];
    float rux = ru[m], ruy = ru[nM + m], ruz = ru[2 * nM + m];
    int out = 0;
    for (int s = 0; s < nSeg; s++) {
        if (seedNode && seedNode[s] >= 0) continue;
        float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, e1[s], e1[nSeg + s], e1[2 * nSeg + s], e2[s],
                             e2[nSeg + s], e2[2 * nSeg + s], tol, align);
        if (d >= 0.0f) {
            if (out < maxCand) reachSeg[(size_t)m * maxCand + out] = s;
            out++;
        }
    }
    reachCount[m] = out;
}

__global__ void __launch_bounds__(T) k_catch_slip(int *__restrict__ boundSeg, const float *__restrict__ fdf,
                                                  const float *__restrict__ fmag, int *__restrict__ cooldown,
                                                  int *__restrict__ stats, int *__restrict__ capStats, KinParams k,
                                                  int step, int seed, int rng, uint32_t stream, int nM) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    int bs = boundSeg[m], cool = cooldown[m];
    catch_slip_one(m, bs, cool, stats, capStats, fdf[m], fmag[m], k, step, seed, rng, stream);
    boundSeg[m] = bs;
    cooldown[m] = cool;
}

__global__ void __launch_bounds__(T) k_nucleotide(int *__restrict__ nuc, const int *__restrict__ boundSeg,
                                                  const float *__restrict__ hist, NucParams p, int step, int seed,
                                                  int rng, uint32_t stream, int nM) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    nuc[m] = nucleotide_one(m, nuc[m], boundSeg[m] >= 0, hist, p, step, seed, rng, stream);
}

__global__ void __launch_bounds__(T) k_bond(const float *__restrict__ mc, const float *__restrict__ mu,
                                            const float *__restrict__ my, const float *__restrict__ mbRG,
                                            const float *__restrict__ fc, const float *__restrict__ fu,
                                            const float *__restrict__ fy, const float *__restrict__ fbRG,
                                            const float *__restrict__ fLen, const int *__restrict__ boundSeg,
                                            const float *__restrict__ bindArc, const int *__restrict__ nuc,
                                            float *__restrict__ bondData, XbParams xp, int nM, int nSeg) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    float out[13];
#pragma unroll
    for (int k = 0; k < 13; k++) out[k] = 0.0f;
    int s = boundSeg[m];
    if (s >= 0) bond_one(m, nM, nSeg, s, mc, mu, my, mbRG, fc, fu, fy, fbRG, fLen, bindArc[m], nuc[m], xp, out);
    float *d = bondData + (size_t)m * XB_STRIDE;
#pragma unroll
    for (int k = 0; k < 13; k++) d[k] = out[k];
}

__global__ void __launch_bounds__(T) k_apply_head(const float *__restrict__ bondData, float *__restrict__ F,
                                                  float *__restrict__ Tq, int nM) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    int nB = 3 * nM, h = 3 * m + 2;
    const float *d = bondData + (size_t)m * XB_STRIDE;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        F[a * nB + h] = F[a * nB + h] + d[a];  // float + float (CrossBridgeSystem.java:150-155)
        Tq[a * nB + h] = Tq[a * nB + h] + d[3 + a];
    }
}

__global__ void __launch_bounds__(T) k_register_force_dot(const float *__restrict__ bondData,
                                                          const int *__restrict__ boundSeg, float *__restrict__ fdf,
                                                          float *__restrict__ fmag, float *__restrict__ hist,
                                                          int *__restrict__ place, int nM) {
    int m = blockIdx.x * T + threadIdx.x;
    if (m >= nM) return;
    if (boundSeg[m] >= 0) {
        const float *d = bondData + (size_t)m * XB_STRIDE;
        float fd = d[12];
        fdf[m] = fd;
        float fx = d[0], fy = d[1], fz = d[2];
        fmag[m] = (float)sqrt((double)(fx * fx + fy * fy + fz * fz));
        int p = place[m];
        hist[m * 10 + p] = fd;
        place[m] = (p + 1) % 10;
    } else {
        fdf[m] = 0.0f;
        fmag[m] = 0.0f;
        place[m] = 0;
        for (int k = 0; k < 10; k++) hist[m * 10 + k] = 0.0f;
    }
}

__global__ void __launch_bounds__(T) k_csr_gather(const int *__restrict__ off, const int *__restrict__ list,
                                                  const float *__restrict__ data, int stride, int doff,
                                                  float *__restrict__ F, float *__restrict__ Tq, int nKeys) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= nKeys) return;
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    csr_payload_sum(s, off, list, data, stride, doff, f, t);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        F[a * nKeys + s] = (float)(F[a * nKeys + s] + f[a]);
        Tq[a * nKeys + s] = (float)(Tq[a * nKeys + s] + t[a]);
    }
}

__global__ void __launch_bounds__(T) k_joints(const float *__restrict__ coord, const float *__restrict__ uVec,
                                              const float *__restrict__ len, const float *__restrict__ bTG,
                                              const float *__restrict__ bRG, float *__restrict__ F,
                                              float *__restrict__ Tq, const int *__restrict__ nuc, JointParams jp,
                                              int nB) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= nB) return;
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    joints_one(s, nB, coord, uVec, len, bTG, bRG, nuc[s / 3], jp, f, t);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        F[a * nB + s] = (float)(F[a * nB + s] + f[a]);
        Tq[a * nB + s] = (float)(Tq[a * nB + s] + t[a]);
    }
}

__global__ void __launch_bounds__(T) k_dimer(const float *__restrict__ coord, const float *__restrict__ uVec,
                                             const float *__restrict__ len, const float *__restrict__ bTG,
                                             const float *__restrict__ bRG, float *__restrict__ F,
                                             float *__restrict__ Tq, const int *__restrict__ motorA,
                                             const int *__restrict__ motorB, const int *__restrict__ parallel,
                                             const int *__restrict__ boundSeg, DimerParams dp, int nB, int nD) {
    int d = blockIdx.x * T + threadIdx.x;
    if (d >= nD) return;
    int mA = motorA[d], mB = motorB[d];
    bool both = boundSeg[mA] >= 0 && boundSeg[mB] >= 0;
    dimer_one(mA, mB, parallel[d] != 0, both, nB, coord, uVec, len, bTG, bRG, F, Tq, dp);
}

__global__ void __launch_bounds__(T) k_tether(const float *__restrict__ mC, const float *__restrict__ mU,
                                              const float *__restrict__ mLen, const float *__restrict__ mBTG,
                                              const float *__restrict__ mBRG, float *__restrict__ mF,
                                              float *__restrict__ mT, const float *__restrict__ bbC,
                                              const float *__restrict__ bbU, const float *__restrict__ bbInv,
                                              const int *__restrict__ hbs, const int *__restrict__ tm,
                                              const float *__restrict__ axial, float *__restrict__ miniData,
                                              MiniParams mp, int nMB, int nBb, int nD) {
    int d = blockIdx.x * T + threadIdx.x;
    if (d >= nD) return;
    tether_one(d, nMB, nBb, hbs[d], 3 * tm[d], axial[d], mC, mU, mLen, mBTG, mBRG, mF, mT, bbC, bbU, bbInv, miniData,
               mp);
}

// attachKey = [node(nA) | motor(nA)]; boundSeg null = NodeSystem.tether, else tetherBoundStiffen
__global__ void __launch_bounds__(T) k_node_tether(const float *__restrict__ mC, const float *__restrict__ mU,
                                                   const float *__restrict__ mLen, const float *__restrict__ mBTG,
                                                   float *__restrict__ mF, float *__restrict__ mT,
                                                   const float *__restrict__ nC, const float *__restrict__ nU,
                                                   const float *__restrict__ nY, const float *__restrict__ nInv,
                                                   const int *__restrict__ attachKey, const float *__restrict__ radial,
                                                   const float *__restrict__ coeffK, const int *__restrict__ boundSeg,
                                                   float *__restrict__ nodeData, double dt, double boundCoeff,
                                                   int nMB, int nN, int nA) {
    int a = blockIdx.x * T + threadIdx.x;
    if (a >= nA) return;
    int m = attachKey[nA + a];
    bool atEnd1;
    double coeff = node_coeff(coeffK[a], m, boundSeg, boundCoeff, atEnd1);
    node_tether_one(a, nMB, nN, nA, attachKey[a], 3 * m, atEnd1, coeff, dt, mC, mU, mLen, mBTG, mF, mT, nC, nU, nY,
                    nInv, radial, nodeData);
}

// NodeSystem.xbridgeInject (NodeSystem.java:141-166)
__global__ void __launch_bounds__(T) k_node_xbridge(const float *__restrict__ bondData, const float *__restrict__ nU,
                                                    const float *__restrict__ nY, const int *__restrict__ attachKey,
                                                    const float *__restrict__ radial,
                                                    const float *__restrict__ coeffK, const int *__restrict__ boundSeg,
                                                    float *__restrict__ nodeData, int nN, int nA) {
    int a = blockIdx.x * T + threadIdx.x;
    if (a >= nA) return;
    int k = attachKey[a], m = attachKey[nA + a];
    bool dimer = coeffK[a] > 0.0f;
    double nux = nU[k], nuy = nU[nN + k], nuz = nU[2 * nN + k];
    double nyx = nY[k], nyy = nY[nN + k], nyz = nY[2 * nN + k];
    double nzx = nuy * nyz - nuz * nyy, nzy = nuz * nyx - nux * nyz, nzz = nux * nyy - nuy * nyx;
    double ru = radial[a], ry = radial[nA + a], rz = radial[2 * nA + a];
    double Rx = ru * nux + ry * nyx + rz * nzx;
    double Ry = ru * nuy + ry * nyy + rz * nzy;
    double Rz = ru * nuz + ry * nyz + rz * nzz;
    double fx = 0, fy = 0, fz = 0;
    if (boundSeg[m] >= 0) {
        const float *b = bondData + (size_t)m * XB_STRIDE;
        fx += b[0]; fy += b[1]; fz += b[2];
    }
    if (dimer && boundSeg[m + 1] >= 0) {
        const float *b = bondData + (size_t)(m + 1) * XB_STRIDE;
        fx += b[0]; fy += b[1]; fz += b[2];
    }
    double Rmx = Rx * 1.0e-6, Rmy = Ry * 1.0e-6, Rmz = Rz * 1.0e-6;
    double Tx = Rmy * fz - Rmz * fy, Ty = Rmz * fx - Rmx * fz, Tz = Rmx * fy - Rmy * fx;
    float *o = nodeData + (size_t)MINI_STRIDE * a;
    o[0] = (float)(o[0] + fx); o[1] = (float)(o[1] + fy); o[2] = (float)(o[2] + fz);
    o[3] = (float)(o[3] + Tx); o[4] = (float)(o[4] + Ty); o[5] = (float)(o[5] + Tz);
}

}  // namespace

#define CHK(c, name) return check_launch(c, name)

extern "C" {

int sbx_publish_head_from_body(sbx_ctx *c, const float *bodyCoord, const float *bodyUVec, const float *bodySegLength,
                               float *head, float *uVec, float *rodUVec, const int *counts, int nM) {
    (void)counts;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nM >= 0, "publish_head_from_body: nM < 0");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, bodyCoord && bodyUVec && bodySegLength && head && uVec && rodUVec, "publish_head: null array");
    hipLaunchKernelGGL(k_publish_head, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, bodyCoord, bodyUVec,
                       bodySegLength, head, uVec, rodUVec, nM);
    CHK(c, "k_publish_head");
}

int sbx_publish_motors(sbx_ctx *c, const float *head, const float *reach, float *center, float *boundingRadius,
                       int *ownerStore, int *ownerSlot, const int *publishParams, const int *counts, int nM,
                       int cap) {
    (void)counts;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, publishParams && nM >= 0, "publish_motors: null publishParams or nM < 0");
    int base = publishParams[0];
    SBX_REQUIRE(c, base >= 0 && base + nM <= cap, "publish_motors: slots [%d, %d) outside capacity %d", base,
                base + nM, cap);
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, head && reach && center && boundingRadius, "publish_motors: null array");
    hipLaunchKernelGGL(k_publish_motors, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, head, reach, center,
                       boundingRadius, ownerStore, ownerSlot, nM, cap, base);
    CHK(c, "k_publish_motors");
}

int sbx_grid_reachable(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                       const float *segEnd2, const float *gridParams, const int *gridDims,
                       const int *gridCellOffsets, const int *gridCellContents, const int *ownerStore,
                       const int *ownerSlot, int *motorReachSeg, int *motorReachCount, const float *kinParams,
                       const int *counts, int nSeg, int maxCand) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, gridParams && gridDims && kinParams && counts && maxCand > 0 && nSeg >= 0,
                "grid_reachable: bad params");
    int nM = counts[0];
    SBX_REQUIRE(c, nM >= 0, "grid_reachable: nM < 0");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, head && uVec && rodUVec && segEnd1 && segEnd2 && gridCellOffsets && gridCellContents &&
                       ownerStore && ownerSlot && motorReachSeg && motorReachCount,
                "grid_reachable: null array");
    ReachGrid g = {gridParams[0], gridParams[1], gridParams[2], gridParams[4], gridDims[0], gridDims[1], gridDims[2]};
    hipLaunchKernelGGL(k_grid_reachable, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, head, uVec, rodUVec, segEnd1,
                       segEnd2, g, gridCellOffsets, gridCellContents, ownerStore, ownerSlot, motorReachSeg,
                       motorReachCount, maxCand, kinParams[7], kinParams[8], nM, nSeg);
    CHK(c, "k_grid_reachable");
}

int sbx_bind_nearest(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                     const float *segEnd2, const int *motorCandSeg, const int *motorCandCount, int *boundSeg,
                     float *bindArc, const float *kinParams, const int *counts, int nSeg, int maxCand) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, kinParams && counts && maxCand > 0 && nSeg >= 0, "bind_nearest: bad params");
    int nM = counts[0];
    if (nM <= 0) return SBX_OK;
    SBX_REQUIRE(c, head && uVec && rodUVec && segEnd1 && segEnd2 && motorCandSeg && motorCandCount && boundSeg &&
                       bindArc,
                "bind_nearest: null array");
    hipLaunchKernelGGL(k_bind_nearest, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, head, uVec, rodUVec, segEnd1,
                       segEnd2, motorCandSeg, motorCandCount, maxCand, boundSeg, bindArc, kinParams[7], kinParams[8],
                       nM, nSeg, (const int *)nullptr);
    CHK(c, "k_bind_nearest");
}

int sbx_bind_nearest_node_aware(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec,
                                const float *segEnd1, const float *segEnd2, const int *motorCandSeg,
                                const int *motorCandCount, int *boundSeg, float *bindArc, const int *seedNode,
                                const float *kinParams, const int *counts, int nSeg, int maxCand) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, kinParams && counts && maxCand > 0 && nSeg >= 0, "bind_nearest_node_aware: bad params");
    int nM = counts[0];
    if (nM <= 0) return SBX_OK;
    SBX_REQUIRE(c, head && uVec && rodUVec && segEnd1 && segEnd2 && motorCandSeg && motorCandCount && boundSeg &&
                       bindArc && seedNode,
                "bind_nearest_node_aware: null array");
    hipLaunchKernelGGL(k_bind_nearest, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, head, uVec, rodUVec, segEnd1,
                       segEnd2, motorCandSeg, motorCandCount, maxCand, boundSeg, bindArc, kinParams[7], kinParams[8],
                       nM, nSeg, seedNode);
    CHK(c, "k_bind_nearest(node aware)");
}

static int brute_launch(sbx_ctx *c, const char *what, const float *head, const float *uVec, const float *rodUVec,
                        const float *segEnd1, const float *segEnd2, int *reachSeg, int *reachCount,
                        const int *seedNode, const float *kinParams, const int *counts, int nSeg, int maxCand) {
    SBX_REQUIRE(c, kinParams && counts && maxCand > 0 && nSeg >= 0, "brute_reachable: bad params");
    int nM = counts[0];
    if (nM <= 0) return SBX_OK;
    SBX_REQUIRE(c, head && uVec && rodUVec && reachSeg && reachCount && (nSeg == 0 || (segEnd1 && segEnd2)),
                "brute_reachable: null array");
    hipLaunchKernelGGL(k_brute_reachable, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, head, uVec, rodUVec,
                       segEnd1, segEnd2, reachSeg, reachCount, seedNode, kinParams[7], kinParams[8], nM, nSeg,
                       maxCand);
    CHK(c, what);
}

int sbx_brute_reachable(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                        const float *segEnd2, int *reachSeg, int *reachCount, const float *kinParams,
                        const int *counts, int nSeg, int maxCand) {
    SBX_CHECK_CTX(c);
    return brute_launch(c, "k_brute_reachable", head, uVec, rodUVec, segEnd1, segEnd2, reachSeg, reachCount, nullptr,
                        kinParams, counts, nSeg, maxCand);
}

int sbx_brute_reachable_node_aware(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec,
                                   const float *segEnd1, const float *segEnd2, int *reachSeg, int *reachCount,
                                   const int *seedNode, const float *kinParams, const int *counts, int nSeg,
                                   int maxCand) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, seedNode || nSeg == 0, "brute_reachable_node_aware: null seedNode");
    return brute_launch(c, "k_brute_reachable(node aware)", head, uVec, rodUVec, segEnd1, segEnd2, reachSeg,
                        reachCount, seedNode, kinParams, counts, nSeg, maxCand);
}

int sbx_catch_slip_release(sbx_ctx *c, int *boundSeg, const float *forceDotFil, const float *forceMag,
                           int *cooldown, int *stats, int *capStats, const float *kinParams, const int *counts,
                           int nM, int rng, uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, kinParams && counts && nM >= 0, "catch_slip_release: bad params");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "catch_slip_release: unknown rng");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, boundSeg && forceDotFil && forceMag && cooldown && stats && capStats, "catch_slip: null array");
    hipLaunchKernelGGL(k_catch_slip, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, boundSeg, forceDotFil, forceMag,
                       cooldown, stats, capStats, kin_params(kinParams), counts[1], counts[2], rng, stream, nM);
    CHK(c, "k_catch_slip");
}

int sbx_nucleotide_cycle(sbx_ctx *c, int *nucleotideState, const int *boundSeg, const float *forceDotHist,
                         const float *nucParams, const int *counts, int nM, int rng, uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nucParams && counts && nM >= 0, "nucleotide_cycle: bad params");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "nucleotide_cycle: unknown rng");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, nucleotideState && boundSeg && forceDotHist, "nucleotide_cycle: null array");
    hipLaunchKernelGGL(k_nucleotide, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, nucleotideState, boundSeg,
                       forceDotHist, nuc_params(nucParams), counts[1], counts[2], rng, stream, nM);
    CHK(c, "k_nucleotide");
}

int sbx_bond_forces(sbx_ctx *c, const float *motorCoord, const float *motorUVec, const float *motorYVec,
                    const float *motorBRotGam, const float *filCoord, const float *filUVec, const float *filYVec,
                    const float *filBRotGam, const float *filSegLength, const int *boundSeg, const float *bindArc,
                    const int *nucleotideState, float *bondData, const float *xbParams, int nM, int nSeg) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, xbParams && nM >= 0 && nSeg >= 0, "bond_forces: bad params");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, xbParams[3] != 0.0f, "bond_forces: dt must be non-zero");
    SBX_REQUIRE(c, motorCoord && motorUVec && motorYVec && motorBRotGam && filCoord && filUVec && filYVec &&
                       filBRotGam && filSegLength && boundSeg && bindArc && nucleotideState && bondData,
                "bond_forces: null array");
    hipLaunchKernelGGL(k_bond, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, motorCoord, motorUVec, motorYVec,
                       motorBRotGam, filCoord, filUVec, filYVec, filBRotGam, filSegLength, boundSeg, bindArc,
                       nucleotideState, bondData, xb_params(xbParams), nM, nSeg);
    CHK(c, "k_bond");
}

int sbx_apply_head_force(sbx_ctx *c, const float *bondData, float *bodyForceSum, float *bodyTorqueSum, int nM) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nM >= 0, "apply_head_force: nM < 0");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, bondData && bodyForceSum && bodyTorqueSum, "apply_head_force: null array");
    hipLaunchKernelGGL(k_apply_head, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, bondData, bodyForceSum,
                       bodyTorqueSum, nM);
    CHK(c, "k_apply_head");
}

int sbx_register_force_dot(sbx_ctx *c, const float *bondData, const int *boundSeg, float *forceDotFil,
                           float *forceMag, float *forceDotHist, int *forceDotPlace, int nM) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nM >= 0, "register_force_dot: nM < 0");
    if (nM == 0) return SBX_OK;
    SBX_REQUIRE(c, bondData && boundSeg && forceDotFil && forceMag && forceDotHist && forceDotPlace,
                "register_force_dot: null array");
    hipLaunchKernelGGL(k_register_force_dot, dim3(blocks_for(nM, T)), dim3(T), 0, c->stream, bondData, boundSeg,
                       forceDotFil, forceMag, forceDotHist, forceDotPlace, nM);
    CHK(c, "k_register_force_dot");
}

int sbx_motor_seg_gather(sbx_ctx *c, const int *segMotorOffsets, const int *segMotorMyo, const float *bondData,
                         float *filForceSum, float *filTorqueSum, int nSeg) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSeg >= 0, "motor_seg_gather: nSeg < 0");
    if (nSeg == 0) return SBX_OK;
    SBX_REQUIRE(c, segMotorOffsets && segMotorMyo && bondData && filForceSum && filTorqueSum,
                "motor_seg_gather: null array");
    hipLaunchKernelGGL(k_csr_gather, dim3(blocks_for(nSeg, T)), dim3(T), 0, c->stream, segMotorOffsets, segMotorMyo,
                       bondData, XB_STRIDE, 6, filForceSum, filTorqueSum, nSeg);
    CHK(c, "k_csr_gather(motor)");
}

int sbx_motor_joints(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const int *nucleotideState, const float *jointParams, int nB) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, jointParams && nB >= 0 && nB % 3 == 0, "motor_joints: need nB = 3*nMotors");
    if (nB == 0) return SBX_OK;
    SBX_REQUIRE(c, jointParams[0] != 0.0f, "motor_joints: dt must be non-zero");
    SBX_REQUIRE(c, coord && uVec && segLength && bTransGam && bRotGam && forceSum && torqueSum && nucleotideState,
                "motor_joints: null array");
    hipLaunchKernelGGL(k_joints, dim3(blocks_for(nB, T)), dim3(T), 0, c->stream, coord, uVec, segLength, bTransGam,
                       bRotGam, forceSum, torqueSum, nucleotideState, joint_params(jointParams), nB);
    CHK(c, "k_joints");
}

int sbx_dimer_couple(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const int *motorA, const int *motorB, const int *parallel, const float *dimerParams,
                     const int *boundSeg, int nB, int nD) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dimerParams && nB >= 0 && nD >= 0, "dimer_couple: bad params");
    if (nD == 0) return SBX_OK;
    SBX_REQUIRE(c, dimerParams[0] != 0.0f, "dimer_couple: dt must be non-zero");
    SBX_REQUIRE(c, coord && uVec && segLength && bTransGam && bRotGam && forceSum && torqueSum && motorA && motorB &&
                       parallel && boundSeg,
                "dimer_couple: null array");
    hipLaunchKernelGGL(k_dimer, dim3(blocks_for(nD, T)), dim3(T), 0, c->stream, coord, uVec, segLength, bTransGam,
                       bRotGam, forceSum, torqueSum, motorA, motorB, parallel, boundSeg, dimer_params(dimerParams), nB,
                       nD);
    CHK(c, "k_dimer");
}

int sbx_mini_tether(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                    const float *mBTransGam, const float *mBRotGam, float *mForceSum, float *mTorqueSum,
                    const float *bbCoord, const float *bbUVec, const float *bbInvDragY, const int *headBackboneSlot,
                    const int *motorA, const float *attachAxial, float *miniData, const float *miniParams, int nMB,
                    int nBb, int nD) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, miniParams && nMB >= 0 && nBb >= 0 && nD >= 0, "mini_tether: bad params");
    if (nD == 0) return SBX_OK;
    SBX_REQUIRE(c, miniParams[0] != 0.0f, "mini_tether: dt must be non-zero");
    SBX_REQUIRE(c, mCoord && mUVec && mSegLength && mBTransGam && mBRotGam && mForceSum && mTorqueSum && bbCoord &&
                       bbUVec && bbInvDragY && headBackboneSlot && motorA && attachAxial && miniData,
                "mini_tether: null array");
    hipLaunchKernelGGL(k_tether, dim3(blocks_for(nD, T)), dim3(T), 0, c->stream, mCoord, mUVec, mSegLength,
                       mBTransGam, mBRotGam, mForceSum, mTorqueSum, bbCoord, bbUVec, bbInvDragY, headBackboneSlot,
                       motorA, attachAxial, miniData, mini_params(miniParams), nMB, nBb, nD);
    CHK(c, "k_tether");
}

static int node_tether_launch(sbx_ctx *c, const char *what, const float *mCoord, const float *mUVec,
                              const float *mSegLength, const float *mBTransGam, float *mForceSum, float *mTorqueSum,
                              const float *nCoord, const float *nUVec, const float *nYVec, const float *nInvTransY,
                              const int *attachKey, const float *radial, const float *attachCoeffK,
                              const int *boundSeg, float *nodeData, const float *nodeParams, int nMB, int nN, int nA) {
    SBX_REQUIRE(c, nodeParams && nMB >= 0 && nN >= 0 && nA >= 0, "node_tether: bad params");
    if (nA == 0) return SBX_OK;
    SBX_REQUIRE(c, nodeParams[0] != 0.0f, "node_tether: dt must be non-zero");
    SBX_REQUIRE(c, nN > 0 && nMB > 0, "node_tether: attachments without nodes or motors");
    SBX_REQUIRE(c, mCoord && mUVec && mSegLength && mBTransGam && mForceSum && mTorqueSum && nCoord && nUVec &&
                       nYVec && nInvTransY && attachKey && radial && attachCoeffK && nodeData,
                "node_tether: null array");
    hipLaunchKernelGGL(k_node_tether, dim3(blocks_for(nA, T)), dim3(T), 0, c->stream, mCoord, mUVec, mSegLength,
                       mBTransGam, mForceSum, mTorqueSum, nCoord, nUVec, nYVec, nInvTransY, attachKey, radial,
                       attachCoeffK, boundSeg, nodeData, (double)nodeParams[0],
                       boundSeg ? (double)nodeParams[1] : 0.0, nMB, nN, nA);
    CHK(c, what);
}

int sbx_node_tether(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                    const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                    const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachKey,
                    const float *radial, const float *attachCoeffK, float *nodeData, const float *nodeParams, int nMB,
                    int nN, int nA) {
    SBX_CHECK_CTX(c);
    return node_tether_launch(c, "k_node_tether", mCoord, mUVec, mSegLength, mBTransGam, mForceSum, mTorqueSum,
                              nCoord, nUVec, nYVec, nInvTransY, attachKey, radial, attachCoeffK, nullptr, nodeData,
                              nodeParams, nMB, nN, nA);
}

int sbx_node_tether_bound_stiffen(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                                  const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                                  const float *nUVec, const float *nYVec, const float *nInvTransY,
                                  const int *attachKey, const float *radial, const float *attachCoeffK,
                                  const int *boundSeg, float *nodeData, const float *nodeParams, int nMB, int nN,
                                  int nA) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, boundSeg || nA == 0, "node_tether_bound_stiffen: null boundSeg");
    return node_tether_launch(c, "k_node_tether(bound stiffen)", mCoord, mUVec, mSegLength, mBTransGam, mForceSum,
                              mTorqueSum, nCoord, nUVec, nYVec, nInvTransY, attachKey, radial, attachCoeffK,
                              boundSeg, nodeData, nodeParams, nMB, nN, nA);
}

int sbx_node_xbridge_inject(sbx_ctx *c, const float *bondData, const float *nUVec, const float *nYVec,
                            const int *attachKey, const float *radial, const float *attachCoeffK,
                            const int *boundSeg, float *nodeData, int nN, int nA) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nN >= 0 && nA >= 0, "node_xbridge_inject: negative counts");
    if (nA == 0) return SBX_OK;
    SBX_REQUIRE(c, nN > 0, "node_xbridge_inject: attachments without nodes");
    SBX_REQUIRE(c, bondData && nUVec && nYVec && attachKey && radial && attachCoeffK && boundSeg && nodeData,
                "node_xbridge_inject: null array");
    hipLaunchKernelGGL(k_node_xbridge, dim3(blocks_for(nA, T)), dim3(T), 0, c->stream, bondData, nUVec, nYVec,
                       attachKey, radial, attachCoeffK, boundSeg, nodeData, nN, nA);
    CHK(c, "k_node_xbridge");
}

int sbx_backbone_gather(sbx_ctx *c, const int *bbDimerOffsets, const int *bbDimerList, const float *miniData,
                        float *bbForceSum, float *bbTorqueSum, int nBb) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nBb >= 0, "backbone_gather: nBb < 0");
    if (nBb == 0) return SBX_OK;
    SBX_REQUIRE(c, bbDimerOffsets && bbDimerList && miniData && bbForceSum && bbTorqueSum,
                "backbone_gather: null array");
    hipLaunchKernelGGL(k_csr_gather, dim3(blocks_for(nBb, T)), dim3(T), 0, c->stream, bbDimerOffsets, bbDimerList,
                       miniData, MINI_STRIDE, 0, bbForceSum, bbTorqueSum, nBb);
    CHK(c, "k_csr_gather(backbone)");
}

}  // extern "C"
