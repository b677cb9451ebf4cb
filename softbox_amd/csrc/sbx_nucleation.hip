// sbx_nucleation.hip -- node nucleation on the GPU (gfx950): the per-node seed requests, the node bond of each
// born slot, the seed tether with its node reaction and the bond's dissolution (NodeNucleationSystem.java).
//
// Every rule is a per-node or per-slot self-write, as in the reference.  The two many-to-one reductions keep
// the reference's order: countBoundFil is an integer histogram (atomic adds, exact in any order) and
// seedTetherNodeReact sums each node's held seeds in slot order, one wavefront per node stepping through the
// slots 64 at a time with a ballot of the slots it holds.
#include "sbx_device.h"
#include "sbx_internal.h"

namespace sbx {
namespace {

constexpr uint32_t SALT_NUC1 = 0x4E554331u;  // "NUC1" (NodeNucleationSystem.java:76)
constexpr uint32_t SALT_NUCD = 0x4E554344u;  // "NUCD" (NodeNucleationSystem.java:253)
constexpr int T = 256;

// four words of one event draw (WANG: base ^ salt and its three xor variants; PHILOX: one counter)
__device__ __forceinline__ void event_bits4(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream,
                                            uint32_t salt, uint32_t h[4]) {
    if (rng == RNG_WANG) {
        uint32_t base = wang_base(item, step, seed) ^ salt;
        h[0] = wang_hash(base);
        h[1] = wang_hash(base ^ 0x9e3779b9u);
        h[2] = wang_hash(base ^ 0x85ebca6bu);
        h[3] = wang_hash(base ^ 0xc2b2ae35u);
        return;
    }
    philox4x32_10(item, (uint32_t)step, salt, 2u, (uint32_t)seed, stream, h);
}

__global__ void __launch_bounds__(T) k_nuc_zero(int *a, int n) {
    int i = blockIdx.x * T + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void __launch_bounds__(T) k_nuc_count(const int *__restrict__ seedNode, int *__restrict__ nodeBoundFil,
                                                  int C, int nN) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= C) return;
    int k = seedNode[s];
    if (k >= 0 && k < nN) atomicAdd(nodeBoundFil + k, 1);
}

// NodeNucleationSystem.emit body (NodeNucleationSystem.java:74-99), float arithmetic as the reference's
__global__ void __launch_bounds__(T) k_nuc_emit(const float *__restrict__ nodeCoord,
                                                 const int *__restrict__ nodeBoundFil, int *__restrict__ acceptFlag,
                                                 float *__restrict__ reqCoord, float *__restrict__ reqUVec,
                                                 float *__restrict__ reqYVec, float pNuc, float halfLen, int formins,
                                                 int step, int seed, int poolOK, int nN, int RC, int rng,
                                                 uint32_t stream) {
    int k = blockIdx.x * T + threadIdx.x;
    if (k >= nN) return;
    acceptFlag[k] = 0;
    if (nodeBoundFil[k] >= formins) return;
    if (poolOK == 0) return;
    uint32_t h[4];
    event_bits4(rng, (uint32_t)k, step, seed, stream, SALT_NUC1, h);
    float u = uniform01(h[0]);
    if (u >= pNuc) return;
    float dx = uniform01(h[1]) * 2.0f - 1.0f;
    float dy = uniform01(h[2]) * 2.0f - 1.0f;
    float dz = uniform01(h[3]) * 2.0f - 1.0f;
    float m2 = dx * dx + dy * dy + dz * dz;
    if (m2 < 1.0e-12f) {
        dx = 1.0f; dy = 0.0f; dz = 0.0f; m2 = 1.0f;
    }
    float inv = 1.0f / (float)sqrt((double)m2);
    dx *= inv; dy *= inv; dz *= inv;
    float ncx = nodeCoord[k], ncy = nodeCoord[nN + k], ncz = nodeCoord[2 * nN + k];
    float ex, ey, ez;
    if (dx < 0.9f && dx > -0.9f) {
        ex = 1.0f; ey = 0.0f; ez = 0.0f;
    } else {
        ex = 0.0f; ey = 1.0f; ez = 0.0f;
    }
    float yx = dy * ez - dz * ey, yy = dz * ex - dx * ez, yz = dx * ey - dy * ex;
    float ym = 1.0f / (float)sqrt((double)(yx * yx + yy * yy + yz * yz));
    yx *= ym; yy *= ym; yz *= ym;
    reqCoord[k] = ncx + halfLen * dx;
    reqCoord[RC + k] = ncy + halfLen * dy;
    reqCoord[2 * RC + k] = ncz + halfLen * dz;
    reqUVec[k] = -dx; reqUVec[RC + k] = -dy; reqUVec[2 * RC + k] = -dz;
    reqYVec[k] = yx; reqYVec[RC + k] = yy; reqYVec[2 * RC + k] = yz;
    acceptFlag[k] = 1;
}

// tagSeeds (mode 0) / initNewborn (mode 1): the born slot of each accepted request (NodeNucleationSystem.java:
// 104-115, 228-244)
__global__ void __launch_bounds__(T) k_nuc_born(const int *__restrict__ rankOffsets, const int *__restrict__ freeList,
                                                 const int *__restrict__ freeOffsets, int *__restrict__ seedNode,
                                                 int *__restrict__ monomerCount, float *__restrict__ segLength,
                                                 int mon, float len, int C, int K) {
    int r = blockIdx.x * T + threadIdx.x;
    if (r >= K) return;
    int rank = rankOffsets[r];
    if (!(rankOffsets[r + 1] > rank)) return;
    if (rank >= freeOffsets[C]) return;
    int slot = freeList[rank];
    if (seedNode) seedNode[slot] = r;
    if (monomerCount) {
        monomerCount[slot] = mon;
        segLength[slot] = len;
    }
}

// the seed tether force of slot s held by node k (NodeNucleationSystem.java:125-137)
__device__ __forceinline__ bool seed_force(int s, int k, int C, int nN, const float *__restrict__ fc,
                                           const float *__restrict__ fu, const float *__restrict__ fl,
                                           const float *__restrict__ fg, const float *__restrict__ nc,
                                           const float *__restrict__ ninv, double fracMove, double dt, double &dx,
                                           double &dy, double &dz, double &inv, double &fm, double &half,
                                           double u[3]) {
    double cx = fc[s], cy = fc[C + s], cz = fc[2 * C + s];
    u[0] = fu[s]; u[1] = fu[C + s]; u[2] = fu[2 * C + s];
    half = 0.5 * fl[s];
    double e2x = cx + half * u[0], e2y = cy + half * u[1], e2z = cz + half * u[2];
    dx = nc[k] - e2x; dy = nc[nN + k] - e2y; dz = nc[2 * nN + k] - e2z;
    double strain = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(strain > 0.0)) return false;
    inv = 1.0 / strain;
    double denom = dt * (1.0 / fg[s] + ninv[k]);
    fm = (denom > 0.0) ? (fracMove * 1.0e-6 * strain / denom) : 0.0;
    return true;
}

__global__ void __launch_bounds__(T) k_nuc_seed_tether(const float *__restrict__ fc, const float *__restrict__ fu,
                                                        const float *__restrict__ fl, const float *__restrict__ fg,
                                                        float *__restrict__ fF, float *__restrict__ fT,
                                                        const float *__restrict__ nc, const float *__restrict__ ninv,
                                                        const int *__restrict__ seedNode, double fracMove, double dt,
                                                        int C, int nN) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= C) return;
    int k = seedNode[s];
    if (k < 0) return;
    double dx, dy, dz, inv, fm, half, u[3];
    if (!seed_force(s, k, C, nN, fc, fu, fl, fg, nc, ninv, fracMove, dt, dx, dy, dz, inv, fm, half, u)) return;
    double Fx = fm * (dx * inv), Fy = fm * (dy * inv), Fz = fm * (dz * inv);
    fF[s] = (float)(fF[s] + Fx);
    fF[C + s] = (float)(fF[C + s] + Fy);
    fF[2 * C + s] = (float)(fF[2 * C + s] + Fz);
    double Rx = half * u[0] * 1.0e-6, Ry = half * u[1] * 1.0e-6, Rz = half * u[2] * 1.0e-6;
    fT[s] = (float)(fT[s] + (Ry * Fz - Rz * Fy));
    fT[C + s] = (float)(fT[C + s] + (Rz * Fx - Rx * Fz));
    fT[2 * C + s] = (float)(fT[2 * C + s] + (Rx * Fy - Ry * Fx));
}

// seedTetherNodeReact: one wavefront per node sweeps the slots 64 at a time; the held ones (a ballot) are added
// by lane 0 in slot order, as the reference's per-node loop over every slot does
__global__ void __launch_bounds__(T) k_nuc_node_react(const float *__restrict__ fc, const float *__restrict__ fu,
                                                       const float *__restrict__ fl, const float *__restrict__ fg,
                                                       float *__restrict__ nF, const float *__restrict__ nc,
                                                       const float *__restrict__ ninv, const int *__restrict__ seedNode,
                                                       double fracMove, double dt, int C, int nN) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (T / 64) + (threadIdx.x >> 6);
    if (k >= nN) return;   // whole wavefronts exit together
    double rx = 0.0, ry = 0.0, rz = 0.0;
    for (int base = 0; base < C; base += 64) {
        int s = base + lane;
        bool held = s < C && seedNode[s] == k;
        unsigned long long m = __ballot(held);
        if (!m) continue;
        double cx = 0, cy = 0, cz = 0;
        if (held) {
            double dx, dy, dz, inv, fm, half, u[3];
            if (seed_force(s, k, C, nN, fc, fu, fl, fg, nc, ninv, fracMove, dt, dx, dy, dz, inv, fm, half, u)) {
                cx = fm * dx * inv; cy = fm * dy * inv; cz = fm * dz * inv;
            } else {
                cx = cy = cz = 0.0;
            }
        }
        while (m) {
            int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            double ax = __shfl(cx, l), ay = __shfl(cy, l), az = __shfl(cz, l);
            rx -= ax; ry -= ay; rz -= az;
        }
    }
    if (lane == 0) {
        nF[k] = (float)(nF[k] + rx);
        nF[nN + k] = (float)(nF[nN + k] + ry);
        nF[2 * nN + k] = (float)(nF[2 * nN + k] + rz);
    }
}

__global__ void __launch_bounds__(T) k_nuc_aim(const float *__restrict__ fu, const float *__restrict__ fRG,
                                                float *__restrict__ fT, const float *__restrict__ aim, double aimCoeff,
                                                double dt, int C) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= C) return;
    double ax = aim[s], ay = aim[C + s], az = aim[2 * C + s];
    if (ax * ax + ay * ay + az * az < 0.25) return;
    double ux = fu[s], uy = fu[C + s], uz = fu[2 * C + s];
    double cx = uy * az - uz * ay, cy = uz * ax - ux * az, cz = ux * ay - uy * ax;
    double kk = aimCoeff * fRG[C + s] / dt;
    fT[s] = (float)(fT[s] + kk * cx);
    fT[C + s] = (float)(fT[C + s] + kk * cy);
    fT[2 * C + s] = (float)(fT[2 * C + s] + kk * cz);
}

__global__ void __launch_bounds__(T) k_nuc_dissolve(int *__restrict__ seedNode, float pDetach, int step, int seed,
                                                     int C, int rng, uint32_t stream) {
    int s = blockIdx.x * T + threadIdx.x;
    if (s >= C) return;
    if (seedNode[s] < 0) return;
    float u = uniform01(event_bits(rng, (uint32_t)s, step, seed, stream, SALT_NUCD));
    if (u < pDetach) seedNode[s] = -1;
}

}  // namespace
}  // namespace sbx

using namespace sbx;

#define NUC_RNG_OK(c, rng, what) \
    SBX_REQUIRE(c, (rng) == SBX_RNG_WANG || (rng) == SBX_RNG_PHILOX, what ": unknown rng %d", rng)

extern "C" {

int sbx_nuc_count_bound_fil(sbx_ctx *c, const int *seedNode, int *nodeBoundFil, int C, int nN) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && nN >= 0, "nuc_count_bound_fil: negative counts");
    SBX_REQUIRE(c, (C == 0 || seedNode) && (nN == 0 || nodeBoundFil), "nuc_count_bound_fil: null array");
    if (nN > 0) hipLaunchKernelGGL(k_nuc_zero, dim3(blocks_for(nN, T)), dim3(T), 0, c->stream, nodeBoundFil, nN);
    if (C > 0 && nN > 0)
        hipLaunchKernelGGL(k_nuc_count, dim3(blocks_for(C, T)), dim3(T), 0, c->stream, seedNode, nodeBoundFil, C, nN);
    return check_launch(c, "k_nuc_count");
}

int sbx_nuc_emit(sbx_ctx *c, const float *nodeCoord, const int *nodeBoundFil, int *acceptFlag, float *reqCoord,
                 float *reqUVec, float *reqYVec, const float *nucParams, int step, int seed, int poolOK, int nN, int RC,
                 int rng, uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nucParams && nN >= 0 && RC >= nN, "nuc_emit: bad arguments (RC must cover nN requests)");
    NUC_RNG_OK(c, rng, "nuc_emit");
    if (nN == 0) return SBX_OK;
    SBX_REQUIRE(c, nodeCoord && nodeBoundFil && acceptFlag && reqCoord && reqUVec && reqYVec, "nuc_emit: null array");
    hipLaunchKernelGGL(k_nuc_emit, dim3(blocks_for(nN, T)), dim3(T), 0, c->stream, nodeCoord, nodeBoundFil,
                       acceptFlag, reqCoord, reqUVec, reqYVec, nucParams[0], nucParams[1], (int)nucParams[2], step, seed,
                       poolOK, nN, RC, rng, stream);
    return check_launch(c, "k_nuc_emit");
}

int sbx_nuc_tag_seeds(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets, int *seedNode,
                      int C, int K) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0, "nuc_tag_seeds: negative counts");
    if (K == 0) return SBX_OK;
    SBX_REQUIRE(c, rankOffsets && freeList && freeOffsets && seedNode, "nuc_tag_seeds: null array");
    hipLaunchKernelGGL(k_nuc_born, dim3(blocks_for(K, T)), dim3(T), 0, c->stream, rankOffsets, freeList, freeOffsets,
                       seedNode, (int *)nullptr, (float *)nullptr, 0, 0.0f, C, K);
    return check_launch(c, "k_nuc_born(tag)");
}

int sbx_nuc_init_newborn(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                         int *monomerCount, float *segLength, const float *seedParams, int C, int K) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && K >= 0 && seedParams, "nuc_init_newborn: bad arguments");
    if (K == 0) return SBX_OK;
    SBX_REQUIRE(c, rankOffsets && freeList && freeOffsets && monomerCount && segLength,
                "nuc_init_newborn: null array");
    hipLaunchKernelGGL(k_nuc_born, dim3(blocks_for(K, T)), dim3(T), 0, c->stream, rankOffsets, freeList, freeOffsets,
                       (int *)nullptr, monomerCount, segLength, (int)seedParams[0], seedParams[1], C, K);
    return check_launch(c, "k_nuc_born(newborn)");
}

int sbx_nuc_seed_tether(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filSegLength,
                        const float *filBTransGam, float *filForceSum, float *filTorqueSum, const float *nodeCoord,
                        const float *nodeInvTransX, const int *seedNode, const float *tetherParams, int C, int nN) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && nN >= 0 && tetherParams, "nuc_seed_tether: bad arguments");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, tetherParams[1] != 0.0f, "nuc_seed_tether: dt must be non-zero");
    SBX_REQUIRE(c, filCoord && filUVec && filSegLength && filBTransGam && filForceSum && filTorqueSum && seedNode &&
                       (nN == 0 || (nodeCoord && nodeInvTransX)),
                "nuc_seed_tether: null array");
    hipLaunchKernelGGL(k_nuc_seed_tether, dim3(blocks_for(C, T)), dim3(T), 0, c->stream, filCoord, filUVec,
                       filSegLength, filBTransGam, filForceSum, filTorqueSum, nodeCoord, nodeInvTransX, seedNode,
                       (double)tetherParams[0], (double)tetherParams[1], C, nN);
    return check_launch(c, "k_nuc_seed_tether");
}

int sbx_nuc_seed_tether_node_react(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filSegLength,
                                   const float *filBTransGam, float *nodeForceSum, const float *nodeCoord,
                                   const float *nodeInvTransX, const int *seedNode, const float *tetherParams, int C,
                                   int nN) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && nN >= 0 && tetherParams, "nuc_seed_tether_node_react: bad arguments");
    if (nN == 0) return SBX_OK;
    SBX_REQUIRE(c, tetherParams[1] != 0.0f, "nuc_seed_tether_node_react: dt must be non-zero");
    SBX_REQUIRE(c, nodeForceSum && nodeCoord && nodeInvTransX &&
                       (C == 0 || (filCoord && filUVec && filSegLength && filBTransGam && seedNode)),
                "nuc_seed_tether_node_react: null array");
    hipLaunchKernelGGL(k_nuc_node_react, dim3(blocks_for(nN, T / 64)), dim3(T), 0, c->stream, filCoord, filUVec,
                       filSegLength, filBTransGam, nodeForceSum, nodeCoord, nodeInvTransX, seedNode,
                       (double)tetherParams[0], (double)tetherParams[1], C, nN);
    return check_launch(c, "k_nuc_node_react");
}

int sbx_nuc_seed_aim_torque(sbx_ctx *c, const float *filUVec, const float *filBRotGam, float *filTorqueSum,
                            const float *aimDir, const float *aimParams, int C) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && aimParams, "nuc_seed_aim_torque: bad arguments");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, aimParams[1] != 0.0f, "nuc_seed_aim_torque: dt must be non-zero");
    SBX_REQUIRE(c, filUVec && filBRotGam && filTorqueSum && aimDir, "nuc_seed_aim_torque: null array");
    hipLaunchKernelGGL(k_nuc_aim, dim3(blocks_for(C, T)), dim3(T), 0, c->stream, filUVec, filBRotGam, filTorqueSum,
                       aimDir, (double)aimParams[0], (double)aimParams[1], C);
    return check_launch(c, "k_nuc_aim");
}

int sbx_nuc_dissolve(sbx_ctx *c, int *seedNode, const float *dissolveParams, int step, int seed, int C, int rng,
                     uint32_t stream) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && dissolveParams, "nuc_dissolve: bad arguments");
    NUC_RNG_OK(c, rng, "nuc_dissolve");
    if (C == 0) return SBX_OK;
    SBX_REQUIRE(c, seedNode, "nuc_dissolve: null array");
    hipLaunchKernelGGL(k_nuc_dissolve, dim3(blocks_for(C, T)), dim3(T), 0, c->stream, seedNode, dissolveParams[0], step,
                       seed, C, rng, stream);
    return check_launch(c, "k_nuc_dissolve");
}

}  // extern "C"
