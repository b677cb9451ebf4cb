/* CPU oracle -- TEST INFRASTRUCTURE ONLY (imported by tests/, smoke() and bench.py's cpu_baseline leg).
 *
 * Protein-node shell: the radial surface tether of motor rods to node spheres (NodeSystem.java) in its
 * plain form and the two scheme variants the reference keeps beside it (bound-stiffened tether,
 * cross-bridge load injection).  The node side of each attachment is written to nodeData[6a..] and
 * gathered per node by the CSR gather of the minifilament backbone (MiniFilamentSystem.backboneGather),
 * keyed by attachNode, exactly as the reference does.
 *
 * Pins: tests/test_oracle_node.py -- the closed-form spring (magnitude coeff 1e-6 strain / (dt (1/g_rod +
 * 1/g_node)), direction toward the surface point), the surface point co-rotating with the node frame,
 * Newton's third law per attachment and per node, singlet vs dimer torque arms, the gather against a
 * brute per-attachment sum and full-system momentum under cross-bridge load (ProteinNodeHarness gates
 * #1a/#1b), the bound-stiffened coefficient switch and the injected cross-bridge load with its torque.
 */
#include <math.h>
#include <stddef.h>

#include "softbox_oracle.h"

#define NODE_STRIDE 6

/* one attachment of NodeSystem.tether / tetherBoundStiffen (NodeSystem.java:66-126, :185-230) */
static void node_tether_one(int a, int k, int rod, int atEnd1, double coeff, int nMB, int nN, int nA, const float *mCoord,
                            const float *mUVec, const float *mSegLength, const float *mBTransGam, float *mForceSum,
                            float *mTorqueSum, const float *nCoord, const float *nUVec, const float *nYVec,
                            const float *nInvTransY, const float *radial, float *nodeData, double dt) {
    double ncx = nCoord[k], ncy = nCoord[nN + k], ncz = nCoord[2 * nN + k];
    double nux = nUVec[k], nuy = nUVec[nN + k], nuz = nUVec[2 * nN + k];
    double nyx = nYVec[k], nyy = nYVec[nN + k], nyz = nYVec[2 * nN + k];
    double nzx = nuy * nyz - nuz * nyy, nzy = nuz * nyx - nux * nyz, nzz = nux * nyy - nuy * nyx;
    double ru = radial[a], ry = radial[nA + a], rz = radial[2 * nA + a];
    double pax = ncx + ru * nux + ry * nyx + rz * nzx;
    double pay = ncy + ru * nuy + ry * nyy + rz * nzy;
    double paz = ncz + ru * nuz + ry * nyz + rz * nzz;
    double rcx = mCoord[rod], rcy = mCoord[nMB + rod], rcz = mCoord[2 * nMB + rod];
    double rux = mUVec[rod], ruy = mUVec[nMB + rod], ruz = mUVec[2 * nMB + rod];
    double rlen = mSegLength[rod];
    double re1x = rcx - 0.5 * rlen * rux, re1y = rcy - 0.5 * rlen * ruy, re1z = rcz - 0.5 * rlen * ruz;
    double fnx = 0, fny = 0, fnz = 0, tnx = 0, tny = 0, tnz = 0;
    double dx = pax - re1x, dy = pay - re1y, dz = paz - re1z;
    double strain = sqrt(dx * dx + dy * dy + dz * dz);
    if (strain > 0.0) {
        double inv = 1.0 / strain;
        double lx = dx * inv, ly = dy * inv, lz = dz * inv;
        double denom = dt * (1.0 / mBTransGam[nMB + rod] + nInvTransY[k]);
        double forceMag = (denom > 0.0) ? (coeff * 1.0e-6 * strain / denom) : 0.0;
        double Fx = forceMag * lx, Fy = forceMag * ly, Fz = forceMag * lz;
        mForceSum[rod] = (float)(mForceSum[rod] + Fx);
        mForceSum[nMB + rod] = (float)(mForceSum[nMB + rod] + Fy);
        mForceSum[2 * nMB + rod] = (float)(mForceSum[2 * nMB + rod] + Fz);
        if (atEnd1) {   /* dimer: the force acts at the rod's end1, arm -L/2 u (m) */
            double Rx = -0.5 * rlen * rux * 1.0e-6, Ry = -0.5 * rlen * ruy * 1.0e-6, Rz = -0.5 * rlen * ruz * 1.0e-6;
            mTorqueSum[rod] = (float)(mTorqueSum[rod] + (Ry * Fz - Rz * Fy));
            mTorqueSum[nMB + rod] = (float)(mTorqueSum[nMB + rod] + (Rz * Fx - Rx * Fz));
            mTorqueSum[2 * nMB + rod] = (float)(mTorqueSum[2 * nMB + rod] + (Rx * Fy - Ry * Fx));
        }
        fnx = -Fx; fny = -Fy; fnz = -Fz;
        if (atEnd1) {   /* node torque arm: the body-frame offset itself (NodeSystem.java:117-120) */
            double Rnx = ru * 1.0e-6, Rny = ry * 1.0e-6, Rnz = rz * 1.0e-6;
            tnx = Rny * fnz - Rnz * fny;
            tny = Rnz * fnx - Rnx * fnz;
            tnz = Rnx * fny - Rny * fnx;
        }
    }
    float *o = nodeData + NODE_STRIDE * a;
    o[0] = (float)fnx; o[1] = (float)fny; o[2] = (float)fnz;
    o[3] = (float)tnx; o[4] = (float)tny; o[5] = (float)tnz;
}

/* The tether over attachments keyed by separate node / motor arrays; boundSeg != NULL selects the
 * bound-stiffened coefficient nodeParams[1] for a motor whose head is bound (tetherBoundStiffen). */
void sbo_node_tether_keys(int nMB, int nN, int nA, const float *mCoord, const float *mUVec, const float *mSegLength,
                          const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                          const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachNode,
                          const int *attachMotor, const float *radial, const float *attachCoeffK, const int *boundSeg,
                          float *nodeData, const float *nodeParams) {
    double dt = nodeParams[0];
    for (int a = 0; a < nA; a++) {
        int m = attachMotor[a];
        double cs = attachCoeffK[a];
        int atEnd1 = cs > 0.0;
        double coeff = atEnd1 ? cs : -cs;
        if (boundSeg && boundSeg[m] >= 0) coeff = nodeParams[1];   /* NodeSystem.java:190-191 */
        node_tether_one(a, attachNode[a], 3 * m, atEnd1, coeff, nMB, nN, nA, mCoord, mUVec, mSegLength, mBTransGam,
                        mForceSum, mTorqueSum, nCoord, nUVec, nYVec, nInvTransY, radial, nodeData, dt);
    }
}

/* NodeSystem.tether (NodeSystem.java:48-127): attachKey = [node(nA) | motor(nA)] */
void sbo_node_tether(int nMB, int nN, int nA, const float *mCoord, const float *mUVec, const float *mSegLength,
                     const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                     const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachKey,
                     const float *radial, const float *attachCoeffK, float *nodeData, const float *nodeParams) {
    sbo_node_tether_keys(nMB, nN, nA, mCoord, mUVec, mSegLength, mBTransGam, mForceSum, mTorqueSum, nCoord, nUVec,
                         nYVec, nInvTransY, attachKey, attachKey + nA, radial, attachCoeffK, NULL, nodeData,
                         nodeParams);
}

/* NodeSystem.tetherBoundStiffen (NodeSystem.java:170-231) */
void sbo_node_tether_bound_stiffen(int nMB, int nN, int nA, const float *mCoord, const float *mUVec,
                                   const float *mSegLength, const float *mBTransGam, float *mForceSum,
                                   float *mTorqueSum, const float *nCoord, const float *nUVec, const float *nYVec,
                                   const float *nInvTransY, const int *attachKey, const float *radial,
                                   const float *attachCoeffK, const int *boundSeg, float *nodeData,
                                   const float *nodeParams) {
    sbo_node_tether_keys(nMB, nN, nA, mCoord, mUVec, mSegLength, mBTransGam, mForceSum, mTorqueSum, nCoord, nUVec,
                         nYVec, nInvTransY, attachKey, attachKey + nA, radial, attachCoeffK, boundSeg, nodeData,
                         nodeParams);
}

/* NodeSystem.xbridgeInject (NodeSystem.java:141-166): the bound heads' cross-bridge force (bondData[13m..+2])
 * added to the attachment's node-side record at the world-frame surface offset, with its torque. */
void sbo_node_xbridge_inject(int nN, int nA, const float *bondData, const float *nUVec, const float *nYVec,
                             const int *attachKey, const float *radial, const float *attachCoeffK,
                             const int *boundSeg, float *nodeData) {
    const int ST = 13;
    for (int a = 0; a < nA; a++) {
        int k = attachKey[a];
        int m = attachKey[nA + a];
        int dimer = attachCoeffK[a] > 0.0f;
        double nux = nUVec[k], nuy = nUVec[nN + k], nuz = nUVec[2 * nN + k];
        double nyx = nYVec[k], nyy = nYVec[nN + k], nyz = nYVec[2 * nN + k];
        double nzx = nuy * nyz - nuz * nyy, nzy = nuz * nyx - nux * nyz, nzz = nux * nyy - nuy * nyx;
        double ru = radial[a], ry = radial[nA + a], rz = radial[2 * nA + a];
        double Rx = ru * nux + ry * nyx + rz * nzx;
        double Ry = ru * nuy + ry * nyy + rz * nzy;
        double Rz = ru * nuz + ry * nyz + rz * nzz;
        double fx = 0, fy = 0, fz = 0;
        if (boundSeg[m] >= 0) { fx += bondData[m * ST]; fy += bondData[m * ST + 1]; fz += bondData[m * ST + 2]; }
        if (dimer && boundSeg[m + 1] >= 0) {
            fx += bondData[(m + 1) * ST]; fy += bondData[(m + 1) * ST + 1]; fz += bondData[(m + 1) * ST + 2];
        }
        double Rmx = Rx * 1.0e-6, Rmy = Ry * 1.0e-6, Rmz = Rz * 1.0e-6;
        double Tx = Rmy * fz - Rmz * fy, Ty = Rmz * fx - Rmx * fz, Tz = Rmx * fy - Rmy * fx;
        float *o = nodeData + NODE_STRIDE * a;
        o[0] = (float)(o[0] + fx); o[1] = (float)(o[1] + fy); o[2] = (float)(o[2] + fz);
        o[3] = (float)(o[3] + Tx); o[4] = (float)(o[4] + Ty); o[5] = (float)(o[5] + Tz);
    }
}

/* ------------------------------------------------------------------ node nucleation (NodeNucleationSystem.java) */
#define SALT_NUC1 0x4E554331u   /* "NUC1" */
#define SALT_NUCD 0x4E554344u   /* "NUCD" */

/* Four words of one event draw.  WANG: the hash of base = key ^ salt and of base ^ 0x9e3779b9 / 0x85ebca6b /
 * 0xc2b2ae35 (NodeNucleationSystem.java:76-83); PHILOX: the four words of counter (item, step, salt, 2), whose
 * first word is sbo_event_bits'. */
void sbo_event_bits4(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream, uint32_t salt,
                     uint32_t out[4]) {
    if (rng == SBO_RNG_WANG) {
        uint32_t base = (item * 1000003u) ^ ((uint32_t)step * 999983u) ^ ((uint32_t)seed * 7919u) ^ salt;
        out[0] = sbo_wang_hash(base);
        out[1] = sbo_wang_hash(base ^ 0x9e3779b9u);
        out[2] = sbo_wang_hash(base ^ 0x85ebca6bu);
        out[3] = sbo_wang_hash(base ^ 0xc2b2ae35u);
        return;
    }
    uint32_t key[2] = {(uint32_t)seed, stream};
    uint32_t c[4] = {item, (uint32_t)step, salt, 2u};
    sbo_philox4x32_10(c, key, out);
}

/* NodeNucleationSystem.countBoundFil (NodeNucleationSystem.java:55-62) */
void sbo_nuc_count_bound_fil(int C, int nN, const int *seedNode, int *nodeBoundFil) {
    for (int k = 0; k < nN; k++) nodeBoundFil[k] = 0;
    for (int s = 0; s < C; s++) {
        int k = seedNode[s];
        if (k >= 0) nodeBoundFil[k]++;
    }
}

/* NodeNucleationSystem.emit (NodeNucleationSystem.java:66-100): per node, a seed request below the formin cap
 * when the pool allows it and the draw falls under P_nuc; requests at stride RC; acceptFlag[k] written for
 * every node.  Float arithmetic as the reference's. */
void sbo_nuc_emit(int nN, int RC, const float *nodeCoord, const int *nodeBoundFil, int *acceptFlag, float *reqCoord,
                  float *reqUVec, float *reqYVec, const float *nucParams, int32_t step, int32_t seed, int poolOK,
                  int rng, uint32_t stream) {
    float pNuc = nucParams[0], halfLen = nucParams[1];
    int formins = (int)nucParams[2];
    for (int k = 0; k < nN; k++) {
        acceptFlag[k] = 0;
        if (nodeBoundFil[k] >= formins) continue;
        if (poolOK == 0) continue;
        uint32_t h[4];
        sbo_event_bits4(rng, (uint32_t)k, step, seed, stream, SALT_NUC1, h);
        float u = sbo_uniform(h[0]);
        if (u >= pNuc) continue;
        float dx = sbo_uniform(h[1]) * 2.0f - 1.0f;
        float dy = sbo_uniform(h[2]) * 2.0f - 1.0f;
        float dz = sbo_uniform(h[3]) * 2.0f - 1.0f;
        float m2 = dx * dx + dy * dy + dz * dz;
        if (m2 < 1.0e-12f) { dx = 1.0f; dy = 0.0f; dz = 0.0f; m2 = 1.0f; }
        float inv = 1.0f / (float)sqrt((double)m2);
        dx *= inv; dy *= inv; dz *= inv;
        float ncx = nodeCoord[k], ncy = nodeCoord[nN + k], ncz = nodeCoord[2 * nN + k];
        float ex, ey, ez;
        if (dx < 0.9f && dx > -0.9f) { ex = 1.0f; ey = 0.0f; ez = 0.0f; } else { ex = 0.0f; ey = 1.0f; ez = 0.0f; }
        float yx = dy * ez - dz * ey, yy = dz * ex - dx * ez, yz = dx * ey - dy * ex;
        float ym = 1.0f / (float)sqrt((double)(yx * yx + yy * yy + yz * yz));
        yx *= ym; yy *= ym; yz *= ym;
        reqCoord[k] = ncx + halfLen * dx; reqCoord[RC + k] = ncy + halfLen * dy; reqCoord[2 * RC + k] = ncz + halfLen * dz;
        reqUVec[k] = -dx; reqUVec[RC + k] = -dy; reqUVec[2 * RC + k] = -dz;
        reqYVec[k] = yx; reqYVec[RC + k] = yy; reqYVec[2 * RC + k] = yz;
        acceptFlag[k] = 1;
    }
}

/* NodeNucleationSystem.tagSeeds (NodeNucleationSystem.java:104-115): request r tethers its born slot to node r */
void sbo_nuc_tag_seeds(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                       int *seedNode) {
    int nFree = freeOffsets[C];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        seedNode[freeList[rank]] = r;
    }
}

/* NodeNucleationSystem.initNewborn (NodeNucleationSystem.java:228-244) */
void sbo_nuc_init_newborn(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                          int *monomerCount, float *segLength, const float *seedParams) {
    int nFree = freeOffsets[C];
    int mon = (int)seedParams[0];
    float len = seedParams[1];
    for (int r = 0; r < K; r++) {
        int rank = rankOffsets[r];
        if (!(rankOffsets[r + 1] > rank)) continue;
        if (rank >= nFree) continue;
        int slot = freeList[rank];
        monomerCount[slot] = mon;
        segLength[slot] = len;
    }
}

/* the seed tether force of slot s held by node k (NodeNucleationSystem.java:125-146): toward the node centre
 * from the barbed end2; returns 0 when the end sits on the centre */
static int seed_tether_force(int s, int k, int C, int nN, const float *filCoord, const float *filUVec,
                             const float *filSegLength, const float *filBTransGam, const float *nodeCoord,
                             const float *nodeInvTransX, double fracMove, double dt, double F[3], double *half,
                             double u[3]) {
    double cx = filCoord[s], cy = filCoord[C + s], cz = filCoord[2 * C + s];
    double ux = filUVec[s], uy = filUVec[C + s], uz = filUVec[2 * C + s];
    double h = 0.5 * filSegLength[s];
    double e2x = cx + h * ux, e2y = cy + h * uy, e2z = cz + h * uz;
    double dx = nodeCoord[k] - e2x, dy = nodeCoord[nN + k] - e2y, dz = nodeCoord[2 * nN + k] - e2z;
    double strain = sqrt(dx * dx + dy * dy + dz * dz);
    *half = h;
    u[0] = ux; u[1] = uy; u[2] = uz;
    if (!(strain > 0.0)) return 0;
    double inv = 1.0 / strain;
    double denom = dt * (1.0 / filBTransGam[s] + nodeInvTransX[k]);
    double fm = (denom > 0.0) ? (fracMove * 1.0e-6 * strain / denom) : 0.0;
    F[0] = fm * (dx * inv); F[1] = fm * (dy * inv); F[2] = fm * (dz * inv);
    return 1;
}

/* NodeNucleationSystem.seedTether (NodeNucleationSystem.java:117-149): force and end2 torque on each held seed */
void sbo_nuc_seed_tether(int C, int nN, const float *filCoord, const float *filUVec, const float *filSegLength,
                         const float *filBTransGam, float *filForceSum, float *filTorqueSum, const float *nodeCoord,
                         const float *nodeInvTransX, const int *seedNode, const float *tetherParams) {
    double fracMove = tetherParams[0], dt = tetherParams[1];
    for (int s = 0; s < C; s++) {
        int k = seedNode[s];
        if (k < 0) continue;
        double F[3], half, u[3];
        if (!seed_tether_force(s, k, C, nN, filCoord, filUVec, filSegLength, filBTransGam, nodeCoord, nodeInvTransX,
                               fracMove, dt, F, &half, u))
            continue;
        filForceSum[s] = (float)(filForceSum[s] + F[0]);
        filForceSum[C + s] = (float)(filForceSum[C + s] + F[1]);
        filForceSum[2 * C + s] = (float)(filForceSum[2 * C + s] + F[2]);
        double Rx = half * u[0] * 1.0e-6, Ry = half * u[1] * 1.0e-6, Rz = half * u[2] * 1.0e-6;
        filTorqueSum[s] = (float)(filTorqueSum[s] + (Ry * F[2] - Rz * F[1]));
        filTorqueSum[C + s] = (float)(filTorqueSum[C + s] + (Rz * F[0] - Rx * F[2]));
        filTorqueSum[2 * C + s] = (float)(filTorqueSum[2 * C + s] + (Rx * F[1] - Ry * F[0]));
    }
}

/* NodeNucleationSystem.seedTetherNodeReact (NodeNucleationSystem.java:161-187): each node adds -F of its held
 * seeds, summed in slot order.  The reference accumulates rx -= fm dx inv; the same products are formed here. */
void sbo_nuc_seed_tether_node_react(int C, int nN, const float *filCoord, const float *filUVec,
                                    const float *filSegLength, const float *filBTransGam, float *nodeForceSum,
                                    const float *nodeCoord, const float *nodeInvTransX, const int *seedNode,
                                    const float *tetherParams) {
    double fracMove = tetherParams[0], dt = tetherParams[1];
    for (int k = 0; k < nN; k++) {
        double ncx = nodeCoord[k], ncy = nodeCoord[nN + k], ncz = nodeCoord[2 * nN + k];
        double rx = 0.0, ry = 0.0, rz = 0.0;
        for (int s = 0; s < C; s++) {
            if (seedNode[s] != k) continue;
            double cx = filCoord[s], cy = filCoord[C + s], cz = filCoord[2 * C + s];
            double ux = filUVec[s], uy = filUVec[C + s], uz = filUVec[2 * C + s];
            double half = 0.5 * filSegLength[s];
            double e2x = cx + half * ux, e2y = cy + half * uy, e2z = cz + half * uz;
            double dx = ncx - e2x, dy = ncy - e2y, dz = ncz - e2z;
            double strain = sqrt(dx * dx + dy * dy + dz * dz);
            if (strain > 0.0) {
                double inv = 1.0 / strain;
                double denom = dt * (1.0 / filBTransGam[s] + nodeInvTransX[k]);
                double fm = (denom > 0.0) ? (fracMove * 1.0e-6 * strain / denom) : 0.0;
                rx -= fm * dx * inv; ry -= fm * dy * inv; rz -= fm * dz * inv;
            }
        }
        nodeForceSum[k] = (float)(nodeForceSum[k] + rx);
        nodeForceSum[nN + k] = (float)(nodeForceSum[nN + k] + ry);
        nodeForceSum[2 * nN + k] = (float)(nodeForceSum[2 * nN + k] + rz);
    }
}

/* NodeNucleationSystem.seedAimTorque (NodeNucleationSystem.java:196-211) */
void sbo_nuc_seed_aim_torque(int C, const float *filUVec, const float *filBRotGam, float *filTorqueSum,
                             const float *aimDir, const float *aimParams) {
    double aimCoeff = aimParams[0], dt = aimParams[1];
    for (int s = 0; s < C; s++) {
        double ax = aimDir[s], ay = aimDir[C + s], az = aimDir[2 * C + s];
        if (ax * ax + ay * ay + az * az < 0.25) continue;
        double ux = filUVec[s], uy = filUVec[C + s], uz = filUVec[2 * C + s];
        double cx = uy * az - uz * ay, cy = uz * ax - ux * az, cz = ux * ay - uy * ax;
        double k = aimCoeff * filBRotGam[C + s] / dt;
        filTorqueSum[s] = (float)(filTorqueSum[s] + k * cx);
        filTorqueSum[C + s] = (float)(filTorqueSum[C + s] + k * cy);
        filTorqueSum[2 * C + s] = (float)(filTorqueSum[2 * C + s] + k * cz);
    }
}

/* NodeNucleationSystem.dissolve (NodeNucleationSystem.java:248-257): a held seed detaches at P_detach */
void sbo_nuc_dissolve(int C, int *seedNode, const float *dissolveParams, int32_t step, int32_t seed, int rng,
                      uint32_t stream) {
    float pDetach = dissolveParams[0];
    for (int s = 0; s < C; s++) {
        if (seedNode[s] < 0) continue;
        float u = sbo_uniform(sbo_event_bits(rng, (uint32_t)s, step, seed, stream, SALT_NUCD));
        if (u < pDetach) seedNode[s] = -1;
    }
}
