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
