/*
 * softbox_oracle.h -- CPU restatement of SoftBox's per-timestep rigid-rod mechanics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests, the
 * smoke() entry point and bench.py's cpu_baseline leg call it to produce the
 * expected outputs that the HIP library (libsoftbox_hip.so) is compared against.
 * The product path never links, loads or calls anything in this directory.
 *
 * Every function restates one reference System method (Java, read as text) with
 * the same float/double types and the same operation order, so that a build with
 * -ffp-contract=off (no FMA contraction, like Java's strict IEEE semantics) gives
 * the reference's FP32-state results.  The reference file:line each function
 * follows is given beside it.  All arrays are planar SoA (x-plane | y-plane |
 * z-plane, plane stride = n) exactly as in RigidRodBody.java:26-57.
 *
 * Parity pins (what checks each function beyond retyping):
 *   drag              closed forms recorded in JOURNAL.md:2527-2532 (tests/test_oracle_drag.py)
 *   wang hash/uniform integer semantics checked against an independent 64-bit
 *                     Python model + statistical moments (tests/test_oracle_rng.py)
 *   philox4x32-10     Random123 known-answer vectors (Salmon et al., SC'11)
 *   brownian/integrate FDT closed forms D = kT/gamma within 3 sigma; zero-noise
 *                     invariance; closed-form single-step rotation
 *   derive            orthonormality invariants
 *   grid/broad phase  O(N^2) brute force set equality; textbook stable sort
 *   chain             frozen deflection known-answer test
 *                     (fixtures/filament_characterization_v1.md:11-25) + Newton 3rd law
 *   csr/gathers       brute gathers (bit-identical by construction) + momentum sum
 *   link force        exact zero at rest; analytic per-step decay
 *                     (CrosslinkerHarness.java:300-347); equal and opposite
 *   link torsion      closed-form angle for constructed geometry
 *   link unbind       closed-form P_break = (c0 + c1 e^{c2 s}) dt over a population
 *   confine           no-op inside; closed-form push for a constructed penetration
 */
#ifndef SOFTBOX_ORACLE_H
#define SOFTBOX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (Constants.java:25-67) ---- */
double sbo_const(int which);   /* 0 kT, 1 aeta, 2 actinMonoRadius, 3 radius, 4 aParallel, 5 aOrthog, 6 aTurning */

/* ---- drag (DragTensorSystem.java:33-97) ---- */
void sbo_rod_drag_si(double lengthUm, double radiusUm, double aeta, double g[6]);
void sbo_sphere_drag_si(double radiusUm, double aeta, double g[6]);
void sbo_drag_run_filament(int n, const int *monomerCount, const int *end1NbrSlot, const int *end2NbrSlot,
                           float *segLength, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff,
                           double aeta, double kT);
void sbo_drag_fixed_length(int n, double lengthUm, double radiusUm, double aeta, double kT,
                           float *segLength, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff);

/* ---- random streams ---- */
enum { SBO_RNG_WANG = 0, SBO_RNG_PHILOX = 1 };
uint32_t sbo_wang_hash(uint32_t seed);
void sbo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* the six raw 32-bit draws of one body-step (Wang: BrownianForceSystem.java:64-70) */
void sbo_brownian_bits(int rng, uint32_t body, int32_t step, int32_t seed, uint32_t stream, uint32_t h[6]);
/* one raw 32-bit event draw keyed by (item, step, seed, salt) (CrosslinkerSystem.java:277) */
uint32_t sbo_event_bits(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream, uint32_t salt);
float sbo_uniform(uint32_t h);  /* (h>>>1)/2147483647f */

/* ---- per-body systems ---- */
void sbo_zero_accumulators(int n, float *forceSum, float *torqueSum);                   /* ChainBendingForceSystem.java:38-45 */
void sbo_brownian_force(int n, float *randForce, float *randTorque, const float *bTransGam, const float *bRotGam,
                        const float *brownTransScale, const float *brownRotScale, float brownianForceMag,
                        int32_t step, int32_t seed, int rng, uint32_t stream, uint32_t bodyBase); /* BrownianForceSystem.java:183-245 */
void sbo_integrate(int n, float *coord, float *uVec, float *yVec, const float *forceSum, const float *torqueSum,
                   const float *randForce, const float *randTorque, const float *bTransGam, const float *bRotGam,
                   float dt);                                                             /* RigidRodLangevinIntegrationSystem.java:33-112 */
void sbo_derive(int n, const float *coord, const float *uVec, float *yVec, float *zVec, float *end1, float *end2,
                const float *segLength);                                                  /* DerivedGeometrySystem.java:140-186 */
void sbo_confine(int n, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                 float *forceSum, float *torqueSum, const float *boxParams, int32_t step);  /* ContainmentSystem.java:58-126 */

/* ---- chain (ChainBendingForceSystem.java:47-331) ---- */
double sbo_accurate_acos(double x);
double sbo_accurate_asin(double s);
double sbo_angle_from_sin_cos(double sin2, double cos);
void sbo_chain_forces(int n, const float *coord, const float *uVec, const float *segLength,
                      const int *end2NbrSlot, const int *end2NbrSide, const int *end1NbrSlot, const int *end1NbrSide,
                      const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                      const float *chainParams);

/* ---- spatial grid (SpatialGrid.java:51-445) ---- */
/* gridParams: [0..2] min, [3] cellSize, [4] invCellSize, [5] cutoff; gridDims: nX nY nZ total.
 * periodic != 0 selects the build's periodic extension (wrapped cells, minimum image). boxLen = nX*cellSize per axis. */
void sbo_body_cell(int cap, int S, const float *center, const float *gridParams, const int *gridDims, int periodic,
                   int *bodyCell);
void sbo_grid_build_serial(int S, int contentsCap, const int *bodyCell, const int *gridDims,
                           int *cellCount, int *gridCellOffsets, int *gridCellContents);
void sbo_broad_phase(int cap, int S, const float *center, const float *boundingRadius, const int *bodyCell,
                     const int *gridCellOffsets, const int *gridCellContents, const int *gridDims,
                     const float *gridParams, int periodic, const float *boxLen, int maxCand,
                     int *candPartner, int *candCount);
void sbo_brute_force(int cap, int S, const float *center, const float *boundingRadius, const float *gridParams,
                     int periodic, const float *boxLen, int maxCand, int *brutePartner, int *bruteCount);
void sbo_publish_filaments(int n, int cap, const float *coord, const float *segLength, float actinRadius,
                           float *center, float *boundingRadius);                         /* FilamentStore.java:275-295 */

/* ---- CSR inverse (CrossBridgeSystem.java:190-217) ---- */
void sbo_csr_build(int nKeys, int nItems, const int *key, int *count, int *offsets, int *list);

/* ---- crosslinks (CrosslinkerSystem.java) ---- */
void sbo_count_active_links(int C, int nSeg, const int *linkState, const int *linkFilA, const int *linkFilB,
                            int *activeLinkCount);                                        /* :694-706 */
void sbo_link_unbind(int nSeg, int nLinks, const float *filUVec, const float *filEnd1,
                     const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2,
                     int *linkState, float *strainHist, int *strainPlace, const float *offParams,
                     int32_t step, int32_t seed, int rng, uint32_t stream);               /* :231-281 */
void sbo_link_forces(int nSeg, int nLinks, const float *filCoord, const float *filUVec, const float *filEnd1,
                     const float *filBTransGam, const int *linkFilA, const int *linkFilB, const float *loc1,
                     const float *loc2, const int *filLinkCt, float *xlinkData, const float *xlParams); /* :51-124 */
void sbo_link_torsion(int nSeg, int nLinks, const float *filUVec, const int *linkFilA, const int *linkFilB,
                      const int *linkState, const int *linkOrientSame, float *torqueMagHist, int *torqueMagPlace,
                      float *xlinkData, const float *torsionParams);                      /* :767-808 */
void sbo_seg_gather_a(int nSeg, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                      const int *linkState, float *filForceSum, float *filTorqueSum);     /* :129-149 */
void sbo_seg_gather_b(int nSeg, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                      const int *linkState, float *filForceSum, float *filTorqueSum);     /* :153-173 */
void sbo_link_brute_gather(int nSeg, int nLinks, const int *linkFilA, const int *linkFilB, const float *xlinkData,
                           const int *linkState, float *bForceSum, float *bTorqueSum);    /* :178-198 */

/* ---- one whole network step (DenseContractileHarness.java:350-429 filament block, no motors) ---- */
typedef struct sbo_net {
    int nRod;
    float *coord, *uVec, *yVec, *zVec, *end1, *end2, *segLength;
    float *bTransGam, *bRotGam, *forceSum, *torqueSum, *randForce, *randTorque;
    float *brownTransScale, *brownRotScale;
    int *end1NbrSlot, *end1NbrSide, *end2NbrSlot, *end2NbrSide;
    float params[2];          /* dt, sqrt(2kT/dt) */
    float chainParams[7];
    int chainOn;
    /* crosslinks */
    int nLinks;
    int *linkFilA, *linkFilB, *linkState, *strainPlace, *linkOrientSame, *torqueMagPlace, *activeLinkCount;
    float *loc1, *loc2, *strainHist, *torqueMagHist, *xlinkData;
    float xlParams[3], offParams[5], torsionParams[2];
    int unbindOn;
    int *segCountA, *segOffsetsA, *segIdxA, *segCountB, *segOffsetsB, *segIdxB;
    /* containment */
    int confineOn;
    float boxParams[8];
    /* grid + broad phase */
    int gridOn, periodic, maxCand;
    float gridParams[6], boxLen[3], actinRadius;
    int gridDims[4];
    float *center, *boundingRadius;
    int *bodyCell, *cellCount, *gridCellOffsets, *gridCellContents, *candPartner, *candCount;
    /* random stream */
    int32_t seed;
    int rng;
    uint32_t stream;
} sbo_net;

void sbo_network_step(sbo_net *net, int32_t step);

#ifdef __cplusplus
}
#endif
#endif
