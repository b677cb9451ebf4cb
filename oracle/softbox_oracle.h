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
 * Parity pins (what checks each function beyond retyping; tests/test_oracle_pins.py,
 * tests/test_oracle_grid_links.py, tests/test_oracle_motor.py):
 *   drag              values recorded in JOURNAL.md:2527-2532 + closed forms
 *   wang hash/uniform Java int semantics walked by hand (unsigned shifts, 32-bit wrap), the
 *                     uniform mapping's edge values, Box-Muller moments (mean 0, var 1, kurtosis 3)
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
                        int32_t step, int32_t seed, int rng, uint32_t stream, uint32_t bodyBase); /* BrownianForceSystem.java:44-106 */
void sbo_integrate(int n, float *coord, float *uVec, float *yVec, const float *forceSum, const float *torqueSum,
                   const float *randForce, const float *randTorque, const float *bTransGam, const float *bRotGam,
                   float dt);                                                             /* RigidRodLangevinIntegrationSystem.java:33-112 */
void sbo_derive(int n, const float *coord, const float *uVec, float *yVec, float *zVec, float *end1, float *end2,
                const float *segLength);                                                  /* DerivedGeometrySystem.java:27-73 */
void sbo_confine(int n, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                 float *forceSum, float *torqueSum, const float *boxParams, int32_t step);  /* ContainmentSystem.java:58-126 */

/* ---- chain (ChainBendingForceSystem.java:47-331) ---- */
double sbo_accurate_acos(double x);
void sbo_set_chain_reference_mobility(int on);  /* 1: sin(accurateAcos(c))^2 as the reference (:196-203) */
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

/* ---- myosin side (softbox_oracle_motor.c) ----
 * motor body: sub-body 3m rod, 3m+1 lever, 3m+2 head (MotorStore.java:45-60), stride nB = 3*nMotors */
void sbo_publish_head_from_body(int nM, const float *bodyCoord, const float *bodyUVec, const float *bodySegLength,
                                float *head, float *uVec, float *rodUVec);                 /* MotorStore.java:300-314 */
void sbo_publish_motors(int nM, int cap, int base, const float *head, const float *reach, float *center,
                        float *boundingRadius, int *ownerStore, int *ownerSlot);           /* MotorStore.java:326-347 */
void sbo_grid_reachable(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                        const float *segEnd1, const float *segEnd2, const float *gridParams, const int *gridDims,
                        const int *gridCellOffsets, const int *gridCellContents, const int *ownerStore,
                        const int *ownerSlot, int maxCand, int *reachSeg, int *reachCount,
                        const float *kinParams);                                          /* BindingDetectionSystem.java:143-228 */
void sbo_brute_reachable(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                         const float *segEnd1, const float *segEnd2, int maxCand, int *reachSeg, int *reachCount,
                         const float *kinParams);                                         /* :260-283 */
void sbo_bind_nearest(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                      const float *segEnd1, const float *segEnd2, const int *reachSeg, const int *reachCount,
                      int maxCand, int *boundSeg, float *bindArc, const float *kinParams); /* :356-389 */
void sbo_brute_reachable_node_aware(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                                    const float *segEnd1, const float *segEnd2, int maxCand, int *reachSeg,
                                    int *reachCount, const int *seedNode,
                                    const float *kinParams);                            /* BindingDetectionSystem.java:410-436 */
void sbo_bind_nearest_node_aware(int nM, int nSeg, const float *head, const float *uVec, const float *rodUVec,
                                 const float *segEnd1, const float *segEnd2, const int *reachSeg,
                                 const int *reachCount, int maxCand, int *boundSeg, float *bindArc,
                                 const int *seedNode, const float *kinParams);          /* :440-473 */
void sbo_catch_slip_release(int nM, int *boundSeg, const float *forceDotFil, const float *forceMag, int *cooldown,
                            int *stats, int *capStats, const float *kinParams, int32_t step, int32_t seed, int rng,
                            uint32_t stream);                                             /* NucleotideCycleSystem.java:77-135 */
void sbo_nucleotide_cycle(int nM, int *nucleotideState, const int *boundSeg, const float *forceDotHist,
                          const float *nucParams, int32_t step, int32_t seed, int rng, uint32_t stream); /* :33-70 */
void sbo_bond_forces(int nM, int nSeg, const float *motorCoord, const float *motorUVec, const float *motorYVec,
                     const float *motorBRotGam, const float *filCoord, const float *filUVec, const float *filYVec,
                     const float *filBRotGam, const float *filSegLength, const int *boundSeg, const float *bindArc,
                     const int *nucleotideState, float *bondData, const float *xbParams); /* CrossBridgeSystem.java:58-142 */
void sbo_apply_head_force(int nM, const float *bondData, float *bodyForceSum, float *bodyTorqueSum); /* :145-157 */
void sbo_register_force_dot(int nM, const float *bondData, const int *boundSeg, float *forceDotFil, float *forceMag,
                            float *forceDotHist, int *forceDotPlace);                    /* :166-187 */
void sbo_motor_seg_gather(int nSeg, const int *segMotorOffsets, const int *segMotorMyo, const float *bondData,
                          float *filForceSum, float *filTorqueSum);                       /* :302-320 */
void sbo_motor_joints(int nB, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                      const float *bRotGam, float *forceSum, float *torqueSum, const int *nucleotideState,
                      const float *jointParams);                                          /* MotorJointSystem.java:72-198 */
void sbo_dimer_couple(int nB, int nD, const float *coord, const float *uVec, const float *segLength,
                      const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                      const int *motorA, const int *motorB, const int *parallel, const float *dimerParams,
                      const int *boundSeg);                                               /* DimerCouplingSystem.java:83-184 */
void sbo_mini_tether(int nMB, int nBb, int nD, const float *mCoord, const float *mUVec, const float *mSegLength,
                     const float *mBTransGam, const float *mBRotGam, float *mForceSum, float *mTorqueSum,
                     const float *bbCoord, const float *bbUVec, const float *bbInvDragY, const int *headBackboneSlot,
                     const int *motorA, const float *attachAxial, float *miniData,
                     const float *miniParams);                                            /* MiniFilamentSystem.java:65-132 */
void sbo_backbone_gather(int nBb, const int *bbDimerOffsets, const int *bbDimerList, const float *miniData,
                         float *bbForceSum, float *bbTorqueSum);                          /* :137-155 */

/* ---- protein-node shell (softbox_oracle_node.c); attachKey = [node(nA) | motor(nA)] ---- */
void sbo_node_tether_keys(int nMB, int nN, int nA, const float *mCoord, const float *mUVec, const float *mSegLength,
                          const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                          const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachNode,
                          const int *attachMotor, const float *radial, const float *attachCoeffK, const int *boundSeg,
                          float *nodeData, const float *nodeParams);
void sbo_node_tether(int nMB, int nN, int nA, const float *mCoord, const float *mUVec, const float *mSegLength,
                     const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                     const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachKey,
                     const float *radial, const float *attachCoeffK, float *nodeData,
                     const float *nodeParams);                                            /* NodeSystem.java:48-127 */
void sbo_node_tether_bound_stiffen(int nMB, int nN, int nA, const float *mCoord, const float *mUVec,
                                   const float *mSegLength, const float *mBTransGam, float *mForceSum,
                                   float *mTorqueSum, const float *nCoord, const float *nUVec, const float *nYVec,
                                   const float *nInvTransY, const int *attachKey, const float *radial,
                                   const float *attachCoeffK, const int *boundSeg, float *nodeData,
                                   const float *nodeParams);                              /* :170-231 */
void sbo_node_xbridge_inject(int nN, int nA, const float *bondData, const float *nUVec, const float *nYVec,
                             const int *attachKey, const float *radial, const float *attachCoeffK,
                             const int *boundSeg, float *nodeData);                       /* :141-166 */
/* ---- node nucleation (NodeNucleationSystem.java); seeds, events and the node bond of each slot ---- */
void sbo_event_bits4(int rng, uint32_t item, int32_t step, int32_t seed, uint32_t stream, uint32_t salt,
                     uint32_t out[4]);
void sbo_nuc_count_bound_fil(int C, int nN, const int *seedNode, int *nodeBoundFil);         /* :55-62 */
void sbo_nuc_emit(int nN, int RC, const float *nodeCoord, const int *nodeBoundFil, int *acceptFlag, float *reqCoord,
                  float *reqUVec, float *reqYVec, const float *nucParams, int32_t step, int32_t seed, int poolOK,
                  int rng, uint32_t stream);                                                 /* :66-100 */
void sbo_nuc_tag_seeds(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                       int *seedNode);                                                      /* :104-115 */
void sbo_nuc_init_newborn(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                          int *monomerCount, float *segLength, const float *seedParams);    /* :228-244 */
void sbo_nuc_seed_tether(int C, int nN, const float *filCoord, const float *filUVec, const float *filSegLength,
                         const float *filBTransGam, float *filForceSum, float *filTorqueSum, const float *nodeCoord,
                         const float *nodeInvTransX, const int *seedNode, const float *tetherParams); /* :117-149 */
void sbo_nuc_seed_tether_node_react(int C, int nN, const float *filCoord, const float *filUVec,
                                    const float *filSegLength, const float *filBTransGam, float *nodeForceSum,
                                    const float *nodeCoord, const float *nodeInvTransX, const int *seedNode,
                                    const float *tetherParams);                            /* :161-187 */
void sbo_nuc_seed_aim_torque(int C, const float *filUVec, const float *filBRotGam, float *filTorqueSum,
                             const float *aimDir, const float *aimParams);                 /* :196-211 */
void sbo_nuc_dissolve(int C, int *seedNode, const float *dissolveParams, int32_t step, int32_t seed, int rng,
                      uint32_t stream);                                                    /* :248-257 */

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
    /* crosslinker formation at a cadence (V2OneXHarness.java:379): NULL = off */
    struct sbo_form *form;
    int formInt;
    /* actin turnover at the biochem cadence (FullSystemDemoHarness.java:650-684): NULL = off */
    struct sbo_turnover *turn;
} sbo_net;

/* ---- crosslinker formation (softbox_oracle_form.c; CrosslinkerSystem.java:298-760, FormationGrid.java,
 * FilIDSystem.java) ---- */
#define SBO_FORM_MAXC 256      /* CrosslinkerSystem.FORM_MAXC (CrosslinkerSystem.java:457) */
#define SBO_STRAIN_WIN 10
#define SBO_TORQUE_WIN 5
int sbo_filid_rounds(int n);                                                              /* FilIDSystem.java:31-34 */
void sbo_filid_init(int n, const int *filState, const int *end2NbrSlot, int *ptr);       /* :38-47 */
void sbo_filid_jump(int n, const int *ptrIn, int *ptrOut, const int *filState);          /* :52-59 */
void sbo_filid(int n, const int *filState, const int *end2NbrSlot, int *filID, int *scratch);
void sbo_fil_fil_candidates(int nSeg, const float *coord, const float *segLength, const int *filID, int reqCap,
                            int *reqFilA, int *reqFilB, const float *formParams, int *formCounts, int periodic,
                            const float *boxLen);                                         /* CrosslinkerSystem.java:405-431 */
void sbo_grid_form_count(int nSeg, const float *coord, const float *segLength, const int *filID,
                         const float *gridParams, const int *gridDims, const int *offsets, const int *contents,
                         int *candCountSeg, const float *formParams, int periodic, const float *boxLen); /* :461-513 */
void sbo_grid_form_scan(int nSeg, const int *candCountSeg, int *candBaseSeg, int reqCap, int *reqFilA, int *reqFilB,
                        int *formCounts);                                                 /* :516-529 */
void sbo_grid_form_emit(int nSeg, const float *coord, const float *segLength, const int *filID,
                        const float *gridParams, const int *gridDims, const int *offsets, const int *contents,
                        const int *candBaseSeg, const int *candCountSeg, int reqCap, int *reqFilA, int *reqFilB,
                        const float *formParams, int periodic, const float *boxLen);     /* :535-604 */
void sbo_form_gates(int nSeg, int nCand, const float *coord, const float *uVec, const float *end1, const float *end2,
                    const float *segLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
                    int *reqOrient, int *gatePass, const float *formParams, int32_t step, int32_t seed, int mode,
                    int rng, uint32_t stream, int periodic, const float *boxLen);         /* :607-676 */
void sbo_form_admit_reduce(int nCand, int reqCap, int nSeg, const int *reqFilA, const int *reqFilB,
                           const int *gatePass, int *minCand);                            /* :680-691 */
void sbo_form_admit(int nCand, int C, const int *reqFilA, const int *reqFilB, const float *reqLoc1,
                    const float *reqLoc2, const int *gatePass, const int *minCand, const int *activeLinkCount,
                    const int *linkState, const int *linkFilA, const int *linkFilB, const float *loc1,
                    const float *loc2, int *acceptFlag, const float *formParams);         /* :713-738 */
int sbo_form_allocate(int K, int C, const int *acceptFlag, const int *reqFilA, const int *reqFilB,
                      const float *reqLoc1, const float *reqLoc2, const int *reqOrient, int *freeCount,
                      int *freeOffsets, int *freeList, int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1,
                      float *loc2, int *linkState, float *strainHist, int *strainPlace, int *linkOrientSame,
                      float *torqueMagHist, int *torqueMagPlace);                         /* :298-357, 742-760 */
typedef struct sbo_form {
    int nSeg, C, reqCap;
    const float *coord, *uVec, *end1, *end2, *segLength;    /* filament store (pose at the pass) */
    const int *filState, *end2NbrSlot;                      /* filState NULL: every slot ACTIVE */
    int *linkFilA, *linkFilB, *linkState, *strainPlace, *linkOrientSame, *torqueMagPlace, *activeLinkCount;
    float *loc1, *loc2, *strainHist, *torqueMagHist;
    float gridParams[6];                                    /* the formation grid (FormationGrid.java:63-76) */
    int gridDims[4];
    int periodic;
    float boxLen[3];
    float formParams[7];                                    /* CrosslinkerStore.setFormParams (:196-206) */
    int mode;
    int formCounts[4];                                      /* [0] = candidates of the last pass */
    int *filID, *filIDScratch, *bodyCell, *cellCount, *gridCellOffsets, *gridCellContents, *candCountSeg,
        *candBaseSeg, *reqFilA, *reqFilB, *reqOrient, *gatePass, *minCand, *acceptFlag, *freeCount, *freeOffsets,
        *freeList, *rankOffsets;
    float *center, *reqLoc1, *reqLoc2;
    int lastFormed;
} sbo_form;
int sbo_formation(sbo_form *f, int32_t step, int32_t seed, int rng, uint32_t stream); /* V2OneXHarness.java:444-464 */

/* ---- actin turnover (softbox_oracle_turnover.c; AgingSystem, DepolySystem, GrowthSystem, SeveringSystem,
 * FilamentBirthSystem, ActinPool) ---- */
#define SBO_FIL_ACTIVE 0      /* FilamentStore.FIL_ACTIVE (FilamentStore.java:112) */
#define SBO_FIL_FREE (-1)     /* FilamentStore.FIL_FREE (:113) */
void sbo_aging_age(int C, const int *filState, float *nucFrac, const float agingParams[2], int fires);
void sbo_aging_growth_atp(int C, const int *grewFlag, const int *monomerCount, float *nucFrac);
void sbo_aging_split_inherit_nuc(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                 float *nucFrac);
void sbo_aging_nucleate_fresh_atp(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                  float *nucFrac);
void sbo_depoly(int C, const int *filState, int *monomerCount, float *coord, const float *uVec,
                const int *end1NbrSlot, const float *nucFrac, int *returnedMon, int *deathFlag,
                const float depolyParams[3], const float depolyRateParams[2], int32_t step, int32_t seed, int rng,
                uint32_t stream, int fires);
void sbo_apply_death(int C, int *filState, int *monomerCount, int *seedNode, int *e1s, int *e1d, int *e2s, int *e2d,
                     float *bts, float *brs, const int *deathFlag);
void sbo_grow(int C, const int *seedNode, int *monomerCount, float *coord, const float *uVec, int *grewFlag,
              const float growParams[3], int32_t step, int32_t seed, int rng, uint32_t stream, int fires);
void sbo_mark_splits(int C, const int *seedNode, const int *monomerCount, const float *coord, const float *uVec,
                     const float *yVec, int *acceptFlag, float *reqCoord, float *reqUVec, float *reqYVec, int RC,
                     const float splitParams[3], int fires);
void sbo_split_wire(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                    int *monomerCount, float *coord, const float *uVec, int *e1s, int *e1d, int *e2s, int *e2d,
                    int *seedNode, const float splitParams[3]);
void sbo_recompute_drag(int C, const int *monomerCount, float *segLength, const int *end1NbrSlot,
                        const int *end2NbrSlot, float *bTransGam, float *bRotGam, float *bTransDiff, float *bRotDiff,
                        const float dragParams[9], int fires);
void sbo_cofilin_accumulate(int C, const int *filState, const float *nucFrac, float *cofFrac,
                            const float cofilinParams[2], int fires);
void sbo_cofilin_dissolve(int C, const int *filState, const int *monomerCount, const float *cofFrac,
                          int *severDeathFlag, int *severReturnedMon, const float cofilinParams[2], int fires);
void sbo_nucleate_fresh_cofilin(int C, int K, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                float *cofFrac);
int sbo_csr_scan(int n, int *counts, int *offsets);
void sbo_birth_alloc(int C, int K, int *filState, int *freeCount, int *freeOffsets, int *freeList, int *acceptFlag,
                     int *rankOffsets, const float *reqCoord, const float *reqUVec, const float *reqYVec, int RC,
                     float *coord, float *uVec, float *yVec, float *bts, float *brs, const float birthParams[2]);
typedef struct sbo_turnover {
    int C;                                    /* slots: the filament store's capacity */
    int *filState, *monomerCount, *seedNode;  /* FilamentStore.filState/monomerCount, NodeNucleationStore.seedNode */
    float *nucFrac;                           /* 3C: ATP | ADP-Pi | ADP (AgingStore.nucFrac) */
    float *cofFrac;                           /* C (SeverStore.cofFrac) */
    float *bTransDiff, *bRotDiff;             /* 3C each, or NULL */
    int *grewFlag, *returnedMon, *deathFlag, *severDeathFlag, *severReturnedMon, *acceptFlag;
    float *reqCoord, *reqUVec, *reqYVec;      /* 3C planar, request r = slot r */
    int *freeCount, *freeOffsets, *freeList, *rankOffsets, *scanOffsets;   /* offsets: C + 1 */
    int interval, seed, growthOn, proxy;
    float growParams[3], splitParams[3], dragParams[9], depolyParams[3], depolyRateParams[2], agingParams[2],
        cofilinParams[2], birthParams[2];
    double onRate, biochemDeltaT, uMPerMonomer;
    double conc;                              /* ActinPool: free actin (uM) and the integer ledger */
    long long taken, returned;
    int last[6];                              /* last cadence: grew, returned, severed monomers, deaths,
                                                 dissolves, splits */
    int cadences;
} sbo_turnover;
void sbo_turnover_cadence(sbo_turnover *t, sbo_net *net, int32_t step);   /* FullSystemDemoHarness.java:650-684 */

void sbo_network_step(sbo_net *net, int32_t step);
void sbo_network_grid(sbo_net *net);                   /* publish + body cell + grid + broad phase */
void sbo_network_forces(sbo_net *net, int32_t step);   /* zero, Brownian, chain, crosslinks + gathers */
void sbo_network_finish(sbo_net *net, int32_t step);   /* confine, integrate, derive */

/* ---- one whole contractile step: motors + minifilament backbones + the filament network ----
 * DenseContractileHarness.cpuStep (DenseContractileHarness.java:350-429) without turnover.  The body
 * view holds filaments in slots [0, nRod) and motor heads in [nRod, nRod + nM) (base = nRod). */
typedef struct sbo_motors {
    int nM;                                   /* motors; sub-bodies nB = 3*nM */
    float *coord, *uVec, *yVec, *zVec, *end1, *end2, *segLength, *bTransGam, *bRotGam;
    float *forceSum, *torqueSum, *randForce, *randTorque, *brownTransScale, *brownRotScale;
    float *head, *hUVec, *rodUVec, *reach;    /* published binding interface, stride nM */
    int *boundSeg, *cooldown, *stats, *capStats, *nucleotideState, *forceDotPlace;
    float *bindArc, *forceDotFil, *forceMag, *forceDotHist;
    float bodyParams[2], jointParams[11], kinParams[14], nucParams[8], xbParams[6];
    int *reachSeg, *reachCount;               /* nM*maxCand, nM */
    float *bondData;                          /* 13*nM */
    int *segMotorCount, *segMotorOffsets, *segMotorMyo;
    int nD;                                   /* dimers */
    int *motorA, *motorB, *parallel;
    float dimerParams[5];
    int nBb, nMD;                             /* backbones, tethered dimers */
    float *bbCoord, *bbUVec, *bbYVec, *bbZVec, *bbEnd1, *bbEnd2, *bbSegLength, *bbTransGam, *bbRotGam;
    float *bbForceSum, *bbTorqueSum, *bbRandForce, *bbRandTorque, *bbBts, *bbBrs, *bbInvDragY;
    int *headBackboneSlot, *tetherMotor;
    float *attachAxial, *miniData;
    int *bbDimerCount, *bbDimerOffsets, *bbDimerList;
    float miniParams[3], bbBodyParams[2];
    int32_t seedMotor, seedBackbone;
    int *ownerStore, *ownerSlot;              /* body view back-pointers, cap = nRod + nM */
    /* node shell: nodeMode 1 (NodeSystem.tether) or 2 (tetherBoundStiffen, boundCoeff miniParams[1]) makes
     * the backbone bodies node spheres and each tethered item an attachment (headBackboneSlot = node,
     * tetherMotor = motor, miniData = nodeData, bbInvDragY = nodeInvTransY); 0 keeps the minifilament tether */
    float *radial, *attachCoeffK;
    int nodeMode;
} sbo_motors;

void sbo_contractile_step(sbo_net *net, sbo_motors *mo, int32_t step);

#ifdef __cplusplus
}
#endif
#endif
