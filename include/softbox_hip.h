/*
 * softbox_hip.h -- C ABI of libsoftbox_hip.so, the MI355X (gfx950) mechanics hot path of SoftBox.
 *
 * This boundary replaces the TornadoVM TaskGraph dispatch of SoftBox's per-timestep rigid-rod
 * Systems.  Each System entry point takes the same arrays, in the same order, as the Java
 * System method it replaces (cited beside it), so a java.lang.foreign downcall can stand in
 * for `.task(name, System::method, args...)`:
 *   - big per-body / per-link arrays are DEVICE pointers (allocated with sbx_alloc or any HIP
 *     allocator on the context's device), planar SoA with plane stride n exactly as the
 *     reference stores (RigidRodBody.java:26-57, FilamentStore.java:29-87);
 *   - the tiny scalar arrays the reference re-uploads every execution (`params`, `counts`,
 *     `chainParams`, `gridParams`, ...) are HOST pointers read at call time (the
 *     EVERY_EXECUTION transfer becomes kernel arguments; no device copy);
 *   - array lengths the Java code takes from getSize() are explicit int arguments.
 * All calls are asynchronous on the context's stream; sbx_pull and sbx_sync synchronise.
 *
 * Errors: every int-returning function returns SBX_OK (0) or a negative SBX_ERR_*.  Argument
 * errors are reported and leave the context usable; HIP/runtime failures are sticky (every
 * later call returns the same code until sbx_clear_error).  sbx_last_error gives the text.
 * Nothing aborts the host process.  One host thread drives one context.
 */
#ifndef SOFTBOX_HIP_H
#define SOFTBOX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBX_OK 0
#define SBX_ERR_ARG -1      /* invalid argument (size, null pointer, out-of-range parameter) */
#define SBX_ERR_HIP -2      /* HIP runtime failure (sticky) */
#define SBX_ERR_NOMEM -3    /* device allocation failed (sticky) */
#define SBX_ERR_STATE -4    /* call not valid in the current state */

#define SBX_RNG_WANG 0      /* reference stream: Wang hash keyed (i*1000003)^(step*999983)^(seed*7919) */
#define SBX_RNG_PHILOX 1    /* counter-based Philox4x32-10 keyed (seed, stream) / (body, step, block) */

#define SBX_MAX_CAND 256    /* SpatialGrid.MAX_CAND (SpatialGrid.java:48) */
#define SBX_XL_STRIDE 12    /* CrosslinkerStore.STRIDE (CrosslinkerStore.java:33) */
#define SBX_STRAIN_WIN 10   /* CrosslinkerStore.STRAIN_WIN (CrosslinkerStore.java:64) */
#define SBX_TORQUE_WIN 5    /* CrosslinkerStore.TORQUE_WIN (CrosslinkerStore.java:118) */

typedef struct sbx_ctx sbx_ctx;
typedef struct sbx_net sbx_net;

/* ------------------------------------------------------------------ context and memory
 * Replaces the TornadoExecutionPlan / DataTransferMode plumbing: FIRST_EXECUTION ->
 * sbx_push once, EVERY_EXECUTION -> host scalar arrays passed per call, UNDER_DEMAND -> sbx_pull. */
int sbx_create(sbx_ctx **out, int device, void *hip_stream); /* hip_stream NULL: library-owned stream */
void sbx_destroy(sbx_ctx *c);
const char *sbx_last_error(const sbx_ctx *c);
void sbx_clear_error(sbx_ctx *c);
int sbx_sync(sbx_ctx *c);
int sbx_alloc(sbx_ctx *c, void **dptr, size_t bytes);
int sbx_free(sbx_ctx *c, void *dptr);
int sbx_push(sbx_ctx *c, void *dst_dev, const void *src_host, size_t bytes);   /* transferToDevice */
int sbx_pull(sbx_ctx *c, void *dst_host, const void *src_dev, size_t bytes);   /* transferToHost (synchronous) */
int sbx_copy(sbx_ctx *c, void *dst_dev, const void *src_dev, size_t bytes);
int sbx_fill_i32(sbx_ctx *c, int *dst_dev, int value, size_t count);           /* kernel fill (graph-safe) */
const char *sbx_version(void);

/* ------------------------------------------------------------------ per-body systems */
/* ChainBendingForceSystem.zeroAccumulators (ChainBendingForceSystem.java:38-45) */
int sbx_zero_accumulators(sbx_ctx *c, float *forceSum, float *torqueSum, int n);

/* BrownianForceSystem.brownianForce (BrownianForceSystem.java:44-106); params[1] = sqrt(2kT/dt),
 * counts[1] = step, counts[2] = seed.  rng/stream select the random stream (stream: replica key). */
int sbx_brownian_force(sbx_ctx *c, float *randForce, float *randTorque, const float *bTransGam, const float *bRotGam,
                       const float *brownTransScale, const float *brownRotScale, const float *params,
                       const int *counts, int n, int rng, uint32_t stream);

/* RigidRodLangevinIntegrationSystem.integrate (RigidRodLangevinIntegrationSystem.java:33-112); params[0] = dt */
int sbx_integrate(sbx_ctx *c, float *coord, float *uVec, float *yVec, const float *forceSum, const float *torqueSum,
                  const float *randForce, const float *randTorque, const float *bTransGam, const float *bRotGam,
                  const float *params, const int *counts, int n);

/* DerivedGeometrySystem.derive (DerivedGeometrySystem.java:27-73) */
int sbx_derive(sbx_ctx *c, const float *coord, const float *uVec, float *yVec, float *zVec, float *end1, float *end2,
               const float *segLength, const int *counts, int n);

/* MechanicsFusion.forceBuild (MechanicsFusion.java:52-114) = zeroAccumulators + brownianForce */
int sbx_force_build(sbx_ctx *c, float *forceSum, float *torqueSum, float *randForce, float *randTorque,
                    const float *bTransGam, const float *bRotGam, const float *brownTransScale,
                    const float *brownRotScale, const float *params, const int *counts, int n, int rng,
                    uint32_t stream);

/* MechanicsFusion.integDerive (MechanicsFusion.java:121-201) */
int sbx_integ_derive(sbx_ctx *c, float *coord, float *uVec, float *yVec, float *zVec, const float *segLength,
                     float *end1, float *end2, const float *forceSum, const float *torqueSum, const float *randForce,
                     const float *randTorque, const float *bTransGam, const float *bRotGam, const float *params,
                     int n);

/* MechanicsFusion.integDeriveConfined (MechanicsFusion.java:209-338); boxParams[7] = dt, counts[1] = step */
int sbx_integ_derive_confined(sbx_ctx *c, float *coord, float *uVec, float *yVec, float *zVec,
                              const float *segLength, float *end1, float *end2, float *forceSum, float *torqueSum,
                              const float *randForce, const float *randTorque, const float *bTransGam,
                              const float *bRotGam, const float *boxParams, const int *counts, int n);

/* ContainmentSystem.confine (ContainmentSystem.java:58-126) */
int sbx_confine(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength, const float *bTransGam,
                float *forceSum, float *torqueSum, const float *boxParams, const int *counts, int n);

/* ChainBendingForceSystem.chainForces (ChainBendingForceSystem.java:105-331) */
int sbx_chain_forces(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const int *end2NbrSlot, const int *end2NbrSide, const int *end1NbrSlot, const int *end1NbrSide,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const float *chainParams, const int *counts, int n);

/* ------------------------------------------------------------------ spatial grid
 * gridParams (host): [0..2] xMin yMin zMin, [3] cellSize, [4] invCellSize, [5] cutoff
 * gridDims (host):   [0..2] nX nY nZ, [3] totalCells;  counts (host): [1] = S active bodies
 * periodic: 0 = reference clamped grid; 1 = periodic extension (cells wrap, minimum image,
 * boxLen = nBins * cellSize per axis, needs nBins >= 3). */

/* FilamentStore.publishToBodyView (FilamentStore.java:275-295); viewParams[0] = actin radius, cap = view capacity */
int sbx_publish_filaments(sbx_ctx *c, const float *coord, const float *segLength, float *center,
                          float *boundingRadius, int *ownerStore, int *ownerSlot, const float *viewParams,
                          int n, int cap);

/* SpatialGrid.bodyCell (SpatialGrid.java:51-75) */
int sbx_body_cell(sbx_ctx *c, const float *center, const float *gridParams, const int *gridDims, const int *counts,
                  int *bodyCell, int cap, int periodic);

/* The whole stable counting sort of SpatialGrid.gridChunkZero/Histogram/Reduce + gridScanLocal/Chunks/Add +
 * gridChunkScatter (SpatialGrid.java:99-160, 237-331): gridCellOffsets[totalCells+1] and gridCellContents[cap],
 * bodies ascending within a cell, terminator clamped to contentsCap (:155-158). */
int sbx_grid_build(sbx_ctx *c, const int *bodyCell, const int *counts, const int *gridDims, int *gridCellOffsets,
                   int *gridCellContents, int contentsCap);

/* SpatialGrid.broadPhase (SpatialGrid.java:340-406); candPartner[cap*maxCand], candCount[cap] (may exceed maxCand) */
int sbx_broad_phase(sbx_ctx *c, const float *center, const float *boundingRadius, const int *bodyCell,
                    const int *gridCellOffsets, const int *gridCellContents, const int *gridDims,
                    const float *gridParams, const int *counts, int *candPartner, int *candCount, int cap,
                    int maxCand, int periodic);

/* ------------------------------------------------------------------ CSR inverse and gathers */
/* CrossBridgeSystem.csrHistogram/csrScan/csrScatter and csrChunk* (CrossBridgeSystem.java:190-298): stable
 * counting sort of items by key (key < 0 skipped): offsets[nKeys+1], list[nItems] ascending item index. */
int sbx_csr_build(sbx_ctx *c, const int *key, int nItems, int nKeys, int *offsets, int *list);

/* No counterpart in the reference: sbx_csr_build with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as sbx_csr_build.  Its scratch is a fixed device array of the library (about
 * 34 MB per GPU, held until the process ends) sized for at most SBX_NOALLOC_MAX_KEYS keys and
 * SBX_NOALLOC_MAX_ITEMS items; larger inputs are rejected with SBX_ERR_ARG.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other host threads, other streams) race. */
#define SBX_NOALLOC_MAX_KEYS 4194304
#define SBX_NOALLOC_MAX_ITEMS 2097152
int sbx_csr_build_noalloc(sbx_ctx *c, const int *key, int nItems, int nKeys, int *offsets, int *list);

/* ------------------------------------------------------------------ crosslinker formation
 * V2OneXHarness.formationCpu (V2OneXHarness.java:444-464).  formParams (host, 7): CrosslinkerStore.formParams;
 * formCounts (host, 4): [1] step, [2] seed, [3] mode.  periodic / boxLen (host): the periodic extension. */
/* FilIDSystem.init + rounds(n) jumps (FilIDSystem.java:31-60; V2OneXHarness.computeFilID :466-470);
 * filState may be NULL (every slot ACTIVE) */
int sbx_filid(sbx_ctx *c, const int *filState, const int *end2NbrSlot, int *filID, int *filIDScratch, int n);
/* FormationGrid.formCandidatesCpu = CrosslinkerSystem.gridFormCount + gridFormScan + gridFormEmit
 * (CrosslinkerSystem.java:461-604) over the formation grid built with sbx_body_cell + sbx_grid_build;
 * formCountsDev (device): [0] = candidates (the request count, clamped to reqCap) */
int sbx_grid_form_candidates(sbx_ctx *c, const float *filCoord, const float *filSegLength, const int *filID,
                             const float *gridParams, const int *gridDims, const int *gridCellOffsets,
                             const int *gridCellContents, int *candCountSeg, int *candBaseSeg, int *reqFilA,
                             int *reqFilB, const float *formParams, int *formCountsDev, int nSeg, int reqCap,
                             int periodic, const float *boxLen);
/* CrosslinkerSystem.formGates (CrosslinkerSystem.java:607-676) */
int sbx_form_gates(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1, const float *filEnd2,
                   const float *filSegLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
                   int *reqOrient, int *gatePass, const float *formParams, const int *formCounts, int nSeg, int reqCap,
                   int rng, uint32_t stream, int periodic, const float *boxLen);
/* CrosslinkerSystem.formAdmitReduce (CrosslinkerSystem.java:680-691) */
int sbx_form_admit_reduce(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const int *gatePass, int *minCand,
                          int reqCap, int nSeg);
/* CrosslinkerSystem.formAdmit (CrosslinkerSystem.java:713-738) */
int sbx_form_admit(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
                   const int *gatePass, const int *minCand, const int *activeLinkCount, const int *linkState,
                   const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2, int *acceptFlag,
                   const float *formParams, int reqCap, int nSeg, int C);
/* CrosslinkerSystem.freeFlags + CrossBridgeSystem.csrScan + freeScatter + csrScan + allocate + placeOrient
 * (CrosslinkerSystem.java:298-357, 742-760); K requests, C link slots */
int sbx_form_allocate(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
                      const int *reqOrient, const int *acceptFlag, int *freeCount, int *freeOffsets, int *freeList,
                      int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1, float *loc2, int *linkState,
                      float *strainHist, int *strainPlace, int *linkOrientSame, float *torqueMagHist,
                      int *torqueMagPlace, int K, int C);

/* ------------------------------------------------------------------ actin turnover
 * Per-segment state over the C slots of a filament store (device arrays): filState (FilamentStore.filState,
 * >= 0 ACTIVE, < 0 FREE), monomerCount, seedNode (NodeNucleationStore.seedNode: the node holding a growing
 * barbed tip, < 0 none), nucFrac (3C planar ATP | ADP-Pi | ADP, AgingStore.nucFrac), cofFrac (C,
 * SeverStore.cofFrac).  The per-System entry points follow the Java methods' arguments; `fires` is the
 * cadence flag of their counts arrays, step / seed the draw keys (Wang: the reference's slot-step-seed base
 * with the salts GROW / DEPO; or Philox). */
/* AgingSystem.age (AgingSystem.java:44-64); agingParams (host, 2): pH, pD */
int sbx_aging_age(sbx_ctx *c, const int *filState, float *nucFrac, const float *agingParams, int fires, int C);
/* AgingSystem.growthAtp (AgingSystem.java:72-84) */
int sbx_aging_growth_atp(sbx_ctx *c, const int *grewFlag, const int *monomerCount, float *nucFrac, int C);
/* AgingSystem.splitInheritNuc (AgingSystem.java:94-108); K requests */
int sbx_aging_split_inherit_nuc(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                                float *nucFrac, int C, int K);
/* DepolySystem.depoly (DepolySystem.java:70-101), or depolyProxy (:115-150) when nucFrac is non-NULL;
 * depolyParams (host, 3): P, halfmono, floor; depolyRateParams (host, 2): pATP, pADP */
int sbx_depoly(sbx_ctx *c, const int *filState, int *monomerCount, float *coord, const float *uVec,
               const int *end1NbrSlot, const float *nucFrac, int *returnedMon, int *deathFlag,
               const float *depolyParams, const float *depolyRateParams, int step, int seed, int fires, int C,
               int rng, uint32_t stream);
/* DepolySystem.applyDeath (DepolySystem.java:159-187): every slot clears its own links to dying slots, so the
 * link break needs no scatter (the same result as the reference's on reciprocal chains, which every
 * turnover operation keeps) */
int sbx_apply_death(sbx_ctx *c, int *filState, int *monomerCount, int *seedNode, int *end1NbrSlot, int *end1NbrSide,
                    int *end2NbrSlot, int *end2NbrSide, float *brownTransScale, float *brownRotScale,
                    const int *deathFlag, int C);
/* GrowthSystem.grow (GrowthSystem.java:73-106); growParams (host, 3): P, halfmono, maxMon */
int sbx_growth_grow(sbx_ctx *c, const int *seedNode, int *monomerCount, float *coord, const float *uVec,
                    int *grewFlag, const float *growParams, int step, int seed, int fires, int C, int rng,
                    uint32_t stream);
/* GrowthSystem.markSplits (GrowthSystem.java:115-145); requests at stride C; splitParams (host, 3) */
int sbx_growth_mark_splits(sbx_ctx *c, const int *seedNode, const int *monomerCount, const float *coord,
                           const float *uVec, const float *yVec, int *acceptFlag, float *reqCoord, float *reqUVec,
                           float *reqYVec, const float *splitParams, int fires, int C);
/* GrowthSystem.splitWire (GrowthSystem.java:155-196); K requests */
int sbx_growth_split_wire(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                          int *monomerCount, float *coord, const float *uVec, int *end1NbrSlot, int *end1NbrSide,
                          int *end2NbrSlot, int *end2NbrSide, int *seedNode, const float *splitParams, int C, int K);
/* GrowthSystem.recomputeDrag (GrowthSystem.java:206-227); dragParams (host, 9); bTransDiff / bRotDiff may be
 * NULL */
int sbx_growth_recompute_drag(sbx_ctx *c, const int *monomerCount, float *segLength, const int *end1NbrSlot,
                              const int *end2NbrSlot, float *bTransGam, float *bRotGam, float *bTransDiff,
                              float *bRotDiff, const float *dragParams, int fires, int C);
/* SeveringSystem.cofilinAccumulate (SeveringSystem.java:37-50); cofilinParams (host, 2): pCof, ratio */
int sbx_cofilin_accumulate(sbx_ctx *c, const int *filState, const float *nucFrac, float *cofFrac,
                           const float *cofilinParams, int fires, int C);
/* SeveringSystem.cofilinDissolve (SeveringSystem.java:59-73) */
int sbx_cofilin_dissolve(sbx_ctx *c, const int *filState, const int *monomerCount, const float *cofFrac,
                         int *severDeathFlag, int *severReturnedMon, const float *cofilinParams, int fires, int C);
/* SeveringSystem.nucleateFreshCofilin (SeveringSystem.java:87-97); K requests */
int sbx_cofilin_fresh(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets, float *cofFrac,
                      int C, int K);
/* FullSystemDemoHarness.allocCpu (FullSystemDemoHarness.java:950-957): FilamentBirthSystem.freeFlags +
 * CrossBridgeSystem.csrScan + freeScatter + csrScan of the K requests + allocate (FilamentBirthSystem.java:41-86).
 * Requests at stride RC; freeCount and acceptFlag come back cleared (csrScan's cursor side effect);
 * birthParams (host, 2): the born slot's Brownian scales */
int sbx_birth_alloc(sbx_ctx *c, int *filState, int *freeCount, int *freeOffsets, int *freeList, int *acceptFlag,
                    int *rankOffsets, const float *reqCoord, const float *reqUVec, const float *reqYVec, float *coord,
                    float *uVec, float *yVec, float *brownTransScale, float *brownRotScale, const float *birthParams,
                    int C, int K, int RC);

/* CrosslinkerSystem.countActiveLinks (CrosslinkerSystem.java:694-706) */
int sbx_count_active_links(sbx_ctx *c, const int *linkState, const int *linkFilA, const int *linkFilB,
                           int *activeLinkCount, int C, int nSeg);

/* CrosslinkerSystem.unbind (CrosslinkerSystem.java:231-281); counts: [0] nLinks [1] step [2] seed */
int sbx_link_unbind(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1,
                    const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2, int *linkState,
                    float *strainHist, int *strainPlace, const float *offParams, const int *counts, int nSeg,
                    int rng, uint32_t stream);

/* CrosslinkerSystem.linkForces (CrosslinkerSystem.java:51-124) */
int sbx_link_forces(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1,
                    const float *filBTransGam, const int *linkFilA, const int *linkFilB, const float *loc1,
                    const float *loc2, const int *filLinkCt, float *xlinkData, const float *xlParams, int nSeg,
                    int nLinks);

/* CrosslinkerSystem.linkTorsion (CrosslinkerSystem.java:767-808) */
int sbx_link_torsion(sbx_ctx *c, const float *filUVec, const int *linkFilA, const int *linkFilB,
                     const int *linkState, const int *linkOrientSame, float *torqueMagHist, int *torqueMagPlace,
                     float *xlinkData, const float *torsionParams, int nSeg, int nLinks);

/* CrosslinkerSystem.segGatherA / segGatherB (CrosslinkerSystem.java:129-173); side 0 = A, 1 = B */
int sbx_seg_gather(sbx_ctx *c, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                   const int *linkState, float *filForceSum, float *filTorqueSum, int nSeg, int side);

/* ------------------------------------------------------------------ myosin side
 * Motor body: sub-body 3m rod, 3m+1 lever, 3m+2 head (MotorStore.java:45-60), planar stride nB = 3*nM.
 * Published per-motor arrays (head, uVec, rodUVec) have stride nM.  counts: [0] nM [1] step [2] seed [3] nSeg. */

/* MotorStore.publishHeadFromBody (MotorStore.java:300-314) */
int sbx_publish_head_from_body(sbx_ctx *c, const float *bodyCoord, const float *bodyUVec, const float *bodySegLength,
                               float *head, float *uVec, float *rodUVec, const int *counts, int nM);
/* MotorStore.publishToBodyView (MotorStore.java:326-347); publishParams[0] = first view slot */
int sbx_publish_motors(sbx_ctx *c, const float *head, const float *reach, float *center, float *boundingRadius,
                       int *ownerStore, int *ownerSlot, const int *publishParams, const int *counts, int nM, int cap);
/* BindingDetectionSystem.gridReachable (BindingDetectionSystem.java:143-228); kinParams[7] reach, [8] alignTol */
int sbx_grid_reachable(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                       const float *segEnd2, const float *gridParams, const int *gridDims,
                       const int *gridCellOffsets, const int *gridCellContents, const int *ownerStore,
                       const int *ownerSlot, int *motorReachSeg, int *motorReachCount, const float *kinParams,
                       const int *counts, int nSeg, int maxCand);
/* BindingDetectionSystem.bindNearest (BindingDetectionSystem.java:356-389): binding events */
int sbx_bind_nearest(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                     const float *segEnd2, const int *motorCandSeg, const int *motorCandCount, int *boundSeg,
                     float *bindArc, const float *kinParams, const int *counts, int nSeg, int maxCand);
/* BindingDetectionSystem.bindNearestNodeAware (BindingDetectionSystem.java:440-473): bindNearest skipping
 * segments held by a node (seedNode >= 0) */
int sbx_bind_nearest_node_aware(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec,
                                const float *segEnd1, const float *segEnd2, const int *motorCandSeg,
                                const int *motorCandCount, int *boundSeg, float *bindArc, const int *seedNode,
                                const float *kinParams, const int *counts, int nSeg, int maxCand);
/* BindingDetectionSystem.bruteReachable (BindingDetectionSystem.java:260-283): every segment, in index order */
int sbx_brute_reachable(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec, const float *segEnd1,
                        const float *segEnd2, int *reachSeg, int *reachCount, const float *kinParams,
                        const int *counts, int nSeg, int maxCand);
/* BindingDetectionSystem.bruteReachableNodeAware (BindingDetectionSystem.java:410-436) */
int sbx_brute_reachable_node_aware(sbx_ctx *c, const float *head, const float *uVec, const float *rodUVec,
                                   const float *segEnd1, const float *segEnd2, int *reachSeg, int *reachCount,
                                   const int *seedNode, const float *kinParams, const int *counts, int nSeg,
                                   int maxCand);
/* NucleotideCycleSystem.catchSlipRelease (NucleotideCycleSystem.java:77-135): unbinding events */
int sbx_catch_slip_release(sbx_ctx *c, int *boundSeg, const float *forceDotFil, const float *forceMag,
                           int *cooldown, int *stats, int *capStats, const float *kinParams, const int *counts,
                           int nM, int rng, uint32_t stream);
/* NucleotideCycleSystem.cycle (NucleotideCycleSystem.java:33-70) */
int sbx_nucleotide_cycle(sbx_ctx *c, int *nucleotideState, const int *boundSeg, const float *forceDotHist,
                         const float *nucParams, const int *counts, int nM, int rng, uint32_t stream);
/* CrossBridgeSystem.bondForces (CrossBridgeSystem.java:58-142); bondData stride 13 */
int sbx_bond_forces(sbx_ctx *c, const float *motorCoord, const float *motorUVec, const float *motorYVec,
                    const float *motorBRotGam, const float *filCoord, const float *filUVec, const float *filYVec,
                    const float *filBRotGam, const float *filSegLength, const int *boundSeg, const float *bindArc,
                    const int *nucleotideState, float *bondData, const float *xbParams, int nM, int nSeg);
/* CrossBridgeSystem.applyHeadForce (CrossBridgeSystem.java:145-157) */
int sbx_apply_head_force(sbx_ctx *c, const float *bondData, float *bodyForceSum, float *bodyTorqueSum, int nM);
/* CrossBridgeSystem.registerForceDot (CrossBridgeSystem.java:166-187) */
int sbx_register_force_dot(sbx_ctx *c, const float *bondData, const int *boundSeg, float *forceDotFil,
                           float *forceMag, float *forceDotHist, int *forceDotPlace, int nM);
/* CrossBridgeSystem.segGather (CrossBridgeSystem.java:302-320); CSR from sbx_csr_build keyed by boundSeg */
int sbx_motor_seg_gather(sbx_ctx *c, const int *segMotorOffsets, const int *segMotorMyo, const float *bondData,
                         float *filForceSum, float *filTorqueSum, int nSeg);
/* MotorJointSystem.joints (MotorJointSystem.java:72-198) */
int sbx_motor_joints(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const int *nucleotideState, const float *jointParams, int nB);
/* DimerCouplingSystem.couple (DimerCouplingSystem.java:83-184); needs a disjoint pairing (each motor in <= 1 dimer) */
int sbx_dimer_couple(sbx_ctx *c, const float *coord, const float *uVec, const float *segLength,
                     const float *bTransGam, const float *bRotGam, float *forceSum, float *torqueSum,
                     const int *motorA, const int *motorB, const int *parallel, const float *dimerParams,
                     const int *boundSeg, int nB, int nD);
/* MiniFilamentSystem.tether (MiniFilamentSystem.java:65-132) */
int sbx_mini_tether(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                    const float *mBTransGam, const float *mBRotGam, float *mForceSum, float *mTorqueSum,
                    const float *bbCoord, const float *bbUVec, const float *bbInvDragY, const int *headBackboneSlot,
                    const int *motorA, const float *attachAxial, float *miniData, const float *miniParams, int nMB,
                    int nBb, int nD);
/* MiniFilamentSystem.backboneGather (MiniFilamentSystem.java:137-155) */
int sbx_backbone_gather(sbx_ctx *c, const int *bbDimerOffsets, const int *bbDimerList, const float *miniData,
                        float *bbForceSum, float *bbTorqueSum, int nBb);
/* NodeSystem.tether (NodeSystem.java:48-127): attachKey = [node(nA) | motor(nA)], radial 3*nA planar body-frame
 * offsets, attachCoeffK signed (> 0 dimer: force at rod end1 with torque; < 0 singlet: at the rod centre),
 * nodeData 6*nA node-side reactions, nodeParams[0] = dt.  The per-node sum is sbx_backbone_gather over the CSR
 * of attachNode (NodeStore.java:27-33). */
int sbx_node_tether(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                    const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                    const float *nUVec, const float *nYVec, const float *nInvTransY, const int *attachKey,
                    const float *radial, const float *attachCoeffK, float *nodeData, const float *nodeParams, int nMB,
                    int nN, int nA);
/* NodeSystem.tetherBoundStiffen (NodeSystem.java:170-231): a motor whose head is bound uses nodeParams[1] */
int sbx_node_tether_bound_stiffen(sbx_ctx *c, const float *mCoord, const float *mUVec, const float *mSegLength,
                                  const float *mBTransGam, float *mForceSum, float *mTorqueSum, const float *nCoord,
                                  const float *nUVec, const float *nYVec, const float *nInvTransY,
                                  const int *attachKey, const float *radial, const float *attachCoeffK,
                                  const int *boundSeg, float *nodeData, const float *nodeParams, int nMB, int nN,
                                  int nA);
/* NodeSystem.xbridgeInject (NodeSystem.java:141-166): adds the bound heads' cross-bridge force (and its torque
 * at the world-frame surface offset) to nodeData */
int sbx_node_xbridge_inject(sbx_ctx *c, const float *bondData, const float *nUVec, const float *nYVec,
                            const int *attachKey, const float *radial, const float *attachCoeffK,
                            const int *boundSeg, float *nodeData, int nN, int nA);
/* NodeNucleationSystem.countBoundFil (NodeNucleationSystem.java:55-62): seeds held by each node */
int sbx_nuc_count_bound_fil(sbx_ctx *c, const int *seedNode, int *nodeBoundFil, int C, int nN);
/* NodeNucleationSystem.emit (NodeNucleationSystem.java:66-100): per-node seed requests at stride RC >= nN;
 * nucParams (host, 3): P_nuc, seed half-length, forminsPerNode; poolOK 0 blocks every request */
int sbx_nuc_emit(sbx_ctx *c, const float *nodeCoord, const int *nodeBoundFil, int *acceptFlag, float *reqCoord,
                 float *reqUVec, float *reqYVec, const float *nucParams, int step, int seed, int poolOK, int nN, int RC,
                 int rng, uint32_t stream);
/* NodeNucleationSystem.tagSeeds (NodeNucleationSystem.java:104-115): after sbx_birth_alloc, request r's slot
 * is held by node r */
int sbx_nuc_tag_seeds(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets, int *seedNode,
                      int C, int K);
/* NodeNucleationSystem.initNewborn (NodeNucleationSystem.java:228-244); seedParams (host, 2): monomers, length */
int sbx_nuc_init_newborn(sbx_ctx *c, const int *rankOffsets, const int *freeList, const int *freeOffsets,
                         int *monomerCount, float *segLength, const float *seedParams, int C, int K);
/* NodeNucleationSystem.seedTether (NodeNucleationSystem.java:117-149); tetherParams (host, 2): fracMove, dt */
int sbx_nuc_seed_tether(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filSegLength,
                        const float *filBTransGam, float *filForceSum, float *filTorqueSum, const float *nodeCoord,
                        const float *nodeInvTransX, const int *seedNode, const float *tetherParams, int C, int nN);
/* NodeNucleationSystem.seedTetherNodeReact (NodeNucleationSystem.java:161-187): -F of each node's held seeds,
 * in slot order, onto the node's forceSum */
int sbx_nuc_seed_tether_node_react(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filSegLength,
                                   const float *filBTransGam, float *nodeForceSum, const float *nodeCoord,
                                   const float *nodeInvTransX, const int *seedNode, const float *tetherParams, int C,
                                   int nN);
/* NodeNucleationSystem.seedAimTorque (NodeNucleationSystem.java:196-211); aimParams (host, 2): coeff, dt */
int sbx_nuc_seed_aim_torque(sbx_ctx *c, const float *filUVec, const float *filBRotGam, float *filTorqueSum,
                            const float *aimDir, const float *aimParams, int C);
/* NodeNucleationSystem.dissolve (NodeNucleationSystem.java:248-257); dissolveParams (host, 1): P_detach */
int sbx_nuc_dissolve(sbx_ctx *c, int *seedNode, const float *dissolveParams, int step, int seed, int C, int rng,
                     uint32_t stream);

/* ------------------------------------------------------------------ whole network step
 * Replaces the filament block of DenseContractileHarness.buildPlan / cpuStep (DenseContractileHarness.java:
 * 350-429, 460-606) as ONE library call per step: [publish + grid + broad phase] -> [crosslink unbind ->
 * count -> force + torsion] -> per-rod fused zero + Brownian + chain + gather A + gather B + confine +
 * integrate + derive.  The store arrays stay caller-owned device memory in the reference layout. */
typedef struct sbx_filaments {
    int n;
    float *coord, *uVec, *yVec;               /* canonical pose (updated in place; see sbx_net_publish) */
    float *zVec, *end1, *end2;                /* derived, written every step */
    float *segLength, *bTransGam, *bRotGam, *brownTransScale, *brownRotScale;
    float *forceSum, *torqueSum;              /* optional: NULL keeps the accumulated F/T in registers */
    int *end1NbrSlot, *end1NbrSide, *end2NbrSlot, *end2NbrSide;
} sbx_filaments;

typedef struct sbx_crosslinks {
    int nLinks;                                /* pool capacity C; 0 = no crosslinks */
    int *linkFilA, *linkFilB;
    float *loc1, *loc2;
    int *linkState;
    float *strainHist;                         /* C * SBX_STRAIN_WIN */
    int *strainPlace;
    int *linkOrientSame;
    float *torqueMagHist;                      /* C * SBX_TORQUE_WIN */
    int *torqueMagPlace;
    float *xlinkData;                          /* C * SBX_XL_STRIDE */
    int *activeLinkCount;                      /* nSeg */
} sbx_crosslinks;

typedef struct sbx_net_params {
    float params[2];        /* FilamentStore.params: dt, sqrt(2kT/dt) */
    float chainParams[7];   /* FilamentStore.chainParams (FilamentStore.java:83-87) */
    int chainOn;
    float xlParams[3];      /* CrosslinkerStore.xlParams: restLength, fracMoveBase, dt */
    float offParams[5];     /* CrosslinkerStore.offParams: const, coeff, exp, dt, restLength */
    float torsionParams[2]; /* CrosslinkerStore.torsionParams: spring, active */
    int unbindOn;
    float boxParams[8];     /* MechanicsFusion boxParams (MechanicsFusion.java:29-31) */
    int confineOn;
    int gridOn;             /* 1: grid + broad phase every step (per-step rebuild) */
    int periodic;
    int maxCand;            /* per-body candidate slice, 1..SBX_MAX_CAND (SpatialGrid.MAX_CAND = 256) */
    float gridParams[6];
    int gridDims[4];
    float boxLen[3];
    float actinRadius;      /* viewParams[0] of publishToBodyView */
    int seed;
    int rng;
    uint32_t stream;
    /* crosslinker formation before the step's unbinding on every step that is a multiple of formInt
     * (V2OneXHarness.java:379, formationCpu :444-464); the formation grid is FormationGrid's
     * (FormationGrid.java:63-76), periodic with the network; formReqCap candidate slots */
    int formOn;
    int formInt;
    float formParams[7];    /* CrosslinkerStore.setFormParams (CrosslinkerStore.java:196-206) */
    int formMode;           /* 0 both orientations, 1 parallel, -1 antiparallel */
    float formGridParams[6];
    int formGridDims[4];
    int formReqCap;
} sbx_net_params;

/* Myosin stage of the whole step (DenseContractileHarness.cpuStep, DenseContractileHarness.java:350-395):
 * publish heads -> grid over filaments + heads -> per motor: gridReachable + catchSlipRelease +
 * bindNearest + cycle -> per sub-body joints -> per dimer couple -> per dimer tether -> per backbone
 * gather + confine + Brownian + integrate + derive -> per motor bond + head force -> per sub-body
 * Brownian + integrate + derive (+ registerForceDot) -> filament block with the motor -> segment gather.
 * Requires the clamped grid (periodic = 0) with gridOn; the filament-pair broad phase is not run. */
typedef struct sbx_motors {
    int nM;                                    /* motors; body sub-bodies nB = 3*nM */
    float *coord, *uVec, *yVec, *zVec, *end1, *end2, *segLength, *bTransGam, *bRotGam;
    float *brownTransScale, *brownRotScale;
    float *forceSum, *torqueSum;               /* body accumulators (required) */
    float *head, *hUVec, *rodUVec, *reach;     /* published binding interface, stride nM */
    int *boundSeg, *cooldown, *stats, *capStats, *nucleotideState, *forceDotPlace;
    float *bindArc, *forceDotFil, *forceMag, *forceDotHist, *bondData;
    int *reachSeg, *reachCount;                /* nM*maxCand, nM */
    int nD;                                    /* dimers (disjoint motor pairs) */
    int *motorA, *motorB, *parallel;
    int nBb, nMD;                              /* backbones, tethered dimers */
    float *bbCoord, *bbUVec, *bbYVec, *bbZVec, *bbEnd1, *bbEnd2, *bbSegLength, *bbTransGam, *bbRotGam;
    float *bbBrownTransScale, *bbBrownRotScale, *bbInvDragY;
    float *bbForceSum, *bbTorqueSum;           /* optional: NULL keeps backbone F/T in registers */
    int *headBackboneSlot, *tetherMotor;
    float *attachAxial, *miniData;
    /* node shell (sbx_motor_params.nodeMode != 0): the backbone bodies are node spheres and each tethered item
     * is an attachment: headBackboneSlot = attachNode, tetherMotor = attachMotor, miniData = nodeData,
     * bbInvDragY = nodeInvTransY (nBb entries); attachAxial unused */
    float *radial, *attachCoeffK;
} sbx_motors;

typedef struct sbx_motor_params {
    float bodyParams[2];    /* MotorStore.bodyParams: dt, sqrt(2kT/dt) */
    float jointParams[11];  /* MotorStore.jointParams */
    float kinParams[14];    /* MotorStore.kinParams */
    float nucParams[8];     /* MotorStore.nucParams */
    float xbParams[6];      /* CrossBridgeSystem xbParams */
    float dimerParams[5];   /* DimerStore.dimerParams */
    float miniParams[3];    /* MiniFilamentStore.miniParams */
    float bbBodyParams[2];  /* MiniFilamentStore.bbBodyParams */
    int seedMotor, seedBackbone;
    int nodeMode;           /* 0 MiniFilamentSystem.tether; 1 NodeSystem.tether (miniParams[0] = dt);
                               2 NodeSystem.tetherBoundStiffen (miniParams[1] = boundCoeff) */
} sbx_motor_params;

int sbx_net_attach_motors(sbx_net *net, const sbx_motors *m, const sbx_motor_params *p);

/* ------------------------------------------------------------------ skin broad phase
 * No counterpart in the reference (SpatialGrid.broadPhase, SpatialGrid.java:340-404, rebuilds the grid
 * every step).  The candidate slices candPartner/candCount stay identical to the per-step rebuild's:
 * per-body neighbour lists are built at a rebuild with reach + skin_i + skin_j and filtered exactly every
 * step.  A body that moves more than its skin since the rebuild goes (by the step's own guard) onto an
 * outside-skin list that every body also tests, and finds its own candidates by walking the rebuild's
 * coarse grid; past extraCap every body tests every other.  Rebuilds are scheduled by the host between
 * steps (graph replays): on the first step, after sbx_net_publish, every `horizon` steps (0: never), and
 * once the latest finished step reports more than extraTrigger bodies outside their skins.  The skin
 * only changes the cost.  While a skin mode is on, the fine grid CSR (gridCellOffsets/Contents) is not
 * rebuilt; bodyCell still is. */
#define SBX_SKIN_REBUILD 0   /* the reference's per-step grid rebuild + broad phase */
#define SBX_SKIN_FIXED 1     /* every body gets skinDefault */
#define SBX_SKIN_LEARNED 2   /* per coarse cell from the skin model (sbx_mlp_*), clamped to [skinMin, skinMax] */
typedef struct sbx_skin_params {
    float skinDefault;   /* um; fixed mode, and the first build of learned mode */
    float skinMin, skinMax;
    float margin;        /* learned skin = margin * exp(model out0) */
    int listCap;         /* neighbour slots per body (an overflowing body walks the coarse grid instead) */
    int horizon;         /* steps between scheduled rebuilds (0: only the other triggers) */
    int extraTrigger;    /* bodies outside their skins (reported by the latest finished step) that schedule
                          * a rebuild */
    int extraCap;        /* capacity of the outside-skin list */
} sbx_skin_params;
/* modelWeights: device array of sbx_mlp_nparam() floats (learned mode; copied) */
int sbx_net_set_skin(sbx_net *net, int mode, const sbx_skin_params *p, const float *modelWeights);
/* counters -> out[0..cap): rebuilds, bodies outside their skins now, steps in which every body tested every
 * other (outside-skin list overflow), body-steps done by a whole block (more than 64 candidates), list entries
 * at the last build, bodies whose list overflowed at the last build, steps since the last rebuild, skin-mode
 * steps; returns the number written (<= 8) */
int sbx_net_skin_stats(sbx_net *net, long long *out, int cap);
/* device state of the last rebuild (for training): reference centres, skins, coarse cells and CSR;
 * coarseGrid[8] = xMin, yMin, zMin, cell, 1/cell, nX, nY, nZ */
int sbx_net_skin_state(sbx_net *net, float **refCoord, float **skin, int **coarseCell, int **coarseOffsets,
                       int **coarseContents, int *nCoarse, float *coarseGrid);

/* ------------------------------------------------------------------ reach skin (the binding grid of the myosin stage)
 * No counterpart in the reference: BindingDetectionSystem.gridReachable (BindingDetectionSystem.java:143-228)
 * bins every filament segment and head into the binding grid and sorts it every step.  With a reach skin the
 * segments sit in a coarse grid built at a rebuild (cell >= 0.5 L + myoColTol + skin, at least cellFloor);
 * every step each head walks the 27 coarse cells around its current position and keeps the segments that
 * are in its binding-grid stencil (the cell the per-step grid bins them in) and pass reachTestDistSq, in
 * the per-step walk's order (stencil position, then index).  A segment that moved more than its skin since
 * the rebuild is moved by the step's guard (in the filament kernel) onto a list every head also scans; if
 * that list overflows extraCap, heads scan every segment.  So reachSeg/reachCount and every binding event
 * are identical to the per-step rebuild's whatever the skins.  Rebuilds are scheduled by the host between
 * steps (graph replays): on the first step, after turnover, sbx_net_publish or a topology change, every
 * `horizon` steps (0: never), and once the latest finished step reports more than extraTrigger segments
 * outside their skins.  While it is on, the binding grid CSR (cellOffsets/cellContents) is not built. */
typedef struct sbx_reach_skin_params {
    float skinDefault;   /* um: fixed mode, and the first build of learned mode */
    float skinMin, skinMax;  /* learned skins are clamped to [skinMin, skinMax] */
    float margin;        /* learned skin = margin * exp(model out0) */
    float cellFloor;     /* smallest coarse cell, um (0: the binding grid's cell) */
    int horizon;         /* steps between scheduled rebuilds (0: only the other triggers) */
    int extraTrigger;    /* segments outside their skins (reported by the latest finished step) that
                          * schedule a rebuild */
    int extraCap;        /* capacity of the list of segments outside their skins */
} sbx_reach_skin_params;
/* mode SBX_SKIN_REBUILD turns it off; SBX_SKIN_FIXED: the coarse cell from skinDefault; SBX_SKIN_LEARNED: the
 * model predicts each (sampled) cell's skin over the previous build's cells, and the coarse cell takes the
 * smallest skin that leaves at most extraTrigger segments predicted to outrun it (modelWeights: device array of
 * sbx_mlp_nparam() floats, copied); a rebuild after turnover or a caller's edit keeps the model's last choice.
 * Every segment's skin is what the cell leaves it.  Needs motors attached. */
int sbx_net_set_reach_skin(sbx_net *net, int mode, const sbx_reach_skin_params *p, const float *modelWeights);
/* -> out[0..cap): rebuilds (device), segments outside their skins now, steps whose heads scanned every
 * segment, heads placed by the selection path (more than 4 hits on one lane or more than maxCand hits),
 * the coarse cell of the last rebuild in nm, rebuilds for: stale state, horizon, outside-skin count;
 * steps in the mode; returns the number written (<= 9) */
int sbx_net_reach_skin_stats(sbx_net *net, long long *out, int cap);

/* ------------------------------------------------------------------ skin model (MLP 12 -> 64 -> 64 -> 2, GELU)
 * No counterpart in the reference.  Parameters: one device float array of sbx_mlp_nparam() (layout in
 * softbox_amd/csrc/sbx_mlp.h); x[b*12 + k], y[b*2 + o], target like y.  Loss = mean over samples of
 * sum_o 0.5 * outWeight[o] * (y - target)^2; grad has the same layout as the parameters (the trailing
 * input-standardisation entries are not trained). */
/* training data of the skin model from a trajectory (no counterpart in the reference): per coarse cell
 * (edge cellSize) features[c*12 + k] of its members' motion over H steps (coordPrev -> coordNow); per-body running
 * max displacement from x0 and first step beyond a per-body threshold; per-cell max (mode 0, float),
 * min (mode 1, int >= 0) or root mean square (mode 2, float) over members (`fill` when none); per-cell
 * value copied to its members. */
int sbx_skin_features(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells, int n,
                      const float *coordNow, const float *coordPrev, const float *uNow, const float *uPrev,
                      const int *activeLinkCount, float H, float cellSize, int periodic,
                      const float *boxLen /* host [3] */, float *features);
int sbx_skin_track(sbx_ctx *c, const float *coord, const float *x0, int n, int periodic, const float *boxLen,
                   float *maxDisp, const float *threshold, int *breachStep, int step);
int sbx_skin_cell_reduce(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells,
                         const float *bodyFloat, const int *bodyInt, int mode, float fill, float *cellOut);
int sbx_skin_cell_scatter(sbx_ctx *c, const int *coarseOffsets, const int *coarseContents, int nCells,
                          const float *cellVal, float *bodyVal);
int sbx_mlp_nparam(void);
int sbx_mlp_forward(sbx_ctx *c, const float *weights, const float *x, int nSamples, float *y);
size_t sbx_mlp_scratch_floats(int nSamples);
int sbx_mlp_loss_grad(sbx_ctx *c, const float *weights, const float *x, const float *target, int nSamples,
                      const float *outWeight /* host [2] */, float *scratch, float *grad, float *loss /* device [1] */);
int sbx_mlp_adam(sbx_ctx *c, float *weights, const float *grad, float *m, float *v, int t, float lr, float beta1,
                 float beta2, float eps);

int sbx_net_create(sbx_ctx *c, const sbx_filaments *fil, const sbx_crosslinks *xl, const sbx_net_params *p,
                   sbx_net **out);
void sbx_net_destroy(sbx_net *net);
/* advance nSteps steps; step counters step0 .. step0+nSteps-1 (counts[1] of the reference) */
int sbx_net_step(sbx_net *net, int step0, int nSteps);
/* formation counters: out[0] passes, [1] candidates of the last pass, [2] links formed by the last pass,
 * [3] links formed in all; returns the number written (<= 4) */
int sbx_net_form_stats(sbx_net *net, long long *out, int cap);
typedef struct sbx_turnover {
    int C;                                     /* slots: the filament store's n */
    int *filState, *monomerCount, *seedNode;
    float *nucFrac;                            /* 3C */
    float *cofFrac;                            /* C */
    float *bTransDiff, *bRotDiff;              /* 3C each, or NULL (written by recomputeDrag) */
    double *pool;                              /* device [1]: free actin, uM (ActinPool.conc) */
    long long *ledger;                         /* device [2]: monomers taken, returned (ActinPool ledger) */
    int *stats;                                /* device [8]: last cadence's monomers grown, returned by depoly,
                                                  returned by dissolves, deaths, dissolves, splits; [6] cadences */
} sbx_turnover;

typedef struct sbx_turnover_params {
    int interval;                              /* steps per biochem cadence (GrowthStore.biochemCheckInt) */
    int seed;                                  /* growCounts[2] / depolyCounts[2] */
    int growthOn;                              /* growth P = onRate * conc * biochemDeltaT (GrowthStore.refreshRate) */
    int proxy;                                 /* 1: depolyProxy (nucleotide-dependent), 0: fixed-rate depoly */
    float growParams[3], splitParams[3], dragParams[9], depolyParams[3], depolyRateParams[2], agingParams[2],
        cofilinParams[2], birthParams[2];
    double onRate, biochemDeltaT, uMPerMonomer;
} sbx_turnover_params;

/* One biochem cadence of FullSystemDemoHarness.cpuStep (FullSystemDemoHarness.java:650-684; one growth pass,
 * no node nucleation) over a filament store: refreshRate, age, depoly(Proxy), csrScan + put, applyDeath,
 * grow, csrScan, growthAtp (a no-op: the scan cleared grewFlag), take, cofilinAccumulate, cofilinDissolve,
 * csrScan, applyDeath, put, markSplits, allocCpu, splitWire, splitInheritNuc, nucleateFreshCofilin,
 * recomputeDrag; the pool and the stats are updated on the device. */
int sbx_turnover_cadence(sbx_ctx *c, const sbx_filaments *fil, const sbx_turnover *t, const sbx_turnover_params *p,
                         int step, int rng, uint32_t stream);

/* Actin turnover at its biochem cadence inside sbx_net_step: on every step t with t % p->interval == 0 the
 * cadence of sbx_turnover_cadence runs on the current pose before the step (before its grid and formation:
 * FullSystemDemoHarness.cpuStep order); the formation pass then reads t->filState.  The arrays of *t stay the
 * caller's (pass t = NULL to switch turnover off).  Not with the skin broad phase or a batch. */
int sbx_net_set_turnover(sbx_net *net, const sbx_turnover *t, const sbx_turnover_params *p);
/* make the caller's coord/uVec/yVec hold the current pose (the step ping-pongs pose buffers) */
int sbx_net_publish(sbx_net *net);
/* rebuild the crosslink CSR and recount the active links after the caller changed linkFilA/linkFilB or
 * linkState between steps (formation); sbx_net_publish also recounts on the next step */
int sbx_net_topology_changed(sbx_net *net);
/* device pointers of the latest broad-phase output and grid CSR */
int sbx_net_candidates(sbx_net *net, int **candPartner, int **candCount, int **gridCellOffsets,
                       int **gridCellContents, int **bodyCell);
/* per-kernel timing with HIP events on the context stream: on=1 records every launch of sbx_net_step */
int sbx_net_set_profiling(sbx_net *net, int on);
/* accumulated [kernel count] names / total ms / launches; returns number of kernels written (<= cap) */
int sbx_net_kernel_times(sbx_net *net, const char **names, double *total_ms, long long *launches, int cap);
/* No counterpart in the reference (a measurement helper): re-runs the latest step's per-step-rebuild
 * broad phase reps times back to back on its grid, between two HIP events on the context's stream,
 * and returns the mean microseconds per launch.  The candidate slices are rewritten with the same
 * values.  Needs a step taken first; waits for all of the handle's streams. */
int sbx_net_time_broad_phase(sbx_net *net, int reps, float *usPerLaunch);
/* on = 1 (or 2): replay captured steps; 0: eager launches.  The per-step rebuild on the side streams
 * replays one captured step per pose buffer (each step's broad phase runs at the start of the next
 * replay, the call's last one after its last replay); the other forms capture 2 steps as one
 * hipGraph and run a remainder that is not a multiple of 2 eagerly.  Results equal eager launches. */
int sbx_net_use_graph(sbx_net *net, int on);
/* 1 (default): the filament-only step runs its broad phase on a second stream of the library,
 * concurrently with the crosslink and rod kernels (the candidate slices feed nothing later in the
 * step); 0: one stream.  Results are identical; sbx_net_step's work is complete on the context's
 * stream either way.  Off while per-kernel timers are on. */
int sbx_net_set_overlap(sbx_net *net, int on);
/* Launch shape of the step: broad-phase lanes per body (2, 4 or 8), the broad phase's neighbourhoods
 * staged in LDS per block (bpStaged 1, with 4 or 8 lanes) or split by stencil row over a body's lanes
 * (bpStaged 2, with 4 or 8 lanes: each lane walks whole rows; for sparse grids), scan keys per thread (4, 8 or 16: scan tiles of
 * 256 * scanItems cells) and k_rod block size (64, 128 or 256).  No counterpart in the reference; results
 * do not depend on it. */
int sbx_net_set_shape(sbx_net *net, int bpGroup, int bpStaged, int scanItems, int rodThreads);
/* Autotuner (no counterpart in the reference): times trialSteps steps (after 2 warm-up steps) for each
 * launch shape, with and without the concurrent streams (sbx_net_set_overlap), and keeps the fastest.
 * The steps are real steps step0, step0+1, ... of the run (results depend on neither): *advanced = steps
 * taken; shape[5] = {bpGroup, bpStaged, scanItems, rodThreads, overlap} and *msPerStep (may be NULL) report the
 * choice. */
int sbx_net_tune(sbx_net *net, int step0, int trialSteps, int *advanced, int *shape, float *msPerStep);
/* No counterpart in the reference (TornadoVM's task graphs have no stream capture): the rule the step's
 * captures follow.  ops = nOps triples (op: 0 record, 1 wait; stream: 0 the capturing stream, 1 the
 * broad-phase stream, 2 the sort stream; event id).  Returns -1 when every wait on an event recorded inside
 * the same capture is a join, a fork, a same-stream wait or the broad-phase stream waiting on the sort
 * stream; the sort stream waiting on the broad-phase stream's in-capture record (the form that ended the
 * host process in stream capture) returns that triple's index, as does a malformed triple.  Host-only;
 * needs no device. */
int sbx_capture_check(const int *ops, int nOps);
/* The (op, stream, event) triples of the net's last capture, as sbx_capture_check takes them: copies up to
 * cap ints into out and returns the total number of ints.  A capture that breaks the rule is refused with
 * SBX_ERR_STATE instead of being replayed. */
int sbx_net_capture_log(const sbx_net *net, int *out, int cap);
/* No counterpart in the reference: lanes per motor (4, 8, 16, 32 or 64) of the fused reach search + binding;
 * the results do not depend on it.  The per-step grid walk takes at least 16 (4 and 8 apply to the reach
 * skin's search).  sbx_net_tune picks it after the launch shape when motors are attached;
 * sbx_net_motor_group returns the current value. */
int sbx_net_set_motor_group(sbx_net *net, int lanes);
int sbx_net_motor_group(const sbx_net *net);

/* ------------------------------------------------------------------ batched step
 * No counterpart in the reference (its harnesses step one network): nRep independent networks of uneven
 * size, concatenated in HBM and advanced together, one launch per stage for all of them (and one
 * captured graph with sbx_net_use_graph on the handle of sbx_batch_net).  Rods, links and grid cells of
 * replica r start at its rodBase / linkBase / cellBase (sbx_batch_layout): rod and link ids in the
 * batch's topology, CSR and candidate slices are offset by rodBase.  Each replica is binned in its own
 * grid and draws its noise and unbinding events under its own seed and own rod / link indices, so
 * every replica's state is bit-identical to stepping it alone with sbx_net_step.  The replicas must share
 * the rod physics (params, chainParams, chainOn, maxCand, actinRadius, rng, stream) and, among those with
 * crosslinks, the link physics (xlParams, offParams, torsionParams, unbindOn); grid, box and seed may
 * differ; containment and motors are not batched.
 * The replicas' stores are copied in at creation and written back by sbx_batch_publish. */
typedef struct sbx_batch sbx_batch;
int sbx_batch_create(sbx_ctx *c, int nRep, const sbx_filaments *fil, const sbx_crosslinks *xl /* [nRep] or NULL */,
                     const sbx_net_params *p, sbx_batch **out);
void sbx_batch_destroy(sbx_batch *b);
int sbx_batch_step(sbx_batch *b, int step0, int nSteps);
/* copy every replica's pose, derived geometry, accumulated F/T and crosslink state into its own stores */
int sbx_batch_publish(sbx_batch *b);
/* the network handle of the concatenated replicas (graphs, overlap, shape, tuning, timers, candidates) */
int sbx_batch_net(sbx_batch *b, sbx_net **net);
int sbx_batch_layout(sbx_batch *b, int rep, int *rodBase, int *linkBase, int *cellBase);

#ifdef __cplusplus
}
#endif
#endif
