// sbx_network.hip -- the whole filament-network mechanics step as a short fixed kernel sequence.
//
// Replaces the filament block of DenseContractileHarness.buildPlan/cpuStep
// (DenseContractileHarness.java:350-429, 460-606), which the reference runs as ~30 TaskGraph tasks:
//   [grid]  counting sort (count, one-pass scan, scatter, place) -> k_broad_phase
//   [links] k_net_unbind (deaths also decrement both ends' active-link counts) -> k_net_link
//           (force + torsion payload; the counts are k_rod's tally of the last step less this step's deaths)
//   [rods]  k_rod: zero + Brownian + chain + gather A + gather B + confine + integrate + derive,
//           then the active-link count and the cell of the new pose (the next step's binning)
// k_rod keeps forceSum/torqueSum and the Brownian draw in registers and ping-pongs the pose buffers
// (the chain reads neighbours' pre-step pose while each rod writes its own new pose).  The step
// counter lives in device memory so a run of steps can be captured once as a hipGraph and replayed.
// k_body_cell runs only when the pose did not come from k_rod (creation, after sbx_net_publish).
// The crosslink CSR (keyed by linkFilA / linkFilB) depends only on the link topology, so it is built
// at creation and on sbx_net_topology_changed instead of every step; linkState (which unbinding
// changes every step) is applied by the gather guard, as in CrosslinkerSystem.segGatherA/B.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sbx_device.h"
#include "sbx_form.h"
#include "sbx_turnover.h"
#include "sbx_grid.h"
#include "sbx_internal.h"
#include "sbx_links.h"
#include "sbx_mlp.h"
#include "sbx_motor.h"
#include "sbx_skin.h"
#include "sbx_tiles.h"

using namespace sbx;

namespace {

struct Timer {
    const char *name;
    double ms = 0.0;
    long long launches = 0;
};

}  // namespace

struct sbx_net {
    sbx_ctx *ctx = nullptr;
    sbx_filaments fil{};
    sbx_crosslinks xl{};
    sbx_net_params p{};
    int n = 0, L = 0;
    // pose ping-pong: buf[0] = caller arrays, buf[1] = library arrays
    float *coordB[2] = {nullptr, nullptr}, *uB[2] = {nullptr, nullptr}, *yB[2] = {nullptr, nullptr};
    int cur = 0;
    float *ownPose = nullptr;
    // grid
    GridGeom geom{};
    float *boundR = nullptr;
    int *bodyCell = nullptr, *cellOffsets = nullptr, *cellContents = nullptr, *candPartner = nullptr,
        *candCount = nullptr;
    SortScratch sort;
    int derivedValid = 0;  // bodyCell and linkPack describe the current pose (written by the last k_rod)
    int countsValid = 0;   // xl.activeLinkCount matches linkState (kept by the step after a recount)
    float4 *linkPack = nullptr;  // per-rod {c, gpar} {u} {end1} for the crosslink gathers (links only)
    int *cellNext = nullptr;  // k_rod bins the new pose here; swapped with bodyCell after each step
    int *lastCell = nullptr;  // the cells the latest step's grid used (what SpatialGrid.bodyCell left)
    // links
    int *offA = nullptr, *idxA = nullptr, *offB = nullptr, *idxB = nullptr;
    LinkParams lp{};
    ChainParams cp{};
    // step counter
    int *dStep = nullptr;
    // profiling
    int profiling = 0;
    std::vector<Timer> timers;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> eventPool;
    // myosin stage
    int hasMotors = 0;
    sbx_motors mo{};
    sbx_motor_params mp{};
    int *ownerStore = nullptr, *ownerSlot = nullptr;   // body view back-pointers over cap = n + nM
    // segment -> bound heads of this step (k_bond_head): count and up to SEG_HEADS heads per segment; k_rod reads
    // and clears them (CrossBridgeSystem.csrBuild's inverse, CrossBridgeSystem.java:189-298, without the sort)
    int *segHeadCnt = nullptr, *segHeads = nullptr;
    int *bbOff = nullptr, *bbList = nullptr;           // backbone -> tethered dimers CSR (static)
    int *dimerOf = nullptr;                            // motor -> its dimer or -1 (static)
    KinParams kp{};
    NucParams np{};
    XbParams xp{};
    JointParams jp{};
    DimerParams dp{};
    MiniParams mip{};
    ReachGrid rg{};
    // skin broad phase (sbx_net_set_skin)
    int skinMode = 0;                       // SBX_SKIN_REBUILD / _FIXED / _LEARNED
    sbx_skin_params sp{};
    GridGeom *cgDev = nullptr;              // coarse rebuild grid of the last rebuild (device)
    int *cCellsDev = nullptr;               // its cell count (device)
    int cCellsMax = 0;                      // cells at the smallest coarse cell (allocation)
    float maxR = 0.0f, cellFloor = 0.0f;
    int *cCell = nullptr, *cOff = nullptr, *cCont = nullptr;  // coarse CSR of the last rebuild
    float *refC = nullptr, *refU = nullptr, *skin = nullptr;   // pose and skins at the last rebuild
    float *cellSkin = nullptr, *cellHorizon = nullptr;
    int *vlist = nullptr, *vcount = nullptr;
    int *skinFlags = nullptr;
    float *model = nullptr;                 // MLP_NPARAM device floats (learned mode)
    float4 *bodyPack[2] = {nullptr, nullptr};  // {centre, bounding radius} per pose buffer (by k_rod; grid on)
    // broad phase off the main stream, overlapping the mechanics kernels (it feeds nothing in the
    // step); per-step rebuild: the grid sort on a third stream and the grid CSR double-buffered by
    // pose buffer, so the next pose's sort runs while this pose's broad phase still reads its CSR
    hipStream_t side = nullptr, sortS = nullptr;
    // every event record and stream wait enqueued while a step is being captured, as (op, stream, event)
    // triples (stream 0 = the capturing stream, 1 = side, 2 = sort; events numbered by first use), checked
    // against the capture rule before the graph is used (capture_check)
    std::vector<int> capLog;
    std::vector<hipEvent_t> capEv;
    hipEvent_t evFork = nullptr;
    hipEvent_t evRod[2] = {nullptr, nullptr}, evPlaced[2] = {nullptr, nullptr}, evJoin[2] = {nullptr, nullptr};
    int needFork = 1;  // the side streams must be ordered after the main stream before their next work
    // 1: the latest record of some event was made inside a stream capture (a captured record stands
    // for no work scheduled outside its graph) and the main stream has seen only captures and graph
    // replays since; settle_events re-records them eagerly before any wait outside a capture
    int evCaptured = 0;
    int *csrOff[2] = {nullptr, nullptr}, *csrCont[2] = {nullptr, nullptr};  // [0] = cellOffsets/cellContents
    float4 *sortedPose[2] = {nullptr, nullptr};  // bodyPack in contents order (the sort's place pass)
    int *sortedCell[2] = {nullptr, nullptr};     // bodyCell in contents order
    int lastCsr = 0;                             // CSR buffer of the latest step's grid
    // grid histogram per pose buffer: k_rod takes counts[cell]++ and the slot as it bins the new
    // pose (per-step rebuild only); the next step's sort consumes and re-zeroes them
    int *gridCounts[2] = {nullptr, nullptr}, *gridSlot[2] = {nullptr, nullptr};
    int overlap = 1;
    // batched step (sbx_batch_create): replicas concatenated in this network's arrays
    int R = 1;
    RepInfo *reps = nullptr;                          // device table [R]
    int *rodRep = nullptr, *linkRep = nullptr;        // replica of each rod / link
    // launch shape (results do not depend on it; sbx_net_tune picks it by timing steps)
    int bpGroup = 4;      // broad-phase lanes per body (2, 4 or 8)
    int bpStaged = 0;     // 1: each block's neighbourhood staged in LDS; 2: rows split over the lanes (4 or 8)
    int rodThreads = 256; // k_rod block size (64, 128 or 256), or RODS_PAIRED
    int mbGroup = 16;     // lanes per motor in k_motor_bind (16, 32 or 64)
    // graph
    int useGraph = 0;
    int graphSteps = 0;  // non-zero while a capture is being recorded
    int graphCur = 0;    // pose buffer index the captured graph starts from
    int graphLen = 2;    // steps per captured graph (even: the pose ping-pong returns to its start)
    int graphLastCsr = 0;          // CSR buffer of the captured graph's last step
    int *graphLastCell = nullptr;  // cell buffer of the captured graph's last step
    hipGraphExec_t graphExec = nullptr;
    hipGraph_t graph = nullptr;
    // overlapped per-step rebuild: one 1-step graph per pose buffer, each = the previous step's broad
    // phase + this step's sort + this step's mechanics, three chains with no wait between them
    hipGraphExec_t unitExec[2] = {nullptr, nullptr};
    hipGraph_t unitGraph[2] = {nullptr, nullptr};
    int unitLastCsr[2] = {0, 0};
    int *unitLastCell[2] = {nullptr, nullptr};
    int bpPending = 0, bpPendingBuf = 0;  // a broad phase of CSR[buf] still to enqueue
    int csrValid[2] = {0, 0};             // CSR[b] has been built
    int joinLive[2] = {0, 0}, placedLive[2] = {0, 0};  // evJoin / evPlaced recorded in the current capture
    // crosslinker formation every formInt steps (formOn): its own sort scratch, buffers and counters
    int formOn = 0;
    FormPass form{};
    SortScratch formSort;
    int formFilIDValid = 0;     // filID follows the chain topology, which only the caller changes
    int *formInts = nullptr;    // one allocation behind the pass's int buffers
    float *formFloats = nullptr;
    long long formPasses = 0;
    // actin turnover every tp.interval steps (turnOn): the caller's state arrays and the cadence's scratch
    int turnOn = 0;
    sbx_turnover turn{};
    sbx_turnover_params tp{};
    TurnScratch turnScratch;
    // reach skin (sbx_net_set_reach_skin): gridReachable's candidates from a coarse grid of the segments
    int rsMode = 0;                          // SBX_SKIN_REBUILD (off), _FIXED or _LEARNED
    sbx_reach_skin_params rsp{};
    float rsR0 = 0.0f;                       // 0.5 L + myoColTol of the longest segment at enable (um)
    float rsFloor = 0.0f;                    // smallest coarse cell (um)
    long long rsCellsMax = 0;                // coarse cells at the smallest cell (allocation)
    float *rsRef = nullptr, *rsRefU = nullptr;  // segment centres and axes at the last rebuild
    float *rsSkin = nullptr;                 // per-segment skin (< 0: on the extra list)
    int *rsCell = nullptr, *rsOff = nullptr, *rsCont = nullptr;  // coarse cells and CSR of the last rebuild
    int *rsExtra = nullptr;                  // segments outside their skins (rsp.extraCap)
    int *rsWords = nullptr;                  // device words (RS_*)
    float *rsModel = nullptr;                // MLP_NPARAM device floats (learned mode)
    volatile int *rsHost = nullptr;          // pinned: {generation, outside-skin count} of the latest finished step,
                                             // reach skin [0..1] and filament skin [2..3]
    int *rsHostDev = nullptr;                // its device address
    int rsValid = 0;                         // the coarse grid describes the current segments
    int rsGen = 0, rsSince = 0;              // rebuilds enqueued so far; steps since the last one
    long long rsWhy[3] = {0, 0, 0};          // rebuilds for: stale state, horizon, outside-skin count
    long long rsSteps = 0;                   // steps in the mode
    // the filament skin's host schedule (sbx_net_set_skin): outside-skin list, validity, rebuild counts
    int *fsExtra = nullptr;
    int fsValid = 0, fsGen = 0, fsSince = 0;
    long long fsWhy[3] = {0, 0, 0};
    // the host's view of the outside-skin count is at most RS_INFLIGHT replays (or eager steps) old: an event
    // after each, and the host waits for the oldest before enqueueing past it
    static constexpr int RS_INFLIGHT = 2;
    hipEvent_t rsEv[RS_INFLIGHT] = {nullptr, nullptr};
    int rsEvNext = 0, rsEvLive = 0;
};

namespace {

enum {
    T_BIN = 0, T_SORT, T_PAIRS, T_UNBIND, T_COUNT, T_LINK, T_ROD, T_MBIND, T_MJOINT, T_TETHER, T_BACKBONE,
    T_BOND, T_MINTEG, T_SKIN_REBUILD, T_SKIN_PAIRS, T_RS_REBUILD, T_COUNT_TIMERS
};
const char *kTimerNames[T_COUNT_TIMERS] = {"k_body_cell",    "grid_sort",    "k_broad_phase", "k_net_unbind",
                                           "k_net_count",    "k_net_link",   "k_rod",         "k_motor_bind",
                                           "k_motor_joints", "k_net_tether", "k_backbone",    "k_bond_head",
                                           "k_motor_integ",  "skin_rebuild", "k_skin_pairs",  "reach_rebuild"};

// The bound heads of each segment for this step's motor gather (CrossBridgeSystem.csrBuild, CrossBridgeSystem.java:
// 189-298, builds a CSR by a counting sort over every segment; few heads are bound, so each bound head takes a
// slot of its segment in k_bond_head instead and k_rod orders them).
constexpr int SEG_HEADS = 8;


// CrossBridgeSystem.segGather for segment s (CrossBridgeSystem.java:302-320): the bound heads' bond forces and
// torques summed in ascending head order (the CSR's order: a stable sort keeps ascending index within a key),
// then the segment's count cleared for the next step.  More than SEG_HEADS heads: every head is scanned.
__device__ __forceinline__ void seg_heads_sum(int s, int *__restrict__ segHeadCnt, const int *__restrict__ segHeads,
                                              const int *__restrict__ boundSeg, int nM,
                                              const float *__restrict__ data, double f[3], double t[3]) {
    const int cnt = segHeadCnt[s];
    if (cnt == 0) return;
    auto add = [&](int m) {
        const float *d = data + (size_t)m * XB_STRIDE + 6;
        f[0] += d[0]; f[1] += d[1]; f[2] += d[2];
        t[0] += d[3]; t[1] += d[4]; t[2] += d[5];
    };
    if (cnt <= SEG_HEADS) {
        const int *h = segHeads + (size_t)s * SEG_HEADS;
        int prev = -1;
        for (int k = 0; k < cnt; k++) {  // the next larger head: a selection over the (few) slots
            int next = 0x7fffffff;
            for (int q = 0; q < cnt; q++) {
                const int m = h[q];
                if (m > prev && m < next) next = m;
            }
            add(next);
            prev = next;
        }
    } else {
        for (int m = 0; m < nM; m++)
            if (boundSeg[m] == s) add(m);
    }
    segHeadCnt[s] = 0;
}

// k_rod launch form selected by rodThreads == 512: 256-thread blocks, two rod groups per block, two waves
// drawing the noise of the other two (k_rod<*, true>)
constexpr int RODS_PAIRED = 512;

struct RodArgs {
    const float *coord, *uVec, *yVec;
    float *coordOut, *uOut, *yOut, *zVec, *end1, *end2;
    const float *segLength, *bTransGam, *bRotGam, *bts, *brs;
    float *forceSum, *torqueSum;
    const int *e1Slot, *e1Side, *e2Slot, *e2Side;
    const int *offA, *idxA, *offB, *idxB, *linkState;
    const float *xlinkData;
    int *segHeadCnt;           // bound heads per segment this step (k_rod clears what it reads)
    const int *segHeads;       // up to SEG_HEADS of them per segment, in any order
    const int *boundSeg;       // every head's segment (a segment with more than SEG_HEADS heads scans them)
    int nM;
    const float *bondData;
    int hasMotors;
    int *activeLinkCount;   // written from the gather's active-link tally (may be null)
    int *nextCell;          // bodyCell of the new pose for the next step's grid (null: not fused)
    int *nextCounts, *nextSlot;  // the new pose's grid histogram (null: not taken here)
    float4 *linkPack;       // packed endpoint data of the new pose for the next step's links (or null)
    const float *skinRef;         // skin guard: centre at the last rebuild (or null)
    float *skin;                  // per-body skin (< 0: on the outside-skin list)
    int *rebuildFlag;             // the outside-skin count
    int *rsExtra;                 // the outside-skin list (filament skin or reach skin)
    int rsCap;                    // its capacity
    int minImage;                 // displacements by minimum image (the filament skin in a periodic box)
    float4 *bodyPack;             // {new centre, bounding radius} for the next step's broad phase (or null)
    const float *boundR;
    GridGeom geom;
    const RepInfo *reps;  // batched step: replica table and each rod's replica (null otherwise)
    const int *rodRep;
    const int *dStep;
    int stepOff;
    int n, chainOn, hasLinks, confineOn, checkInt, rng, seed;
    uint32_t stream;
    float dt, bfm;
    double tau, halfX, halfY, halfZ, coeff;
};

// Per-rod fused step.  The force accumulation keeps the reference's exact float rounding sequence:
// forceSum = 0 -> (float)(f + chain) -> (float)(f + sumA) -> (float)(f + sumB) -> confine -> integrate.
// Rods are laid out CHAIN_SPAN per wave (rod = 63 * wave + lane - 1) so each chain joint is evaluated
// once per rod; lane 0 of a wave only supplies the joint at the end2 of the previous wave's last rod.
// BATCH: the rods of many replicas, each binned in its replica's grid and drawing its noise under the
// replica's seed and own rod index (so every replica steps exactly as it would alone).
// PAIRED: blocks of four waves for two rod groups; waves 2 and 3 draw the Brownian noise of the groups of
// waves 0 and 1 into LDS while those run the chain and the gathers, so the draw (about a third of the
// kernel's instructions) leaves the critical path of the waves that integrate.  The draw is the same
// function of (rod, step, seed), so results do not depend on the form.
template <bool BATCH, bool PAIRED>
__global__ void __launch_bounds__(256) k_rod(RodArgs a, ChainParams cp) {
    __shared__ float noiseLds[PAIRED ? 2 * 64 * 6 : 1];
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    int wave, slot = 0;
    bool drawer = false;
    if constexpr (PAIRED) {
        const int wib = threadIdx.x >> 6;
        slot = wib & 1;
        drawer = wib >= 2;
        wave = blockIdx.x * 2 + slot;
    } else {
        wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    }
    const bool waveLive = wave * CHAIN_SPAN - 1 < n;
    if constexpr (!PAIRED) {
        if (!waveLive) return;  // the whole wave is past the last rod
    }
    const int i = wave * CHAIN_SPAN + lane - 1;
    const bool live = waveLive && i >= 0 && i < n;
    const bool own = live && lane > 0;
    const int ic = live ? i : (i < 0 ? 0 : n - 1);  // in-range index for the loads of idle lanes
    const int step = *a.dStep + a.stepOff;
    uint32_t body = (uint32_t)ic;
    int seed = a.seed, cellBase = 0;
    GridGeom geom = a.geom;
    if constexpr (BATCH) {
        const RepInfo &R = a.reps[a.rodRep[ic]];
        body = (uint32_t)(ic - R.rodBase);
        seed = R.seed;
        cellBase = R.cellBase;
        geom = R.g;
    }
    float gf[3], gt[3];
    if constexpr (PAIRED) {
        if (drawer) {
            if (own) {
                brownian_noise(a.rng, body, step, seed, a.stream, gf, gt);
                float *q = noiseLds + (slot * 64 + lane) * 6;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    q[k] = gf[k];
                    q[3 + k] = gt[k];
                }
            }
            __syncthreads();
            return;
        }
    }
    Pose p;
    float gT[3], gR[3];
    float segLen = 0.0f, tS = 0.0f, rS = 0.0f;
    float f[3] = {0.0f, 0.0f, 0.0f}, t[3] = {0.0f, 0.0f, 0.0f};
    if (waveLive) {  // wave-uniform
#pragma unroll
        for (int k = 0; k < 3; k++) {
            p.c[k] = a.coord[k * n + ic];
            p.u[k] = a.uVec[k * n + ic];
            p.y[k] = a.yVec[k * n + ic];
            gT[k] = a.bTransGam[k * n + ic];
            gR[k] = a.bRotGam[k * n + ic];
        }
        segLen = a.segLength[ic]; tS = a.bts[ic]; rS = a.brs[ic];
        // the crosslink CSR bounds are loaded with the pose, so their latency runs under the chain's
        // arithmetic rather than after it
        int ra0 = 0, ra1 = 0, rb0 = 0, rb1 = 0;
        if (a.hasLinks) {
            ra0 = a.offA[ic]; ra1 = a.offA[ic + 1];
            rb0 = a.offB[ic]; rb1 = a.offB[ic + 1];
        }
        if (a.chainOn) {
            double fd[3] = {0.0, 0.0, 0.0}, td[3] = {0.0, 0.0, 0.0};
            chain_body_wave(n, ic, live, own, lane, p.c, p.u, segLen, gT, gR, a.coord, a.uVec, a.segLength,
                            a.bTransGam, a.bRotGam, a.e2Slot[ic], a.e2Side[ic], a.e1Slot[ic], a.e1Side[ic], cp, fd,
                            td);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                f[k] = (float)(f[k] + fd[k]);
                t[k] = (float)(t[k] + td[k]);
            }
        }
        if (own) {
            if (a.hasLinks) {
                double fa[3] = {0, 0, 0}, ta[3] = {0, 0, 0};
                int active = 0;
                gather_range(ra0, ra1, a.idxA, a.xlinkData, a.linkState, 0, fa, ta, active);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    f[k] = (float)(f[k] + fa[k]);
                    t[k] = (float)(t[k] + ta[k]);
                }
                double fb[3] = {0, 0, 0}, tb[3] = {0, 0, 0};
                gather_range(rb0, rb1, a.idxB, a.xlinkData, a.linkState, 6, fb, tb, active);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    f[k] = (float)(f[k] + fb[k]);
                    t[k] = (float)(t[k] + tb[k]);
                }
                if (a.activeLinkCount) a.activeLinkCount[i] = active;
            }
            if (a.hasMotors) {  // CrossBridgeSystem.segGather (CrossBridgeSystem.java:302-320)
                double fm[3] = {0, 0, 0}, tm[3] = {0, 0, 0};
                seg_heads_sum(i, a.segHeadCnt, a.segHeads, a.boundSeg, a.nM, a.bondData, fm, tm);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    f[k] = (float)(f[k] + fm[k]);
                    t[k] = (float)(t[k] + tm[k]);
                }
            }
            if (a.confineOn && !(a.checkInt > 1 && (step % a.checkInt) != 0))
                confine_body(p.c, p.u, segLen, gT[0], a.tau, a.halfX, a.halfY, a.halfZ, a.coeff, f, t);
            if (a.forceSum) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    a.forceSum[k * n + i] = f[k];
                    a.torqueSum[k * n + i] = t[k];
                }
            }
        }
    }
    if constexpr (PAIRED) {
        __syncthreads();
        if (!own) return;
        const float *q = noiseLds + (slot * 64 + lane) * 6;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            gf[k] = q[k];
            gt[k] = q[3 + k];
        }
    } else {
        if (!own) return;
        brownian_noise(a.rng, body, step, seed, a.stream, gf, gt);
    }
    float rf[3], rt[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        rf[k] = brown_amp(tS, a.bfm, gT[k], gf[k]);
        rt[k] = brown_amp(rS, a.bfm, gR[k], gt[k]);
    }
    Derived d;
    integrate_derive(p, f, t, rf, rt, gT, gR, a.dt, segLen, d);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a.coordOut[k * n + i] = p.c[k];
        a.uOut[k * n + i] = p.u[k];
        a.yOut[k * n + i] = p.y[k];
        a.zVec[k * n + i] = d.z[k];
        a.end1[k * n + i] = d.e1[k];
        a.end2[k * n + i] = d.e2[k];
    }
    if (a.nextCell) {
        const int cell = cellBase + cell_of(p.c[0], p.c[1], p.c[2], geom);
        a.nextCell[i] = cell;
        if (a.nextCounts) a.nextSlot[i] = atomicAdd(&a.nextCounts[cell], 1);
    }
    if (a.skinRef) {  // exact displacement guard: a body beyond its skin goes onto the outside-skin list
        float dx = p.c[0] - a.skinRef[i], dy = p.c[1] - a.skinRef[n + i], dz = p.c[2] - a.skinRef[2 * n + i];
        if (a.minImage) {  // (the reach skin takes raw coordinates: its test and grids use them unwrapped)
            dx = min_image(dx, a.geom.Lx, a.geom.iLx);
            dy = min_image(dy, a.geom.Ly, a.geom.iLy);
            dz = min_image(dz, a.geom.Lz, a.geom.iLz);
        }
        const float sk = a.skin[i];
        if (sk >= 0.0f && dx * dx + dy * dy + dz * dz > sk * sk) {  // first time out: onto the list
            a.skin[i] = -1.0f;
            const int e = atomicAdd(a.rebuildFlag, 1);
            if (e < a.rsCap) a.rsExtra[e] = i;
        }
    }
    if (a.bodyPack) a.bodyPack[i] = make_float4(p.c[0], p.c[1], p.c[2], a.boundR[i]);
    if (a.linkPack) {
        float4 *q = a.linkPack + LINK_PACK * i;
        q[0] = make_float4(p.c[0], p.c[1], p.c[2], gT[0]);
        q[1] = make_float4(p.u[0], p.u[1], p.u[2], 0.0f);
        q[2] = make_float4(d.e1[0], d.e1[1], d.e1[2], 0.0f);
    }
}

// cells of the caller's pose in a batch: rod i in the grid of its replica
__global__ void __launch_bounds__(256) k_body_cell_batch(const float *__restrict__ coord, const RepInfo *__restrict__ reps,
                                                       const int *__restrict__ rodRep, int *__restrict__ bodyCell,
                                                       int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const RepInfo &R = reps[rodRep[i]];
    bodyCell[i] = R.cellBase + cell_of(coord[i], coord[n + i], coord[2 * n + i], R.g);
}

// bodyPack of the caller's pose (before the first step and after sbx_net_publish)
__global__ void __launch_bounds__(256) k_pack_bodies(const float *__restrict__ coord, const float *__restrict__ boundR,
                                                   float4 *__restrict__ pack, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) pack[i] = make_float4(coord[i], coord[n + i], coord[2 * n + i], boundR[i]);
}

// linkPack of the caller's pose (before the first step and after sbx_net_publish)
__global__ void __launch_bounds__(256) k_pack_links(const float *__restrict__ coord, const float *__restrict__ uVec,
                                                    const float *__restrict__ end1,
                                                    const float *__restrict__ bTransGam, float4 *__restrict__ pack,
                                                    int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 *q = pack + LINK_PACK * i;
    q[0] = make_float4(coord[i], coord[n + i], coord[2 * n + i], bTransGam[i]);
    q[1] = make_float4(uVec[i], uVec[n + i], uVec[2 * n + i], 0.0f);
    q[2] = make_float4(end1[i], end1[n + i], end1[2 * n + i], 0.0f);
}

// BATCH: each link draws under its replica's seed and own link index
template <bool BATCH>
__global__ void __launch_bounds__(256) k_net_unbind(const float4 *__restrict__ pack,
                                                    const int *__restrict__ linkFilA,
                                                    const int *__restrict__ linkFilB, const float *__restrict__ loc1,
                                                    const float *__restrict__ loc2, int *__restrict__ linkState,
                                                    float *__restrict__ strainHist, int *__restrict__ strainPlace,
                                                    LinkParams lp, const int *__restrict__ dStep, int stepOff,
                                                    int seed, int rng, uint32_t stream, int *__restrict__ activeLinks,
                                                    const RepInfo *__restrict__ reps,
                                                    const int *__restrict__ linkRep, int nSeg, int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    int st = linkState[k];
    if (st < 0) return;
    int step = *dStep + stepOff;
    const int a = linkFilA[k], b = linkFilB[k];
    LinkEnd A = link_end_packed(a, pack), B = link_end_packed(b, pack);
    uint32_t key = (uint32_t)k;
    if constexpr (BATCH) {
        const RepInfo &R = reps[linkRep[k]];
        key = (uint32_t)(k - R.linkBase);
        seed = R.seed;
    }
    int ns = link_unbind_one(k, key, A, B, loc1[k], loc2[k], strainHist, strainPlace, lp, step, seed, rng, stream,
                             st);
    if (ns != st) {
        linkState[k] = ns;
        if (ns < 0) {  // death: the link no longer counts at either end (CSR A and CSR B each held it)
            atomicSub(&activeLinks[a], 1);
            atomicSub(&activeLinks[b], 1);
        }
    }
}

// CrosslinkerSystem.countActiveLinks (CrosslinkerSystem.java:694-706) for every segment: the
// active-link counts the step keeps from then on (k_net_unbind decrements, k_rod re-tallies)
__global__ void __launch_bounds__(256) k_link_counts(const int *__restrict__ offA, const int *__restrict__ idxA,
                                                     const int *__restrict__ offB, const int *__restrict__ idxB,
                                                     const int *__restrict__ linkState, int *__restrict__ activeLinks,
                                                     int n) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < n) activeLinks[s] = active_links(s, offA, idxA, offB, idxB, linkState);
}

// linkForces + linkTorsion for one link (CrosslinkerSystem.java:62-123, 96-128)
__global__ void __launch_bounds__(256) k_net_link(const float4 *__restrict__ pack,
                                                  const int *__restrict__ linkFilA, const int *__restrict__ linkFilB,
                                                  const float *__restrict__ loc1, const float *__restrict__ loc2,
                                                  const int *__restrict__ activeLinks,
                                                  const int *__restrict__ linkState,
                                                  const int *__restrict__ linkOrientSame,
                                                  float *__restrict__ torqueMagHist, int *__restrict__ torqueMagPlace,
                                                  float *__restrict__ xlinkData, LinkParams lp, int nSeg, int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    float pay[12];
#pragma unroll
    for (int j = 0; j < 12; j++) pay[j] = 0.0f;
    int a = linkFilA[k], b = linkFilB[k];
    if (a >= 0 && b >= 0) {
        const int ctA = activeLinks[a], ctB = activeLinks[b];
        LinkEnd A = link_end_packed(a, pack), B = link_end_packed(b, pack);
        link_force_one(A, B, loc1[k], loc2[k], ctA, ctB, lp, pay);
        if (lp.torsionOn && linkState[k] >= 0)
            link_torsion_one(k, A, B, linkOrientSame[k] != 0, torqueMagHist, torqueMagPlace, lp, pay);
    }
    float *d = xlinkData + (size_t)k * XL_STRIDE;
#pragma unroll
    for (int j = 0; j < 12; j++) d[j] = pay[j];
}

__global__ void k_set_step(int *dStep, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dStep = v;
}
__global__ void k_add_step(int *dStep, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dStep += v;
}

__global__ void __launch_bounds__(256) k_bounding_radius(const float *__restrict__ segLength, float actinRadius,
                                                         float *__restrict__ boundR, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) boundR[i] = 0.5f * segLength[i] + actinRadius;  // FilamentStore.java:291
}

// ---------------------------------------------------------------- myosin stage kernels
// body cell over the view: filaments [0, n) bin their centre, motors [n, n + nM) their head tip
__global__ void __launch_bounds__(256) k_body_cell_view(const float *__restrict__ coord, int n,
                                                        const float *__restrict__ head, int nM, GridGeom g,
                                                        int *__restrict__ bodyCell) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n + nM) return;
    if (i < n) bodyCell[i] = cell_of(coord[i], coord[n + i], coord[2 * n + i], g);
    else {
        int m = i - n;
        bodyCell[i] = cell_of(head[m], head[nM + m], head[2 * nM + m], g);
    }
}

__global__ void __launch_bounds__(256) k_owner_view(int *__restrict__ ownerStore, int *__restrict__ ownerSlot, int n,
                                                    int nM) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n + nM) return;
    ownerStore[i] = i < n ? 0 : 1;
    ownerSlot[i] = i < n ? i : i - n;
}

struct MotorArgs {
    sbx_motors m;
    const float *e1, *e2;          // filament end1/end2 (derived last step)
    const int *off, *cont, *ownerStore, *ownerSlot;
    ReachGrid rg;
    KinParams kp;
    NucParams np;
    int maxCand, nSeg, seed, rng;
    uint32_t stream;
    const int *dStep;
    int stepOff;
};

// publishHeadFromBody (MotorStore.java:300-314) for the grid + binding of this step
__global__ void __launch_bounds__(256) k_pub_heads(sbx_motors m) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= m.nM) return;
    int nM = m.nM, nB = 3 * nM, h = 3 * k + 2, r = 3 * k;
    float hl = m.segLength[h];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        m.head[a * nM + k] = m.coord[a * nB + h] + 0.5f * hl * m.uVec[a * nB + h];
        m.hUVec[a * nM + k] = m.uVec[a * nB + h];
        m.rodUVec[a * nM + k] = m.uVec[a * nB + r];
    }
}

// gridReachable + catchSlipRelease + bindNearest + cycle for one motor (each reads only its own kinetic
// state and the step-start filament geometry, so the per-motor fusion equals the reference's four passes).
// A group of MB_GROUP lanes takes one motor: its reach stencil (BindingDetectionSystem.gridReachable,
// BindingDetectionSystem.java:157-227: the 3x3x3 cells around the head's cell, clamped at the grid's
// faces, (z, y, x) order, ascending index within a cell) is mapped to contiguous positions of the
// cell-sorted contents, and the lanes take consecutive positions.  A hit's place in reachSeg is the
// group's count so far plus the hits at earlier positions of the round (ballot), so the list is the
// reference's in its order; bindNearest (:367-387) becomes an argmin over (distance, list place) of the
// first maxCand hits, which is the serial scan's strict-< winner.  The group size G (16, 32 or 64 lanes,
// sbx_net_set_motor_group) trades lanes for rounds: few motors over crowded stencils (heads clustered on
// nodes) want wide groups, many motors over sparse stencils narrow ones; results do not depend on it.
constexpr int MB_ROUND = 2;

// The rest of the fused motor pass, on one lane per motor: reachCount, catchSlipRelease, bindNearest's
// winner (bestSeg, the first of the minimal distances in list order; -1: no hit), the nucleotide cycle.
__device__ __forceinline__ void motor_kinetics(const MotorArgs &a, int k, int cnt, int bestSeg, float mx, float my,
                                               float mz) {
    const sbx_motors &m = a.m;
    const int nSeg = a.nSeg;
    const int step = *a.dStep + a.stepOff;
    m.reachCount[k] = cnt;
    int bs = m.boundSeg[k], cool = m.cooldown[k];
    catch_slip_one(k, bs, cool, m.stats, m.capStats, m.forceDotFil[k], m.forceMag[k], a.kp, step, a.seed, a.rng,
                   a.stream);
    if (bs == FREE_BINDABLE && bestSeg >= 0) {
        const int sg = bestSeg;
        m.bindArc[k] = bind_arc(mx, my, mz, a.e1[sg], a.e1[nSeg + sg], a.e1[2 * nSeg + sg], a.e2[sg],
                                a.e2[nSeg + sg], a.e2[2 * nSeg + sg]);
        bs = sg;
    }
    m.boundSeg[k] = bs;
    m.cooldown[k] = cool;
    m.nucleotideState[k] = nucleotide_one(k, m.nucleotideState[k], bs >= 0, m.forceDotHist, a.np, step, a.seed,
                                          a.rng, a.stream);
}

template <int G>
__global__ void __launch_bounds__(256) k_motor_bind(MotorArgs a) {
    constexpr int U = MB_ROUND;
    const int lane = threadIdx.x & 63, gl = lane % G, grp = lane / G;
    const int k = (blockIdx.x * 256 + threadIdx.x) / G;
    const sbx_motors &m = a.m;
    if (k >= m.nM) return;  // whole groups leave together (k is uniform in a group)
    const int nM = m.nM, nSeg = a.nSeg;
    const float mx = m.head[k], my = m.head[nM + k], mz = m.head[2 * nM + k];
    const float mux = m.hUVec[k], muy = m.hUVec[nM + k], muz = m.hUVec[2 * nM + k];
    const float rux = m.rodUVec[k], ruy = m.rodUVec[nM + k], ruz = m.rodUVec[2 * nM + k];
    const ReachGrid &rg = a.rg;
    const int cx = clamp_bin((mx - rg.xMin) * rg.invCell, rg.nX);
    const int cy = clamp_bin((my - rg.yMin) * rg.invCell, rg.nY);
    const int cz = clamp_bin((mz - rg.zMin) * rg.invCell, rg.nZ);
    GridGeom g{};
    g.nX = rg.nX; g.nY = rg.nY; g.nZ = rg.nZ;
    g.periodic = 0;  // the reach stencil is clamped at the faces, periodic or not
    int pre[18], shift[18];
    const int total = stencil_map(cx + cy * rg.nX + cz * rg.nX * rg.nY, g, a.off, pre, shift);
    const unsigned long long groupBits = (G == 64 ? ~0ull : ((1ull << G) - 1)) << (grp * G);
    const unsigned long long below = (1ull << lane) - 1;
    int *slice = m.reachSeg + (size_t)k * a.maxCand;
    const float tol = a.kp.myoColTol, align = a.kp.alignTol;
    int cnt = 0, bestPos = 0x7fffffff, bestSeg = -1;
    float bestD = 1.0e30f;
    for (int base = 0; base < total; base += G * U) {
        int seg[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = base + u * G + gl;
            const int j = q < total ? a.cont[stencil_at9(pre, shift, q)] : -1;
            // the step's body view (k_owner_view): filament segments are bodies [0, nSeg) with ownerSlot = j,
            // heads come after; so ownerStore[j] == 0 is j < nSeg and needs no load
            seg[u] = (j >= 0 && j < nSeg) ? j : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            float d = -1.0f;
            const int sg = seg[u];
            if (sg >= 0)
                d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, a.e1[sg], a.e1[nSeg + sg],
                               a.e1[2 * nSeg + sg], a.e2[sg], a.e2[nSeg + sg], a.e2[2 * nSeg + sg], tol, align);
            const bool hit = d >= 0.0f;
            const unsigned long long bits = __ballot(hit) & groupBits;
            if (hit) {
                const int slot = cnt + __popcll(bits & below);
                if (slot < a.maxCand) {
                    slice[slot] = sg;
                    if (d < bestD) { bestD = d; bestPos = slot; bestSeg = sg; }
                }
            }
            cnt += __popcll(bits);
        }
    }
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
        const float od = __shfl_xor(bestD, o, 64);
        const int op = __shfl_xor(bestPos, o, 64), os = __shfl_xor(bestSeg, o, 64);
        if (od < bestD || (od == bestD && op < bestPos)) { bestD = od; bestPos = op; bestSeg = os; }
    }
    if (gl != 0) return;
    motor_kinetics(a, k, cnt, bestSeg, mx, my, mz);
}

// ---------------------------------------------------------------- reach skin (sbx_net_set_reach_skin)
// No counterpart in the reference, which bins every segment and head into the binding grid and sorts it
// every step for gridReachable (BindingDetectionSystem.java:157-227).  A hit (head h, segment s) has
// |h - c_s| < 0.5 L_s + myoColTol: the contact point lies on the segment and within myoColTol of the head.
// So a segment kept in a coarse grid at its centre c0 of the last rebuild, with the skin
// sk = cell - (0.5 L_s + myoColTol) - 2 slack, lies within one coarse cell of every head it can reach while
// |c_s - c0| <= sk: the 27 coarse cells around the head's current cell hold all of them (clamping at the
// faces keeps neighbouring cells neighbours).  Segments past their skin (the guard in k_rod) or too long
// for the cell are on the outside-skin list, which every head scans too; past its capacity heads scan every
// segment.  A candidate is a hit if its binding-grid cell (the per-step binning of its current centre) lies
// in the head's clamped stencil and it passes reachTestDistSq; its place in the per-step list is its key
// (stencil position << 26 | index), so hits are ranked by key and bindNearest is the argmin over
// (distance, key) of the first maxCand of them.
constexpr int RS_EXTRA = 0;      // segments outside their skins since the last rebuild (> extraCap: overflow)
constexpr int RS_REBUILDS = 1;   // rebuilds so far
constexpr int RS_MAXSKIN = 2;    // learned: 1 once the model has run for the rebuild being built
constexpr int RS_GEN = 3;        // generation of the last rebuild (the host's count)
constexpr int RS_BRUTE = 4;      // steps whose heads scanned every segment (outside-skin list overflow)
constexpr int RS_SELECT = 5;     // heads placed by the selection path
constexpr int RS_LASTSKIN = 6;   // learned: the skin the model's last evaluation chose (float bits)
constexpr int RS_GEOM = 8;       // 8 words: xMin, yMin, zMin, cell, invCell (float bits), nX, nY, nZ
constexpr int RS_HIST = 16;      // learned mode: segments per predicted-skin bin (log bins over [skinMin, skinMax])
constexpr int RS_HBINS = 64;
constexpr int RS_WORDS = RS_HIST + RS_HBINS;
constexpr int RS_PER_LANE = 2;   // hits a lane holds in registers (more on any lane: the selection path)
constexpr int RS_KEY_BITS = 26;
constexpr unsigned RS_KEY_MASK = (1u << RS_KEY_BITS) - 1u;
constexpr int RS_MODEL_CELLS = 16384;  // coarse cells the model evaluates per rebuild (at most about)
constexpr float RS_SLACK = 1.0e-3f;  // um, covers float rounding of the reach bound, the cells and the guard

struct ReachArgs {
    const float *coord;    // step-start segment centres (planar, stride nSeg)
    const float *skin;     // per-segment skin (< 0: on the outside-skin list)
    const int *off, *cont; // coarse CSR of the last rebuild
    const int *extra;      // the outside-skin list
    int *words;            // RS_* words
    int *hostWords;        // pinned {generation, outside-skin count} (device address)
    int extraCap;
    GridGeom fine;         // the binding grid as the per-step path bins the segments (k_body_cell_view)
};

// gridReachable + catchSlipRelease + bindNearest + cycle per motor from the coarse grid (G lanes per head).
template <int G>
__global__ void __launch_bounds__(256) k_motor_bind_skin(MotorArgs a, ReachArgs r) {
    constexpr int U = MB_ROUND;
    const int lane = threadIdx.x & 63, gl = lane % G, grp = lane / G;
    const int k = (blockIdx.x * 256 + threadIdx.x) / G;
    const sbx_motors &m = a.m;
    const int nExtra = r.words[RS_EXTRA];
    const bool brute = nExtra > r.extraCap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the host's (lagged) view of the guard
        r.hostWords[1] = nExtra;
        r.hostWords[0] = r.words[RS_GEN];
        if (brute) r.words[RS_BRUTE] += 1;
    }
    if (k >= m.nM) return;  // whole groups leave together
    const int nM = m.nM, nSeg = a.nSeg;
    // publishHeadFromBody (MotorStore.java:300-314), which k_pub_heads runs for the per-step path
    const int nB = 3 * nM, hb = 3 * k + 2, rb = 3 * k;
    const float hl = m.segLength[hb];
    const float mux = m.uVec[hb], muy = m.uVec[nB + hb], muz = m.uVec[2 * nB + hb];
    const float mx = m.coord[hb] + 0.5f * hl * mux, my = m.coord[nB + hb] + 0.5f * hl * muy,
                mz = m.coord[2 * nB + hb] + 0.5f * hl * muz;
    const float rux = m.uVec[rb], ruy = m.uVec[nB + rb], ruz = m.uVec[2 * nB + rb];
    if (gl == 0) {
        m.head[k] = mx; m.head[nM + k] = my; m.head[2 * nM + k] = mz;
        m.hUVec[k] = mux; m.hUVec[nM + k] = muy; m.hUVec[2 * nM + k] = muz;
        m.rodUVec[k] = rux; m.rodUVec[nM + k] = ruy; m.rodUVec[2 * nM + k] = ruz;
    }
    const ReachGrid &rg = a.rg;
    // the head's binding-grid stencil, as the per-step walk clamps it
    const int cx = clamp_bin((mx - rg.xMin) * rg.invCell, rg.nX);
    const int cy = clamp_bin((my - rg.yMin) * rg.invCell, rg.nY);
    const int cz = clamp_bin((mz - rg.zMin) * rg.invCell, rg.nZ);
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < rg.nX ? cx + 1 : rg.nX - 1;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < rg.nY ? cy + 1 : rg.nY - 1;
    const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz + 1 < rg.nZ ? cz + 1 : rg.nZ - 1;
    // the coarse cells around the head's current position
    GridGeom cg{};
    cg.nX = r.words[RS_GEOM + 5];
    cg.nY = r.words[RS_GEOM + 6];
    cg.nZ = r.words[RS_GEOM + 7];
    cg.periodic = 0;
    const float cInv = __int_as_float(r.words[RS_GEOM + 4]);
    const int ccx = clamp_bin((mx - __int_as_float(r.words[RS_GEOM])) * cInv, cg.nX);
    const int ccy = clamp_bin((my - __int_as_float(r.words[RS_GEOM + 1])) * cInv, cg.nY);
    const int ccz = clamp_bin((mz - __int_as_float(r.words[RS_GEOM + 2])) * cInv, cg.nZ);
    int pre[18], shift[18];
    int nGrid = 0;
    if (!brute) nGrid = stencil_map(ccx + ccy * cg.nX + ccz * cg.nX * cg.nY, cg, r.off, pre, shift);
    const int total = brute ? nSeg : nGrid + nExtra;
    const float tol = a.kp.myoColTol, align = a.kp.alignTol;
    // q-th candidate, or -1
    auto cand = [&](int q) -> int {
        if (q >= total) return -1;
        if (brute) return q;
        if (q < nGrid) return r.cont[stencil_at9(pre, shift, q)];
        return r.extra[q - nGrid];
    };
    // reach distance^2 of candidate q's segment s (< 0: no hit) and its key in the per-step list.  A grid entry
    // that is on the outside-skin list is skipped (the list supplies it); the skin and both ends load together,
    // the centre (for the binding-grid cell) only for a reach hit.
    auto test = [&](int q, int s, unsigned &key) -> float {
        const float sk = (!brute && q < nGrid) ? r.skin[s] : 0.0f;
        const float e1x = a.e1[s], e1y = a.e1[nSeg + s], e1z = a.e1[2 * nSeg + s];
        const float e2x = a.e2[s], e2y = a.e2[nSeg + s], e2z = a.e2[2 * nSeg + s];
        if (sk < 0.0f) return -1.0f;
        const float d = reach_test(mx, my, mz, mux, muy, muz, rux, ruy, ruz, e1x, e1y, e1z, e2x, e2y, e2z, tol, align);
        if (d < 0.0f) return -1.0f;
        const int3 b = cell3_of(r.coord[s], r.coord[nSeg + s], r.coord[2 * nSeg + s], r.fine);
        if (b.x < x0 || b.x > x1 || b.y < y0 || b.y > y1 || b.z < z0 || b.z > z1) return -1.0f;
        key = ((unsigned)((b.z - z0) * 9 + (b.y - y0) * 3 + (b.x - x0)) << RS_KEY_BITS) | (unsigned)s;
        return d;
    };
    const unsigned long long groupBits = (G == 64 ? ~0ull : ((1ull << G) - 1)) << (grp * G);
    unsigned hk[RS_PER_LANE];
    float hd[RS_PER_LANE];
#pragma unroll
    for (int q = 0; q < RS_PER_LANE; q++) {
        hk[q] = 0xffffffffu;
        hd[q] = 1.0e30f;
    }
    int nh = 0, cnt = 0;
    for (int base = 0; base < total; base += G * U) {
        int sg[U];
#pragma unroll
        for (int u = 0; u < U; u++) sg[u] = cand(base + u * G + gl);
#pragma unroll
        for (int u = 0; u < U; u++) {
            unsigned key = 0;
            float d = -1.0f;
            if (sg[u] >= 0) d = test(base + u * G + gl, sg[u], key);
            const bool hit = d >= 0.0f;
            cnt += __popcll(__ballot(hit) & groupBits);
            if (hit) {
#pragma unroll
                for (int q = 0; q < RS_PER_LANE; q++)
                    if (q == nh) {
                        hk[q] = key;
                        hd[q] = d;
                    }
                nh++;
            }
        }
    }
    int *slice = m.reachSeg + (size_t)k * a.maxCand;
    const bool spill = (__ballot(nh > RS_PER_LANE) & groupBits) != 0;
    unsigned bestK = 0xffffffffu;
    float bestD = 1.0e30f;
    if (!spill && cnt <= a.maxCand) {
        if (cnt > 0) {
            // a hit's place = the group's hits with a smaller key (keys are distinct)
            int rank[RS_PER_LANE] = {};
            const int gbase = grp * G;
            for (int j = 0; j < G; j++) {
                const int nj = __shfl(nh, gbase + j, 64);
#pragma unroll
                for (int p = 0; p < RS_PER_LANE; p++) {
                    const unsigned kj = __shfl(hk[p], gbase + j, 64);
#pragma unroll
                    for (int q = 0; q < RS_PER_LANE; q++) rank[q] += (p < nj && kj < hk[q]) ? 1 : 0;
                }
            }
#pragma unroll
            for (int q = 0; q < RS_PER_LANE; q++)
                if (q < nh) {
                    slice[rank[q]] = (int)(hk[q] & RS_KEY_MASK);
                    if (hd[q] < bestD || (hd[q] == bestD && hk[q] < bestK)) {
                        bestD = hd[q];
                        bestK = hk[q];
                    }
                }
#pragma unroll
            for (int o = G / 2; o >= 1; o >>= 1) {
                const float od = __shfl_xor(bestD, o, 64);
                const unsigned ok = __shfl_xor(bestK, o, 64);
                if (od < bestD || (od == bestD && ok < bestK)) {
                    bestD = od;
                    bestK = ok;
                }
            }
        }
    } else {
        // selection: the keep smallest keys one per pass over the candidates (dense heads only)
        const int keep = cnt < a.maxCand ? cnt : a.maxCand;
        unsigned lo = 0;
        for (int p = 0; p < keep; p++) {
            unsigned mk = 0xffffffffu;
            float md = 1.0e30f;
            for (int base = 0; base < total; base += G) {
                const int s = cand(base + gl);
                if (s >= 0) {
                    unsigned key = 0;
                    const float d = test(base + gl, s, key);
                    if (d >= 0.0f && key >= lo && key < mk) {
                        mk = key;
                        md = d;
                    }
                }
            }
#pragma unroll
            for (int o = G / 2; o >= 1; o >>= 1) {
                const unsigned ok = __shfl_xor(mk, o, 64);
                const float od = __shfl_xor(md, o, 64);
                if (ok < mk) {
                    mk = ok;
                    md = od;
                }
            }
            if (gl == 0) slice[p] = (int)(mk & RS_KEY_MASK);
            if (md < bestD) {  // list order is key order: strict < keeps the first of equal distances
                bestD = md;
                bestK = mk;
            }
            lo = mk + 1u;
        }
        if (gl == 0) atomicAdd(&r.words[RS_SELECT], 1);
    }
    if (gl != 0) return;
    motor_kinetics(a, k, cnt, bestK != 0xffffffffu ? (int)(bestK & RS_KEY_MASK) : -1, mx, my, mz);
}

// A rebuild (enqueued by the host between steps): [learned: per-cell model skins] -> geometry -> bin -> sort.
struct ReachBuild {
    int n, extraCap, learned, gen, sampleStride;
    int keepLearned;                 // learned mode, model not run: keep its last skin
    int trigger;                     // the outside-skin count that schedules a rebuild
    float since;                     // steps since the previous build (learned features)
    const float *coord, *uVec, *segLength;
    const int *links;                // activeLinkCount (learned features; may be null)
    float *ref, *refU, *skin;
    int *cell, *extra, *words;
    const int *off, *cont;           // the previous build's CSR (learned features)
    const float *model;
    float tol, R0, floorCell, skinDefault, skinMin, skinMax, margin;
    float ox, oy, oz, spanX, spanY, spanZ;  // the binding grid's origin and extent
    long long cellsMax;
};

// Learned mode: the model's skin for occupied coarse cells of the previous build (features of their members'
// motion since that build), tallied per skin bin with the cell's member count.  Thread t evaluates the cell
// that segment t * sampleStride leads (is first in), so a rebuild evaluates at most about n / sampleStride
// cells whatever the grid's size; the tallies estimate every segment's predicted skin.
__global__ void __launch_bounds__(256) k_rs_assign(ReachBuild b) {
    __shared__ float ws[MLP_NPARAM];  // the model: 21 KB, more than the scalar cache holds
    const long long sl = (long long)(blockIdx.x * 256 + threadIdx.x) * b.sampleStride;
    const int s = sl < b.n ? (int)sl : 0;
    const int c = sl < b.n ? b.cell[s] : -1;
    const bool leads = c >= 0 && b.cont[b.off[c]] == s;
    if (!__syncthreads_or(leads)) return;  // the whole block: no cell to evaluate
    for (int q = threadIdx.x; q < MLP_NPARAM; q += 256) ws[q] = b.model[q];
    __syncthreads();
    if (!leads) return;
    GridGeom g{};  // raw displacements, as the guard takes them
    float f[SKIN_FEATURES], y[MLP_OUT];
    skin_cell_features(c, b.off, b.cont, b.n, b.coord, b.ref, b.uVec, b.refU, b.links, b.since,
                       __int_as_float(b.words[RS_GEOM + 3]), g, f);
    mlp_forward(ws, f, y);
    const float sk = fminf(b.skinMax, fmaxf(b.skinMin, b.margin * expf(y[0])));
    const float u = logf(sk / b.skinMin) / logf(b.skinMax / b.skinMin);  // [0, 1]
    const int bin = min(RS_HBINS - 1, max(0, (int)(u * (float)RS_HBINS)));
    atomicAdd(&b.words[RS_HIST + bin], b.off[c + 1] - b.off[c]);
    atomicMax(&b.words[RS_MAXSKIN], 1);  // the model ran
}

// One thread: this build's coarse grid (cell = the reach of the longest segment + the skin, at least the floor)
// over the binding grid's extent; opens the outside-skin list.  Learned mode: the smallest skin (histogram bin
// edge) whose cell leaves at most extraTrigger segments predicted to outrun their skins over the horizon (scaled
// from the sampled cells); those few go on the outside-skin list without bringing the next rebuild forward.  So
// the model grows the cell above the floor only where it predicts more motion than the floor cell's skins hold.
__global__ void k_rs_geom(ReachBuild b) {
    if (threadIdx.x != 0) return;
    float maxSkin = b.skinDefault;
    const float last = __int_as_float(b.words[RS_LASTSKIN]);
    if (b.keepLearned && last > 0.0f) maxSkin = last;  // a stale rebuild keeps the model's last choice
    if (b.learned && b.words[RS_MAXSKIN] > 0) {
        maxSkin = b.skinMax;
        for (int bin = 0; bin < RS_HBINS; bin++) {
            const float sk = b.skinMin * powf(b.skinMax / b.skinMin, (float)(bin + 1) / (float)RS_HBINS);
            const float cell = fmaxf(b.floorCell, b.R0 + sk + 2.0f * RS_SLACK);
            const float allow = cell - b.R0 - 2.0f * RS_SLACK;  // the skin the cell leaves a segment
            long long above = 0;  // segments predicted past it: bins whose lower edge is at or above it
            for (int k = 0; k < RS_HBINS; k++) {
                const float lo = b.skinMin * powf(b.skinMax / b.skinMin, (float)k / (float)RS_HBINS);
                if (lo >= allow) above += (long long)b.words[RS_HIST + k] * b.sampleStride;
            }
            if (above <= b.trigger) {
                maxSkin = fminf(b.skinMax, sk);
                break;
            }
        }
        b.words[RS_LASTSKIN] = __float_as_int(maxSkin);
    }
    b.words[RS_MAXSKIN] = 0;
    for (int k = 0; k < RS_HBINS; k++) b.words[RS_HIST + k] = 0;
    const float cell = fmaxf(b.floorCell, b.R0 + maxSkin + 2.0f * RS_SLACK);
    int nX = max(1, (int)ceilf(b.spanX / cell)), nY = max(1, (int)ceilf(b.spanY / cell));
    int nZ = max(1, (int)ceilf(b.spanZ / cell));
    if ((long long)nX * nY * nZ > b.cellsMax) nX = nY = nZ = 1;  // one cell holds every segment: still complete
    b.words[RS_GEOM] = __float_as_int(b.ox);
    b.words[RS_GEOM + 1] = __float_as_int(b.oy);
    b.words[RS_GEOM + 2] = __float_as_int(b.oz);
    b.words[RS_GEOM + 3] = __float_as_int(cell);
    b.words[RS_GEOM + 4] = __float_as_int(1.0f / cell);
    b.words[RS_GEOM + 5] = nX;
    b.words[RS_GEOM + 6] = nY;
    b.words[RS_GEOM + 7] = nZ;
    b.words[RS_EXTRA] = 0;
    b.words[RS_REBUILDS] += 1;
    b.words[RS_GEN] = b.gen;
}

// Per segment: reference centre (and axis), skin, coarse cell; a segment too long for the cell goes on the list.
__global__ void __launch_bounds__(256) k_rs_bin(ReachBuild b) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= b.n) return;
    const int n = b.n;
    const float x = b.coord[s], y = b.coord[n + s], z = b.coord[2 * n + s];
    b.ref[s] = x;
    b.ref[n + s] = y;
    b.ref[2 * n + s] = z;
    if (b.refU) {
        b.refU[s] = b.uVec[s];
        b.refU[n + s] = b.uVec[n + s];
        b.refU[2 * n + s] = b.uVec[2 * n + s];
    }
    const float cell = __int_as_float(b.words[RS_GEOM + 3]), inv = __int_as_float(b.words[RS_GEOM + 4]);
    const float sk = cell - (0.5f * b.segLength[s] + b.tol) - 2.0f * RS_SLACK;
    if (sk > 0.0f) {
        const int nX = b.words[RS_GEOM + 5], nY = b.words[RS_GEOM + 6], nZ = b.words[RS_GEOM + 7];
        const int ix = clamp_bin((x - __int_as_float(b.words[RS_GEOM])) * inv, nX);
        const int iy = clamp_bin((y - __int_as_float(b.words[RS_GEOM + 1])) * inv, nY);
        const int iz = clamp_bin((z - __int_as_float(b.words[RS_GEOM + 2])) * inv, nZ);
        b.skin[s] = sk;
        b.cell[s] = ix + nX * (iy + nY * iz);
    } else {
        b.skin[s] = -1.0f;
        b.cell[s] = -1;  // not in the coarse CSR (the sort skips negative keys)
        const int e = atomicAdd(&b.words[RS_EXTRA], 1);
        if (e < b.extraCap) b.extra[e] = s;
    }
}

// zero + MotorJointSystem.joints for one sub-body: F = (float)(0 + joints)
// ... and DimerCouplingSystem.couple (DimerCouplingSystem.java:83-184) for a dimer's rods and levers: each of those
// sub-bodies' threads evaluates the coupling terms that act on it (from the step-start pose, which the joints
// also read) and adds them after its joints, in the reference's order (variant 0, variant 1, the levers'
// torsion), so no separate read-modify-write pass over the dimers is needed.  dimerOf[m]: the dimer of motor m or -1.
__global__ void __launch_bounds__(256) k_motor_joints0(sbx_motors m, JointParams jp, DimerParams dp,
                                                       const int *__restrict__ dimerOf) {
    int s = blockIdx.x * 256 + threadIdx.x;
    int nB = 3 * m.nM;
    if (s >= nB) return;
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    joints_one(s, nB, m.coord, m.uVec, m.segLength, m.bTransGam, m.bRotGam, m.nucleotideState[s / 3], jp, f, t);
    float F[3], T[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        F[a] = (float)(0.0f + f[a]);
        T[a] = (float)(0.0f + t[a]);
    }
    const int mm = s / 3, role = s - 3 * mm;
    const int d = (dimerOf && role < 2) ? dimerOf[mm] : -1;
    if (d >= 0) {
        const int mA = m.motorA[d], mB = m.motorB[d], par = m.parallel[d] != 0;
        const bool sideA = mm == mA;
        if (role == 0) {
            for (int v = 0; v < 2; v++) {
                double FA[3], TA[3], TB[3];
                if (!dimer_variant(v, mA, mB, par, nB, m.coord, m.uVec, m.segLength, m.bTransGam, m.bRotGam, dp, FA,
                                   TA, TB))
                    continue;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    F[a] = sideA ? (float)(F[a] + FA[a]) : (float)(F[a] - FA[a]);
                    T[a] = (float)(T[a] + (sideA ? TA[a] : TB[a]));
                }
            }
        } else if (par) {
            double TL[3];
            const bool both = m.boundSeg[mA] >= 0 && m.boundSeg[mB] >= 0;
            if (!both && dimer_lever(mA, mB, nB, m.uVec, m.bRotGam, dp, TL)) {
#pragma unroll
                for (int a = 0; a < 3; a++) T[a] = sideA ? (float)(T[a] + TL[a]) : (float)(T[a] - TL[a]);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        m.forceSum[a * nB + s] = F[a];
        m.torqueSum[a * nB + s] = T[a];
    }
}

__global__ void __launch_bounds__(256) k_dimer_of(const int *__restrict__ motorA, const int *__restrict__ motorB,
                                                  int nD, int *__restrict__ dimerOf) {
    int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= nD) return;
    dimerOf[motorA[d]] = d;
    dimerOf[motorB[d]] = d;
}


// the tether of each attachment: MiniFilamentSystem.tether, or with nodeMode NodeSystem.tether (1) /
// tetherBoundStiffen (2, boundCoeff in miniParams[1]) onto node spheres in the backbone slots
__global__ void __launch_bounds__(256) k_net_tether(sbx_motors m, MiniParams mp, int nodeMode) {
    int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= m.nMD) return;
    if (nodeMode) {
        int mm = m.tetherMotor[d];
        bool atEnd1;
        double coeff = node_coeff(m.attachCoeffK[d], mm, nodeMode == 2 ? m.boundSeg : nullptr, mp.fracMove, atEnd1);
        node_tether_one(d, 3 * m.nM, m.nBb, m.nMD, m.headBackboneSlot[d], 3 * mm, atEnd1, coeff, mp.dt, m.coord,
                        m.uVec, m.segLength, m.bTransGam, m.forceSum, m.torqueSum, m.bbCoord, m.bbUVec, m.bbYVec,
                        m.bbInvDragY, m.radial, m.miniData);
        return;
    }
    tether_one(d, 3 * m.nM, m.nBb, m.headBackboneSlot[d], 3 * m.tetherMotor[d], m.attachAxial[d], m.coord, m.uVec,
               m.segLength, m.bTransGam, m.bRotGam, m.forceSum, m.torqueSum, m.bbCoord, m.bbUVec, m.bbInvDragY,
               m.miniData, mp);
}

struct BackboneArgs {
    sbx_motors m;
    const int *bbOff, *bbList;
    float dt, bfm;
    int confineOn, checkInt, seed, rng;
    uint32_t stream;
    double tau, halfX, halfY, halfZ, coeff;
    const int *dStep;
    int stepOff;
};

// one backbone (or node): zero -> backboneGather -> confine -> Brownian -> integrate -> derive (in place)
__global__ void __launch_bounds__(256) k_backbone(BackboneArgs a) {
    int b = blockIdx.x * 256 + threadIdx.x;
    const sbx_motors &m = a.m;
    int nBb = m.nBb;
    if (b >= nBb) return;
    double fd[3] = {0, 0, 0}, td[3] = {0, 0, 0};
    csr_payload_sum(b, a.bbOff, a.bbList, m.miniData, MINI_STRIDE, 0, fd, td);
    int step = *a.dStep + a.stepOff;
    float f[3], t[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f[k] = (float)(0.0f + fd[k]);
        t[k] = (float)(0.0f + td[k]);
    }
    Pose p;
    float gT[3], gR[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p.c[k] = m.bbCoord[k * nBb + b];
        p.u[k] = m.bbUVec[k * nBb + b];
        p.y[k] = m.bbYVec[k * nBb + b];
        gT[k] = m.bbTransGam[k * nBb + b];
        gR[k] = m.bbRotGam[k * nBb + b];
    }
    float len = m.bbSegLength[b];
    if (a.confineOn && !(a.checkInt > 1 && (step % a.checkInt) != 0))
        confine_body(p.c, p.u, len, gT[0], a.tau, a.halfX, a.halfY, a.halfZ, a.coeff, f, t);
    if (m.bbForceSum) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            m.bbForceSum[k * nBb + b] = f[k];
            m.bbTorqueSum[k * nBb + b] = t[k];
        }
    }
    float gf[3], gt[3], rf[3], rt[3];
    brownian_noise(a.rng, (uint32_t)b, step, a.seed, a.stream, gf, gt);
    float tS = m.bbBrownTransScale[b], rS = m.bbBrownRotScale[b];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        rf[k] = brown_amp(tS, a.bfm, gT[k], gf[k]);
        rt[k] = brown_amp(rS, a.bfm, gR[k], gt[k]);
    }
    Derived d;
    integrate_derive(p, f, t, rf, rt, gT, gR, a.dt, len, d);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m.bbCoord[k * nBb + b] = p.c[k];
        m.bbUVec[k * nBb + b] = p.u[k];
        m.bbYVec[k * nBb + b] = p.y[k];
        m.bbZVec[k * nBb + b] = d.z[k];
        m.bbEnd1[k * nBb + b] = d.e1[k];
        m.bbEnd2[k * nBb + b] = d.e2[k];
    }
}

// CrossBridgeSystem.bondForces + applyHeadForce for one motor (reads the step-start filament pose)
// (also gives a bound head its slot in its segment's list for this step's motor gather in k_rod)
__global__ void __launch_bounds__(256) k_bond_head(sbx_motors m, const float *__restrict__ fc,
                                                   const float *__restrict__ fu, const float *__restrict__ fy,
                                                   const float *__restrict__ fbRG, const float *__restrict__ fLen,
                                                   XbParams xp, int nSeg, int *__restrict__ segHeadCnt,
                                                   int *__restrict__ segHeads) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= m.nM) return;
    float out[13];
#pragma unroll
    for (int j = 0; j < 13; j++) out[j] = 0.0f;
    int s = m.boundSeg[k];
    if (s >= 0)
        bond_one(k, m.nM, nSeg, s, m.coord, m.uVec, m.yVec, m.bRotGam, fc, fu, fy, fbRG, fLen, m.bindArc[k],
                 m.nucleotideState[k], xp, out);
    if (s >= 0 && s < nSeg) {
        const int slot = atomicAdd(&segHeadCnt[s], 1);
        if (slot < SEG_HEADS) segHeads[(size_t)s * SEG_HEADS + slot] = k;
    }
    float *d = m.bondData + (size_t)k * XB_STRIDE;
#pragma unroll
    for (int j = 0; j < 13; j++) d[j] = out[j];
    int nB = 3 * m.nM, h = 3 * k + 2;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        m.forceSum[a * nB + h] = m.forceSum[a * nB + h] + out[a];
        m.torqueSum[a * nB + h] = m.torqueSum[a * nB + h] + out[3 + a];
    }
}

// Brownian + integrate + derive of one motor sub-body (in place); the head thread also runs
// CrossBridgeSystem.registerForceDot for its motor (CrossBridgeSystem.java:166-187)
__global__ void __launch_bounds__(256) k_motor_integ(sbx_motors m, float dt, float bfm, int seed, int rng,
                                                     uint32_t stream, const int *__restrict__ dStep, int stepOff) {
    int s = blockIdx.x * 256 + threadIdx.x;
    int nB = 3 * m.nM;
    if (s >= nB) return;
    int step = *dStep + stepOff;
    Pose p;
    float f[3], t[3], gT[3], gR[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p.c[k] = m.coord[k * nB + s];
        p.u[k] = m.uVec[k * nB + s];
        p.y[k] = m.yVec[k * nB + s];
        f[k] = m.forceSum[k * nB + s];
        t[k] = m.torqueSum[k * nB + s];
        gT[k] = m.bTransGam[k * nB + s];
        gR[k] = m.bRotGam[k * nB + s];
    }
    float gf[3], gt[3], rf[3], rt[3];
    brownian_noise(rng, (uint32_t)s, step, seed, stream, gf, gt);
    float tS = m.brownTransScale[s], rS = m.brownRotScale[s];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        rf[k] = brown_amp(tS, bfm, gT[k], gf[k]);
        rt[k] = brown_amp(rS, bfm, gR[k], gt[k]);
    }
    Derived d;
    integrate_derive(p, f, t, rf, rt, gT, gR, dt, m.segLength[s], d);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m.coord[k * nB + s] = p.c[k];
        m.uVec[k * nB + s] = p.u[k];
        m.yVec[k * nB + s] = p.y[k];
        m.zVec[k * nB + s] = d.z[k];
        m.end1[k * nB + s] = d.e1[k];
        m.end2[k * nB + s] = d.e2[k];
    }
    if (s % 3 == 2) {
        int mi = s / 3;
        if (m.boundSeg[mi] >= 0) {
            const float *bd = m.bondData + (size_t)mi * XB_STRIDE;
            float fd = bd[12];
            m.forceDotFil[mi] = fd;
            float fx = bd[0], fy = bd[1], fz = bd[2];
            m.forceMag[mi] = (float)sqrt((double)(fx * fx + fy * fy + fz * fz));
            int pl = m.forceDotPlace[mi];
            m.forceDotHist[mi * 10 + pl] = fd;
            m.forceDotPlace[mi] = (pl + 1) % 10;
        } else {
            m.forceDotFil[mi] = 0.0f;
            m.forceMag[mi] = 0.0f;
            m.forceDotPlace[mi] = 0;
            for (int k = 0; k < 10; k++) m.forceDotHist[mi * 10 + k] = 0.0f;
        }
    }
}

// ---------------------------------------------------------------- skin broad phase kernels
// Every kernel but k_skin_pairs is gated on the device flag SKF_REBUILD, so a captured step
// rebuilds exactly when the guard (in k_rod) fired, with no host round trip.  The coarse (rebuild)
// grid is sized on the device at each rebuild from the largest skin in use, so small skins get a
// small neighbourhood to scan.
struct SkinArgs {
    int n, mode, listCap, maxCand, cCellsMax;
    int gate;                // flag word that opens the rebuild kernels (set by the host around a rebuild)
    int gen;                 // generation of the rebuild being built
    int *extra;              // the outside-skin list
    int extraCap;
    int *hostWords;          // pinned {.., .., generation, outside-skin count} (device address)
    const float *coord, *uVec, *boundR;
    const int *links;        // activeLinkCount (may be null)
    GridGeom geom;           // fine grid (the broad phase's)
    GridGeom *cg;            // coarse grid of the last rebuild (device)
    int *cCells;             // its cell count (device)
    int *cCell, *cOff, *cCont;
    float *refC, *refU, *skin, *cellSkin, *cellHorizon;
    int *vlist, *vcount;
    int *flags;
    const float *__restrict__ model;
    float skinDefault, skinMin, skinMax, margin;
    float maxR, cellFloor;   // largest bounding radius; smallest coarse cell (fine cell or skinMin reach)
    const int *bodyCell;     // fine cells of the current pose
    const float4 *pack;      // {centre, bounding radius} of the current pose
    int *candPartner, *candCount;
};

constexpr float SKIN_SLACK = 1.0e-3f;  // um, covers float rounding of the list and guard distances
constexpr int SKF_MAXSKIN = 6;         // flag word: largest skin of the rebuild (float bits)

// Skins of this rebuild.  Learned mode after the first rebuild: thread c is a coarse cell of the
// last rebuild -- features over its members, model, clamp, then the skin of every member;
// otherwise thread i is a body and gets the default skin.  Tracks the largest skin (flag word).
__global__ void __launch_bounds__(256) k_skin_assign(SkinArgs a, const float *__restrict__ model) {
    __shared__ float ws[MLP_NPARAM];  // the model: 21 KB, more than the scalar cache holds
    if (a.flags[a.gate] == 0) return;
    int k = blockIdx.x * 256 + threadIdx.x;
    if (a.mode == SBX_SKIN_LEARNED && a.flags[SKF_REBUILDS] > 0 && model) {  // uniform over the grid
        const int nC = *a.cCells;
        const int p0 = k < nC ? a.cOff[k] : 0, p1 = k < nC ? a.cOff[k + 1] : 0;
        if (!__syncthreads_or(p1 > p0)) {  // the whole block: empty cells or none
            if (k < nC) {
                a.cellSkin[k] = 0.0f;
                a.cellHorizon[k] = 0.0f;
            }
            return;
        }
        for (int q = threadIdx.x; q < MLP_NPARAM; q += 256) ws[q] = model[q];
        __syncthreads();
        if (k >= nC) return;
        if (p1 == p0) {  // an empty cell has no member to give a skin: no model evaluation
            a.cellSkin[k] = 0.0f;
            a.cellHorizon[k] = 0.0f;
            return;
        }
        float f[SKIN_FEATURES], y[MLP_OUT];
        float H = (float)(a.flags[SKF_SINCE] > 0 ? a.flags[SKF_SINCE] : 1);
        skin_cell_features(k, a.cOff, a.cCont, a.n, a.coord, a.refC, a.uVec, a.refU, a.links, H,
                           1.0f / a.cg->invCell, a.geom, f);
        mlp_forward(ws, f, y);  // every lane reads the same weight: LDS broadcasts
        float sk = fminf(a.skinMax, fmaxf(a.skinMin, a.margin * expf(y[0])));
        a.cellSkin[k] = sk;
        a.cellHorizon[k] = expf(y[1]);
        for (int p = p0; p < p1; p++) a.skin[a.cCont[p]] = sk;
        atomicMax(&a.flags[SKF_MAXSKIN], __float_as_int(sk));  // positive floats order as ints
    } else {
        if (k >= a.n) return;
        a.skin[k] = a.skinDefault;
        if (k == 0) atomicMax(&a.flags[SKF_MAXSKIN], __float_as_int(a.skinDefault));
    }
}

// The coarse grid of this rebuild, cell >= 2 maxR + cutoff + 2 maxSkin (+ slack), covering the fine
// grid's extent (clamped) or tiling the cubic box (periodic, >= 3 cells per axis); every thread
// derives it (same arithmetic), thread 0 stores it and opens the rebuild's counters.
__device__ __forceinline__ GridGeom skin_coarse_geom(const SkinArgs &a, float maxSkin) {
    float need = 2.0f * a.maxR + a.geom.cutoff + 2.0f * maxSkin + 2.0f * SKIN_SLACK;
    float cell = fmaxf(need, a.cellFloor);
    GridGeom g = a.geom;
    if (g.periodic) {
        int nb = (int)floorf(g.Lx / cell);
        if (nb < 3) nb = 3;  // skinMax was checked against the box on the host
        g.nX = g.nY = g.nZ = nb;
        g.invCell = (float)nb / g.Lx;
    } else {
        float span = 1.0f / a.geom.invCell;
        g.nX = max(1, (int)ceilf(span * a.geom.nX / cell));
        g.nY = max(1, (int)ceilf(span * a.geom.nY / cell));
        g.nZ = max(1, (int)ceilf(span * a.geom.nZ / cell));
        g.invCell = 1.0f / cell;
    }
    if ((long long)g.nX * g.nY * g.nZ > a.cCellsMax) g.nX = g.nY = g.nZ = 1;  // bounded by cellFloor (host)
    return g;
}

__global__ void __launch_bounds__(256) k_skin_bin(SkinArgs a) {
    if (a.flags[a.gate] == 0) return;
    int i = blockIdx.x * 256 + threadIdx.x;
    GridGeom g = skin_coarse_geom(a, __int_as_float(a.flags[SKF_MAXSKIN]));
    if (i == 0) {
        *a.cg = g;
        *a.cCells = g.nX * g.nY * g.nZ;
        a.flags[SKF_OVERFLOW] = 0;
        a.flags[SKF_ENTRIES] = 0;
        a.flags[SKF_SINCE] = 0;
        a.flags[SKF_REBUILDS] += 1;
        a.flags[SKF_EXTRA] = 0;  // a fresh outside-skin list
        a.flags[SKF_GEN] = a.gen;
    }
    if (i >= a.n) return;
    float4 q = a.pack[i];
    a.cCell[i] = cell_of(q.x, q.y, q.z, g);
}

// Neighbour lists: every j > i within ri + rj + cutoff + s_i + s_j (+ slack) of i, from the 27
// coarse cells around i.  Also records the pose the guard measures displacement from.
__global__ void __launch_bounds__(BP_THREADS) k_skin_build(SkinArgs a) {
    if (a.flags[a.gate] == 0) return;
    int t = xcd_block(blockIdx.x, gridDim.x) * BP_THREADS + threadIdx.x;
    const int n = a.n;
    if (t >= n) return;
    const GridGeom g = *a.cg;
    int i = a.cCont[t];  // coarse-cell order: neighbouring lanes walk neighbouring cells
    const float4 qi = a.pack[i];
    const float icx = qi.x, icy = qi.y, icz = qi.z;
    float base = qi.w + a.geom.cutoff + a.skin[i] + SKIN_SLACK;
    int pre[18], shift[18];
    const int total = stencil_map(a.cCell[i], g, a.cOff, pre, shift);
    int cnt = 0;
    int *list = a.vlist + (size_t)t * a.listCap;  // per coarse slot: k_skin_pairs reads slot t's list
    for (int q = 0; q < total; q += BP_BATCH) {
        int jj[BP_BATCH];
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) jj[u] = (q + u < total) ? a.cCont[stencil_at(pre, shift, q + u)] : -1;
        float px[BP_BATCH], py[BP_BATCH], pz[BP_BATCH], pr[BP_BATCH];
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) {
            int j = jj[u] > i ? jj[u] : i;
            float4 q = a.pack[j];
            px[u] = q.x;
            py[u] = q.y;
            pz[u] = q.z;
            pr[u] = q.w + a.skin[j];
        }
#pragma unroll
        for (int u = 0; u < BP_BATCH; u++) {
            if (jj[u] <= i) continue;
            float dx = icx - px[u], dy = icy - py[u], dz = icz - pz[u];
            if (g.periodic) {
                dx = min_image(dx, g.Lx, g.iLx);
                dy = min_image(dy, g.Ly, g.iLy);
                dz = min_image(dz, g.Lz, g.iLz);
            }
            float R = base + pr[u];
            if (dx * dx + dy * dy + dz * dz <= R * R) {
                if (cnt < a.listCap) list[cnt] = jj[u];
                cnt++;
            }
        }
    }
    a.vcount[t] = cnt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a.refC[k * n + i] = a.coord[k * n + i];
        a.refU[k * n + i] = a.uVec[k * n + i];
    }
    if (cnt > a.listCap) atomicAdd(&a.flags[SKF_OVERFLOW], 1);
    atomicAdd(&a.flags[SKF_ENTRIES], cnt < a.listCap ? cnt : a.listCap);
}

__global__ void __launch_bounds__(256) k_max_radius(const float *__restrict__ r, int n, int *__restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && r[i] > 0.0f) atomicMax(out, __float_as_int(r[i]));  // positive floats order as ints
}

// One candidate test (the broad phase's predicate, same arithmetic) and its sort key (fine stencil
// position << 26 | index); -1 when j is not a candidate of SpatialGrid.broadPhase.  j's fine cell
// is binned from the same centre k_rod binned it from, so it equals the broad phase's cell of j.
__device__ __forceinline__ int skin_key(int j, int3 ci, float icx, float icy, float icz, float ri, float4 qj,
                                        const GridGeom &g) {
    float dx = icx - qj.x, dy = icy - qj.y, dz = icz - qj.z;
    if (g.periodic) {
        dx = min_image(dx, g.Lx, g.iLx);
        dy = min_image(dy, g.Ly, g.iLy);
        dz = min_image(dz, g.Lz, g.iLz);
    }
    float distSq = dx * dx + dy * dy + dz * dz;
    float reach = ri + qj.w + g.cutoff;
    if (!(distSq <= reach * reach)) return -1;
    int pos = stencil_pos(ci, cell3_of(qj.x, qj.y, qj.z, g), g);
    if (pos < 0) return -1;
    return (pos << 26) | j;
}


constexpr int SP_LANES = 4;   // lanes per body in k_skin_pairs
constexpr int SP_KEYS = 64;   // keys a body buffers in LDS (more: the whole block takes the body)
static_assert((256 / SP_LANES) * SP_KEYS == TILE_BODIES * TILE_KEYS, "one key buffer for both phases");

// One body's candidates for k_skin_pairs: its list (every j > i within reach + both skins at the rebuild) --
// or, if i left its skin or its list overflowed, the 27 coarse cells around its current coarse cell --
// without the bodies on the outside-skin list, plus every j > i on that list; past the list's capacity,
// every j > i.
struct SkinBody {
    int i, nMain, total;
    bool walk, brute;
    float4 qi;
    int3 ci;
    const int *list;
    int pre[18], shift[18];
    __device__ void init(const SkinArgs &a, int t, int nE) {
        brute = nE > a.extraCap;
        i = a.cCont[t];
        qi = a.pack[i];
        ci = cell3_of(qi.x, qi.y, qi.z, a.geom);
        const int cnt = a.vcount[t];
        walk = !brute && (a.skin[i] < 0.0f || cnt > a.listCap);
        list = a.vlist + (size_t)t * a.listCap;
        nMain = cnt;
        if (walk) {
            const GridGeom cg = *a.cg;
            nMain = stencil_map(cell_of(qi.x, qi.y, qi.z, cg), cg, a.cOff, pre, shift);
        }
        total = brute ? a.n : nMain + nE;
    }
    __device__ int cand(const SkinArgs &a, int q) const {  // the q-th candidate body, or -1
        if (q >= total) return -1;
        int j;
        if (brute) {
            j = q;
        } else if (q < nMain) {
            j = walk ? a.cCont[stencil_at(pre, shift, q)] : list[q];
            if (a.skin[j] < 0.0f) return -1;  // on the outside-skin list: taken from there
        } else {
            j = a.extra[q - nMain];
        }
        return j > i ? j : -1;
    }
    __device__ int key(const SkinArgs &a, int j, float4 qj) const {
        return skin_key(j, ci, qi.x, qi.y, qi.z, qi.w, qj, a.geom);
    }
};

// Candidate slices of the current pose (SpatialGrid.broadPhase's output) from the skin state, SP_LANES lanes per
// coarse slot t (body i = cCont[t], so neighbouring groups hold neighbouring bodies).  Each candidate is kept by
// the broad phase's own predicate (skin_key, same arithmetic); the group gathers the kept keys (stencil position
// << 26 | j, distinct) in LDS by ballot and places each at its rank, so the slice is in the broad phase's order.
// A body with more than SP_KEYS keys is then done by the whole block: keys gathered into the whole buffer and
// rank-placed, or (more keys than that) the maxCand smallest kept over chunks.
__global__ void __launch_bounds__(256) k_skin_pairs(SkinArgs a) {
    constexpr int G = SP_LANES, GPB = 256 / G;
    __shared__ int keyS[GPB * SP_KEYS];
    __shared__ int heavyS[GPB], heavyN, foundS;
    const int lane = threadIdx.x & 63, gl = lane % G, gb = threadIdx.x / G;
    const int t = blockIdx.x * GPB + gb;
    const int n = a.n;
    const int nE = a.flags[SKF_EXTRA];
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // no other thread reads these words in this kernel
        a.flags[SKF_SINCE] += 1;
        a.flags[SKF_STEPS] += 1;
        if (nE > a.extraCap) a.flags[SKF_BRUTE] += 1;
        a.hostWords[3] = nE;  // the host's (lagged) view of the guard
        a.hostWords[2] = a.flags[SKF_GEN];
    }
    if (threadIdx.x == 0) heavyN = 0;
    const bool live = t < n;  // every thread reaches the barriers below
    SkinBody b;
    b.init(a, live ? t : 0, nE);
    if (!live) b.total = 0;
    const unsigned long long groupBits = ((1ull << G) - 1) << ((lane / G) * G);
    const unsigned long long below = (1ull << lane) - 1;
    int *kb = keyS + gb * SP_KEYS;
    int hits = 0;
    for (int base = 0; base < b.total; base += 2 * G) {
        int jj[2];
#pragma unroll
        for (int u = 0; u < 2; u++) jj[u] = b.cand(a, base + u * G + gl);
        float4 qj[2];
#pragma unroll
        for (int u = 0; u < 2; u++) qj[u] = a.pack[jj[u] >= 0 ? jj[u] : b.i];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int key = jj[u] >= 0 ? b.key(a, jj[u], qj[u]) : -1;
            const bool hit = key >= 0;
            const unsigned long long bits = __ballot(hit) & groupBits;
            const int pos = hits + __popcll(bits & below);
            if (hit && pos < SP_KEYS) kb[pos] = key;
            hits += __popcll(bits);
        }
    }
    __syncthreads();
    if (live) {
        int *slice = a.candPartner + (size_t)b.i * a.maxCand;
        if (hits <= SP_KEYS) {
            for (int k = gl; k < hits; k += G) {
                const int key = kb[k];
                int rank = 0;
                for (int x = 0; x < hits; x++) rank += kb[x] < key ? 1 : 0;
                if (rank < a.maxCand) slice[rank] = key & KEY_INDEX_MASK;
            }
            if (gl == 0) a.candCount[b.i] = hits;
        } else if (gl == 0) {
            heavyS[atomicAdd(&heavyN, 1)] = t;
        }
    }
    __syncthreads();
    const int nHeavy = heavyN;
    for (int h = 0; h < nHeavy; h++) {  // heavy bodies, one at a time with the whole block
        SkinBody hb;
        hb.init(a, heavyS[h], nE);
        if (threadIdx.x == 0) foundS = 0;
        __syncthreads();
        for (int q = threadIdx.x; q < hb.total; q += 256) {
            const int j = hb.cand(a, q);
            const int key = j >= 0 ? hb.key(a, j, a.pack[j]) : -1;
            if (key >= 0) {
                const int p = atomicAdd(&foundS, 1);
                if (p < GPB * SP_KEYS) keyS[p] = key;
            }
        }
        __syncthreads();
        const int m = foundS;
        int *slice = a.candPartner + (size_t)hb.i * a.maxCand;
        if (m <= GPB * SP_KEYS) {
            block_rank_place(keyS, m, a.maxCand, [&](int rank, int key) { slice[rank] = key & KEY_INDEX_MASK; });
        } else {  // more keys than the buffer: keep the maxCand smallest over chunks (exact for any number)
            auto keyAt = [&](int q) -> int {
                const int j = hb.cand(a, q);
                return j >= 0 ? hb.key(a, j, a.pack[j]) : -1;
            };
            block_keep_smallest(hb.total, keyAt, keyS, &foundS);
            const int w = m < a.maxCand ? m : a.maxCand;
            for (int x = threadIdx.x; x < w; x += 256) slice[x] = keyS[x] & KEY_INDEX_MASK;
        }
        if (threadIdx.x == 0) {
            a.candCount[hb.i] = m;
            atomicAdd(&a.flags[SKF_SELECT], 1);
        }
        __syncthreads();  // before the next heavy body reuses the buffers
    }
}

hipEvent_t take_event(sbx_net *net) {
    if (!net->eventPool.empty()) {
        hipEvent_t e = net->eventPool.back();
        net->eventPool.pop_back();
        return e;
    }
    // timing only: no system-scope fence at the record (no cache writeback and invalidation
    // between the kernels timed); nullptr when the event cannot be created (that launch is untimed)
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) e = nullptr;
    return e;
}

struct Scope {
    sbx_net *net;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    Scope(sbx_net *nt, int i) : net(nt), id(i) {
        if (net->profiling && !net->graphSteps) {
            a = take_event(net);
            b = take_event(net);
            if (!a || !b) {
                if (a) net->eventPool.push_back(a);
                if (b) net->eventPool.push_back(b);
                a = b = nullptr;
            } else if (hipEventRecord(a, net->ctx->stream) != hipSuccess) {
                net->eventPool.push_back(a);
                net->eventPool.push_back(b);
                a = b = nullptr;
            }
        }
    }
    ~Scope() {
        if (a) {
            hipEventRecord(b, net->ctx->stream);
            net->pending.push_back({id, {a, b}});
        }
    }
};

// ---------------------------------------------------------------- the capture rule
// The step's side streams form a pipeline: the capturing stream forks the sort stream, the sort stream
// feeds the broad-phase stream, and the capturing stream joins both.  A capture in which the sort stream
// waited on an event the broad-phase stream recorded inside the same capture (against the pipeline) has
// ended the host process twice (no GPU fault; the cause is outside this library).  Every event record and
// stream wait of the step goes through ev_record / ev_wait, which log it while a step is being captured;
// capture_check accepts the capture only if every wait on an event recorded inside the same capture is a
// join (by the capturing stream), a fork (of the capturing stream's record), a wait on the same stream, or
// a wait downstream (the broad-phase stream on the sort stream).  Waits on records made before the capture
// (the previous graph's) are unrestricted.  Stream ids: 0 capturing, 1 broad phase, 2 sort.
enum { CAP_RECORD = 0, CAP_WAIT = 1 };

int capture_check(const int *ops, int nOps) {
    static const int rank[3] = {0, 2, 1};  // pipeline order: capturing, sort, broad phase
    // recorder[e]: stream whose record of event e is the latest inside this capture, -1 before any
    std::vector<int> recorder;
    for (int k = 0; k < nOps; k++) {
        const int op = ops[3 * k], st = ops[3 * k + 1], ev = ops[3 * k + 2];
        if (ev < 0 || st < 0 || st > 2 || (op != CAP_RECORD && op != CAP_WAIT)) return k;
        if ((int)recorder.size() <= ev) recorder.resize(ev + 1, -1);
        if (op == CAP_RECORD) {
            recorder[ev] = st;
        } else {
            const int rec = recorder[ev];
            if (st != 0 && rec > 0 && rec != st && rank[rec] > rank[st]) return k;
        }
    }
    return -1;
}

int stream_id(const sbx_net *net, hipStream_t s) {
    if (s == net->ctx->stream) return 0;
    if (s == net->side) return 1;
    if (s == net->sortS) return 2;
    return 3;  // not one of the step's streams: capture_check reports it
}

void cap_log(sbx_net *net, int op, hipStream_t s, hipEvent_t e) {
    int id = -1;
    for (size_t k = 0; k < net->capEv.size(); k++)
        if (net->capEv[k] == e) id = (int)k;
    if (id < 0) {
        id = (int)net->capEv.size();
        net->capEv.push_back(e);
    }
    net->capLog.push_back(op);
    net->capLog.push_back(stream_id(net, s));
    net->capLog.push_back(id);
}

int ev_record(sbx_net *net, hipEvent_t e, hipStream_t s, const char *what) {
    if (net->graphSteps) cap_log(net, CAP_RECORD, s, e);
    return check_hip(net->ctx, hipEventRecord(e, s), what);
}

int ev_wait(sbx_net *net, hipStream_t s, hipEvent_t e, const char *what) {
    if (net->graphSteps) cap_log(net, CAP_WAIT, s, e);
    return check_hip(net->ctx, hipStreamWaitEvent(s, e, 0), what);
}

// before a capture begins / after it ends
void cap_begin(sbx_net *net) {
    net->capLog.clear();
    net->capEv.clear();
}

int cap_verdict(sbx_net *net) {
    const int bad = capture_check(net->capLog.data(), (int)net->capLog.size() / 3);
    if (bad < 0) return SBX_OK;
    return fail(net->ctx, SBX_ERR_STATE, "capture: operation %d (stream %d %s event %d) breaks the capture rule", bad,
                net->capLog[3 * bad + 1], net->capLog[3 * bad] == CAP_WAIT ? "waits on" : "records",
                net->capLog[3 * bad + 2]);
}

int drain_timers(sbx_net *net) {
    if (net->pending.empty()) return SBX_OK;
    int r = check_hip(net->ctx, hipStreamSynchronize(net->ctx->stream), "profiling sync");
    if (r) return r;
    for (auto &pe : net->pending) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
        net->timers[pe.first].ms += ms;
        net->timers[pe.first].launches += 1;
        net->eventPool.push_back(pe.second.first);
        net->eventPool.push_back(pe.second.second);
    }
    net->pending.clear();
    return SBX_OK;
}

// The pinned words through which the skins' search kernels report the outside-skin count (allocated once).
int host_words(sbx_net *net) {
    if (net->rsHost) return SBX_OK;
    sbx_ctx *c = net->ctx;
    void *h = nullptr;
    int r = check_hip(c, hipHostMalloc(&h, 4 * sizeof(int), hipHostMallocDefault), "hipHostMalloc");
    if (r) return r;
    void *d = nullptr;
    if ((r = check_hip(c, hipHostGetDevicePointer(&d, h, 0), "hipHostGetDevicePointer"))) {
        hipHostFree(h);
        return r;
    }
    net->rsHost = (volatile int *)h;
    net->rsHostDev = (int *)d;
    for (int k = 0; k < 4; k++) net->rsHost[k] = k % 2 ? 0 : -1;
    return SBX_OK;
}

// A reach-skin rebuild on the main stream (between steps or graph replays, never inside a capture): [learned:
// the model's skins over the previous build's cells] -> geometry -> bin -> sort of the segments into the
// coarse grid.  why: 0 stale state, 1 horizon, 2 outside-skin count.
int rs_rebuild(sbx_net *net, int why) {
    sbx_ctx *c = net->ctx;
    hipStream_t s = c->stream;
    const int n = net->n;
    ReachBuild b{};
    b.n = n; b.extraCap = net->rsp.extraCap;
    b.sampleStride = n > RS_MODEL_CELLS ? (n + RS_MODEL_CELLS - 1) / RS_MODEL_CELLS : 1;
    // the model reads each segment's motion since the previous build: not after turnover or a caller's edit
    // (why 0), where segments were born, moved or resized; those rebuilds keep its last skin
    b.learned = net->rsMode == SBX_SKIN_LEARNED && net->rsGen > 0 && net->rsModel && why != 0;
    b.keepLearned = net->rsMode == SBX_SKIN_LEARNED && !b.learned;
    b.trigger = net->rsp.extraTrigger;
    b.gen = net->rsGen + 1;
    b.since = (float)(net->rsSince > 0 ? net->rsSince : 1);
    b.coord = net->coordB[net->cur]; b.uVec = net->uB[net->cur]; b.segLength = net->fil.segLength;
    b.links = net->L > 0 ? net->xl.activeLinkCount : nullptr;
    b.ref = net->rsRef; b.refU = net->rsRefU; b.skin = net->rsSkin;
    b.cell = net->rsCell; b.extra = net->rsExtra; b.words = net->rsWords;
    b.off = net->rsOff; b.cont = net->rsCont; b.model = net->rsModel;
    b.tol = net->kp.myoColTol; b.R0 = net->rsR0; b.floorCell = net->rsFloor;
    b.skinDefault = net->rsp.skinDefault; b.skinMin = net->rsp.skinMin; b.skinMax = net->rsp.skinMax;
    b.margin = net->rsp.margin;
    const float fine = 1.0f / net->rg.invCell;
    b.ox = net->rg.xMin; b.oy = net->rg.yMin; b.oz = net->rg.zMin;
    b.spanX = fine * (float)net->rg.nX; b.spanY = fine * (float)net->rg.nY; b.spanZ = fine * (float)net->rg.nZ;
    b.cellsMax = net->rsCellsMax;
    int r;
    {
        Scope sc(net, T_RS_REBUILD);
        if (b.learned)
            hipLaunchKernelGGL(k_rs_assign, dim3(blocks_for((n + b.sampleStride - 1) / b.sampleStride, 256)),
                               dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_rs_geom, dim3(1), dim3(64), 0, s, b);
        hipLaunchKernelGGL(k_rs_bin, dim3(blocks_for(n, 256)), dim3(256), 0, s, b);
        if ((r = check_launch(c, "reach skin rebuild"))) return r;
        if ((r = counting_sort(c, net->sort, net->rsCell, n, (int)net->rsCellsMax, n, net->rsOff, net->rsCont)))
            return r;
    }
    net->rsGen = b.gen;
    net->rsValid = 1;
    net->rsSince = 0;
    net->rsWhy[why]++;
    return SBX_OK;
}

// Before a step or a graph replay: rebuild if the coarse grid is stale, the horizon has passed, or the latest
// finished step reported more than extraTrigger segments outside their skins (read from pinned memory without
// waiting; a late view only delays the rebuild, the search stays exact).
int rs_maybe_rebuild(sbx_net *net) {
    if (!net->rsMode) return SBX_OK;
    int why = -1;
    if (!net->rsValid) why = 0;
    else if (net->rsp.horizon > 0 && net->rsSince >= net->rsp.horizon) why = 1;
    else if (net->rsHost[0] == net->rsGen && net->rsHost[1] > net->rsp.extraTrigger) why = 2;
    return why < 0 ? SBX_OK : rs_rebuild(net, why);
}

// After a step or a replay: count it, and bound how far the host runs ahead of the device (so the
// outside-skin count it reads is recent).
int rs_advance(sbx_net *net, int steps) {
    const bool fs = net->skinMode != SBX_SKIN_REBUILD && !net->hasMotors;
    if (!net->rsMode && !fs) return SBX_OK;
    if (fs) net->fsSince += steps;
    if (net->rsMode) {
        net->rsSince += steps;
        net->rsSteps += steps;
    }
    sbx_ctx *c = net->ctx;
    const int k = net->rsEvNext;
    int r = SBX_OK;
    if (net->rsEvLive == sbx_net::RS_INFLIGHT) {  // slot k holds the oldest record: wait for it
        r = check_hip(c, hipEventSynchronize(net->rsEv[k]), "reach skin throttle");
        net->rsEvLive--;
    }
    if (!r && !net->rsEv[k]) r = check_hip(c, hipEventCreateWithFlags(&net->rsEv[k], hipEventDisableTiming), "event");
    if (!r) r = check_hip(c, hipEventRecord(net->rsEv[k], c->stream), "reach skin record");
    if (r) return r;
    net->rsEvLive++;
    net->rsEvNext = (k + 1) % sbx_net::RS_INFLIGHT;
    return SBX_OK;
}

// Binding and minifilament mechanics of one step (DenseContractileHarness.cpuStep,
// DenseContractileHarness.java:350-395), before the filament block.  Every kernel reads the
// step-start filament pose (the current ping-pong buffer); motor and backbone bodies integrate in
// place after the last kernel that reads their step-start pose (k_bond_head).
int enqueue_myosin(sbx_net *net, const float *coord, const float *uVec, const float *yVec, int stepOff) {
    sbx_ctx *c = net->ctx;
    hipStream_t s = c->stream;
    const int n = net->n;
    const sbx_motors &m = net->mo;
    const int nM = m.nM, nB = 3 * nM, cap = n + nM;
    int r;
    {
        // (the reach skin's search publishes the heads itself)
        Scope sc(net, T_BIN);
        if (!net->rsMode) {
            hipLaunchKernelGGL(k_pub_heads, dim3(blocks_for(nM, 256)), dim3(256), 0, s, m);
            hipLaunchKernelGGL(k_body_cell_view, dim3(blocks_for(cap, 256)), dim3(256), 0, s, coord, n, m.head, nM,
                               net->geom, net->bodyCell);
        }
        if ((r = check_launch(c, "k_pub_heads/k_body_cell_view"))) return r;
    }
    if (!net->rsMode) {
        Scope sc(net, T_SORT);
        if ((r = counting_sort(c, net->sort, net->bodyCell, cap, net->p.gridDims[3], cap, net->cellOffsets,
                               net->cellContents)))
            return r;
    }
    {
        Scope sc(net, T_MBIND);
        MotorArgs a;
        a.m = m;
        a.e1 = net->fil.end1; a.e2 = net->fil.end2;
        a.off = net->cellOffsets; a.cont = net->cellContents;
        a.ownerStore = net->ownerStore; a.ownerSlot = net->ownerSlot;
        a.rg = net->rg; a.kp = net->kp; a.np = net->np;
        a.maxCand = net->p.maxCand; a.nSeg = n; a.seed = net->mp.seedMotor; a.rng = net->p.rng;
        a.stream = net->p.stream; a.dStep = net->dStep; a.stepOff = stepOff;
        const int G = net->mbGroup;
        if (net->rsMode) {
            ReachArgs ra;
            ra.coord = coord; ra.skin = net->rsSkin; ra.off = net->rsOff; ra.cont = net->rsCont;
            ra.extra = net->rsExtra; ra.words = net->rsWords; ra.hostWords = net->rsHostDev;
            ra.extraCap = net->rsp.extraCap; ra.fine = net->geom;
            auto kern = G == 64 ? k_motor_bind_skin<64>
                                : (G == 32 ? k_motor_bind_skin<32>
                                           : (G == 16 ? k_motor_bind_skin<16>
                                                      : (G == 8 ? k_motor_bind_skin<8> : k_motor_bind_skin<4>)));
            hipLaunchKernelGGL(kern, dim3(blocks_for(nM, 256 / G)), dim3(256), 0, s, a, ra);
        } else {
            const int Gp = G < 16 ? 16 : G;  // the per-step walk takes 16, 32 or 64 lanes
            auto kern = Gp == 64 ? k_motor_bind<64> : (Gp == 32 ? k_motor_bind<32> : k_motor_bind<16>);
            hipLaunchKernelGGL(kern, dim3(blocks_for(nM, 256 / Gp)), dim3(256), 0, s, a);
        }
    }
    {
        Scope sc(net, T_MJOINT);
        // with the dimer coupling of each dimer's rods and levers
        hipLaunchKernelGGL(k_motor_joints0, dim3(blocks_for(nB, 256)), dim3(256), 0, s, m, net->jp, net->dp,
                           m.nD > 0 ? net->dimerOf : nullptr);
    }
    if (m.nMD > 0) {
        Scope sc(net, T_TETHER);
        hipLaunchKernelGGL(k_net_tether, dim3(blocks_for(m.nMD, 256)), dim3(256), 0, s, m, net->mip, net->mp.nodeMode);
    }
    if (m.nBb > 0) {
        Scope sc(net, T_BACKBONE);
        BackboneArgs b;
        b.m = m;
        b.bbOff = net->bbOff; b.bbList = net->bbList;
        b.dt = net->mp.bbBodyParams[0]; b.bfm = net->mp.bbBodyParams[1];
        b.confineOn = net->p.confineOn; b.checkInt = (int)net->p.boxParams[6];
        b.seed = net->mp.seedBackbone; b.rng = net->p.rng; b.stream = net->p.stream;
        b.tau = net->p.boxParams[0];
        b.halfX = 0.5 * net->p.boxParams[1] - net->p.boxParams[4];
        b.halfY = 0.5 * net->p.boxParams[2] - net->p.boxParams[4];
        b.halfZ = 0.5 * net->p.boxParams[3] - net->p.boxParams[4];
        b.coeff = net->p.boxParams[5];
        b.dStep = net->dStep; b.stepOff = stepOff;
        hipLaunchKernelGGL(k_backbone, dim3(blocks_for(m.nBb, 256)), dim3(256), 0, s, b);
    }
    {
        Scope sc(net, T_BOND);
        hipLaunchKernelGGL(k_bond_head, dim3(blocks_for(nM, 256)), dim3(256), 0, s, m, coord, uVec, yVec,
                           net->fil.bRotGam, net->fil.segLength, net->xp, n, net->segHeadCnt, net->segHeads);
    }
    {
        Scope sc(net, T_MINTEG);
        hipLaunchKernelGGL(k_motor_integ, dim3(blocks_for(nB, 256)), dim3(256), 0, s, m, net->mp.bodyParams[0],
                           net->mp.bodyParams[1], net->mp.seedMotor, net->p.rng, net->p.stream, net->dStep, stepOff);
    }
    return check_launch(c, "myosin stage");
}

SkinArgs skin_args(sbx_net *net, const float *coord, const float *uVec) {
    SkinArgs a;
    a.n = net->n; a.mode = net->skinMode; a.listCap = net->sp.listCap; a.maxCand = net->p.maxCand;
    a.gate = SKF_REBUILD;
    a.gen = net->fsGen + 1;
    a.extra = net->fsExtra; a.extraCap = net->sp.extraCap; a.hostWords = net->rsHostDev;
    a.cCellsMax = net->cCellsMax;
    a.coord = coord; a.uVec = uVec; a.boundR = net->boundR;
    a.links = net->L > 0 ? net->xl.activeLinkCount : nullptr;
    a.geom = net->geom; a.cg = net->cgDev; a.cCells = net->cCellsDev;
    a.cCell = net->cCell; a.cOff = net->cOff; a.cCont = net->cCont;
    a.refC = net->refC; a.refU = net->refU; a.skin = net->skin;
    a.cellSkin = net->cellSkin; a.cellHorizon = net->cellHorizon;
    a.vlist = net->vlist; a.vcount = net->vcount;
    a.flags = net->skinFlags; a.model = net->model;
    a.skinDefault = net->sp.skinDefault; a.skinMin = net->sp.skinMin; a.skinMax = net->sp.skinMax;
    a.margin = net->sp.margin;
    a.maxR = net->maxR; a.cellFloor = net->cellFloor;
    a.bodyCell = net->bodyCell;
    a.pack = net->bodyPack[net->cur];
    a.candPartner = net->candPartner; a.candCount = net->candCount;
    return a;
}

// A skin rebuild on the main stream (the host's, between steps or graph replays, never inside a capture):
// skins (model) -> coarse geometry + bin -> coarse sort -> lists; the gate word opens the kernels for it.
// why: 0 stale state, 1 horizon, 2 outside-skin count.
int fs_rebuild(sbx_net *net, int why) {
    sbx_ctx *c = net->ctx;
    hipStream_t s = c->stream;
    const int n = net->n;
    SkinArgs a = skin_args(net, net->coordB[net->cur], net->uB[net->cur]);
    int r;
    {
        Scope sc(net, T_SKIN_REBUILD);
        if ((r = fill_i32(c, net->skinFlags + SKF_REBUILD, 1, 1))) return r;
        if ((r = fill_i32(c, net->skinFlags + SKF_MAXSKIN, 0, 1))) return r;
        int span = net->cCellsMax > n ? net->cCellsMax : n;
        hipLaunchKernelGGL(k_skin_assign, dim3(blocks_for(span, 256)), dim3(256), 0, s, a, a.model);
        hipLaunchKernelGGL(k_skin_bin, dim3(blocks_for(n, 256)), dim3(256), 0, s, a);
        if ((r = check_launch(c, "skin assign/bin"))) return r;
        if ((r = counting_sort(c, net->sort, net->cCell, n, net->cCellsMax, n, net->cOff, net->cCont,
                               net->skinFlags + a.gate, s)))
            return r;
        hipLaunchKernelGGL(k_skin_build, dim3(blocks_for(n, BP_THREADS)), dim3(BP_THREADS), 0, s, a);
        if ((r = check_launch(c, "k_skin_build"))) return r;
        if ((r = fill_i32(c, net->skinFlags + SKF_REBUILD, 0, 1))) return r;
    }
    net->fsGen = a.gen;
    net->fsValid = 1;
    net->fsSince = 0;
    net->fsWhy[why]++;
    return SBX_OK;
}

// Before a step or a graph replay in a skin mode: rebuild on stale state, past the horizon, or when the latest
// finished step reported more than extraTrigger bodies outside their skins (pinned memory, read without waiting).
int fs_maybe_rebuild(sbx_net *net) {
    if (net->skinMode == SBX_SKIN_REBUILD || net->hasMotors) return SBX_OK;
    int why = -1;
    if (!net->fsValid) why = 0;
    else if (net->sp.horizon > 0 && net->fsSince >= net->sp.horizon) why = 1;
    else if (net->rsHost[2] == net->fsGen && net->rsHost[3] > net->sp.extraTrigger) why = 2;
    return why < 0 ? SBX_OK : fs_rebuild(net, why);
}

// the skin broad phase of one step: candidate slices from the lists (main stream)
int enqueue_skin_pairs(sbx_net *net, const float *coord, const float *uVec) {
    sbx_ctx *c = net->ctx;
    SkinArgs a = skin_args(net, coord, uVec);
    Scope sc(net, T_SKIN_PAIRS);
    hipLaunchKernelGGL(k_skin_pairs, dim3(blocks_for(net->n, 256 / SP_LANES)), dim3(256), 0, c->stream, a);
    return check_launch(c, "k_skin_pairs");
}

// Does the step run its broad phase on the side stream?  Only for the filament-only step with the
// grid on, and not while per-kernel timers (events on the main stream) are recorded.
bool overlapped(const sbx_net *net) {
    return net->overlap && net->side && !net->profiling && !net->hasMotors && net->p.gridOn &&
           net->skinMode == SBX_SKIN_REBUILD;
}

// Outside a capture: every event whose latest record was captured is recorded again on the main
// stream, which follows all the work of the captures and replays launched so far (a graph joins its
// side chains before it ends), so a later wait on it is a valid, conservative one.  Called before
// any eager launch that waits on an event.
int settle_events(sbx_net *net) {
    if (!net->evCaptured || net->graphSteps) return SBX_OK;
    sbx_ctx *c = net->ctx;
    for (hipEvent_t e : {net->evFork, net->evRod[0], net->evRod[1], net->evPlaced[0], net->evPlaced[1],
                         net->evJoin[0], net->evJoin[1]}) {
        if (!e) continue;
        int r = ev_record(net, e, c->stream, "settle events");
        if (r) return r;
    }
    net->evCaptured = 0;
    return SBX_OK;
}

// The main stream waits for all side-stream work enqueued so far (every chain ends in an evJoin
// record: the broad phase waits for its sort); the next side work forks anew.
int join_side(sbx_net *net) {
    if (!overlapped(net)) return SBX_OK;
    if (net->evCaptured && !net->graphSteps) {  // only captures and replays since: already joined
        net->needFork = 1;
        return SBX_OK;
    }
    sbx_ctx *c = net->ctx;
    int r = SBX_OK;
    // inside a capture only on the capture's own records; a 1-step graph's last sort ends no evJoin
    // chain, so the sorts are waited for too
    const bool rebuild = net->skinMode == SBX_SKIN_REBUILD && net->p.gridOn && !net->hasMotors;
    for (int b = 0; b < 2 && !r; b++) {
        if (!net->graphSteps || net->joinLive[b])
            r = ev_wait(net, c->stream, net->evJoin[b], "join");
        if (!r && rebuild && (!net->graphSteps || net->placedLive[b]))
            r = ev_wait(net, c->stream, net->evPlaced[b], "join");
    }
    net->needFork = 1;
    return r;
}

// The broad phase on CSR[b] (side stream s2): after that CSR's sort, recorded as evJoin[b].
int enqueue_bp(sbx_net *net, int b, hipStream_t s2, bool ovl, bool waitPlaced) {
    sbx_ctx *c = net->ctx;
    int r;
    Scope sc(net, T_PAIRS);
    if (ovl && waitPlaced && (r = ev_wait(net, s2, net->evPlaced[b], "wait evPlaced")))
        return r;
    r = launch_broad_phase_sorted(c, net->sortedPose[b], net->sortedCell[b], net->csrOff[b], net->csrCont[b],
                                  net->geom, net->reps, net->rodRep, net->bpGroup, net->bpStaged, net->n, net->n,
                                  net->candPartner, net->candCount, net->p.maxCand, s2);
    if (r) return r;
    if (ovl && (r = ev_record(net, net->evJoin[b], s2, "record evJoin"))) return r;
    if (net->graphSteps) net->joinLive[b] = 1;
    net->lastCsr = b;
    return SBX_OK;
}

// One step's launches; stepOff is added to the device step counter.  With overlap the broad phase
// (its output is the step's candidate slices, read by no later kernel of the step) leaves the
// main stream.  Per-step rebuild, pose buffer c = cur, x = 1 - c, three streams:
//   sort:  (wait evRod[x]: the previous k_rod binned, counted and packed this pose)
//          (wait evJoin[c]: the broad phase two steps back last read CSR[c])
//          sort -> CSR[c] (+ pose and cell in contents order) -> evPlaced[c]
//   side:  (wait evPlaced[c]) broad phase on CSR[c] -> evJoin[c]
//   main:  (wait evPlaced[x]: k_rod overwrites the pack, cells and histogram the previous sort read)
//          unbind -> link -> k_rod -> evRod[c]
// so the next step's sort runs while this step's broad phase does.  Skin modes run on one stream
// (pairs -> unbind -> link -> k_rod): the pairs read the skin flags and outside-skin list that the
// guard in k_rod extends; the host enqueues rebuilds between steps.  The side streams are forked from the main stream (needFork) at the start of a call, of a capture
// and after a graph replay, so a wait on an event recorded before the fork is only redundant.
// capK: index of the step in the graph being captured (-1: eager); a capture waits on no event
// recorded outside it.
int enqueue_step(sbx_net *net, int stepOff, int capK = -1, bool deferBp = false) {
    sbx_ctx *c = net->ctx;
    hipStream_t s = c->stream;
    const int n = net->n, L = net->L;
    const float *coord = net->coordB[net->cur], *uVec = net->uB[net->cur], *yVec = net->yB[net->cur];
    const bool ovl = overlapped(net);
    const bool skinOn = net->skinMode != SBX_SKIN_REBUILD && !net->hasMotors;
    hipStream_t s2 = ovl ? net->side : s, sS = ovl ? net->sortS : s;
    const int cb = net->cur, xb = 1 - cb;
    const bool prev1 = capK < 0 || capK >= 1, prev2 = capK < 0 || capK >= 2;  // step s-1 / s-2 recorded here
    int r = settle_events(net);
    if (r) return r;
    if (net->hasMotors) {
        if ((r = enqueue_myosin(net, coord, uVec, yVec, stepOff))) return r;
    } else if (net->p.gridOn) {
        if (!net->derivedValid) return fail(c, SBX_ERR_STATE, "net step: derived pose data not prepared");
        if (ovl && (net->needFork || capK == 0)) {
            if ((r = ev_record(net, net->evFork, s, "record evFork"))) return r;
            if ((r = ev_wait(net, s2, net->evFork, "wait evFork"))) return r;
            if ((r = ev_wait(net, sS, net->evFork, "wait evFork"))) return r;
            net->needFork = 0;
        }
        if (skinOn) {  // one stream: the pairs read the flags and list the guard in k_rod writes
            if ((r = enqueue_skin_pairs(net, coord, uVec))) return r;
        } else {
            if (net->bpPending) {  // the broad phase the previous step left; its sort is ordered before
                if ((r = enqueue_bp(net, net->bpPendingBuf, s2, ovl, !net->graphSteps))) return r;
                net->bpPending = 0;
            }
            {
                Scope sc(net, T_SORT);
                if (ovl && prev1 && (r = ev_wait(net, sS, net->evRod[xb], "wait evRod")))
                    return r;
                if (ovl && prev2 && (r = ev_wait(net, sS, net->evJoin[cb], "wait evJoin")))
                    return r;
                SortPayload pay;
                pay.in = net->bodyPack[cb];
                pay.out = net->sortedPose[cb];
                pay.keyOut = net->sortedCell[cb];
                r = counting_sort_counted(c, net->sort, net->bodyCell, n, net->p.gridDims[3], n, net->csrOff[cb],
                                          net->csrCont[cb], net->gridCounts[cb], net->gridSlot[cb], sS, pay);
                if (r) return r;
                if (ovl && (r = ev_record(net, net->evPlaced[cb], sS, "record evPlaced"))) return r;
                if (net->graphSteps) net->placedLive[cb] = 1;
                else net->csrValid[cb] = 1;
            }
            if (deferBp && ovl) {  // left to the next step (the next 1-step graph's first node)
                net->bpPending = 1;
                net->bpPendingBuf = cb;
            } else if ((r = enqueue_bp(net, cb, s2, ovl, true))) {
                return r;
            }
            if (ovl && prev1 && (r = ev_wait(net, s, net->evPlaced[xb], "wait evPlaced")))
                return r;
        }
    }
    if (L > 0) {
        if (net->p.unbindOn) {
            Scope sc(net, T_UNBIND);
            auto kern = net->reps ? k_net_unbind<true> : k_net_unbind<false>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(L, 256)), dim3(256), 0, s, net->linkPack, net->xl.linkFilA,
                               net->xl.linkFilB, net->xl.loc1, net->xl.loc2, net->xl.linkState, net->xl.strainHist,
                               net->xl.strainPlace, net->lp, net->dStep, stepOff, net->p.seed, net->p.rng,
                               net->p.stream, net->xl.activeLinkCount, net->reps, net->linkRep, n, L);
        }
        {
            Scope sc(net, T_LINK);
            hipLaunchKernelGGL(k_net_link, dim3(blocks_for(L, 256)), dim3(256), 0, s, net->linkPack,
                               net->xl.linkFilA, net->xl.linkFilB, net->xl.loc1, net->xl.loc2,
                               net->xl.activeLinkCount, net->xl.linkState, net->xl.linkOrientSame,
                               net->xl.torqueMagHist,
                               net->xl.torqueMagPlace, net->xl.xlinkData, net->lp, n, L);
        }
        r = check_launch(c, "link kernels");
        if (r) return r;
    }
    RodArgs a;
    a.coord = coord; a.uVec = uVec; a.yVec = yVec;
    int nxt = 1 - net->cur;
    a.coordOut = net->coordB[nxt]; a.uOut = net->uB[nxt]; a.yOut = net->yB[nxt];
    a.zVec = net->fil.zVec; a.end1 = net->fil.end1; a.end2 = net->fil.end2;
    a.segLength = net->fil.segLength; a.bTransGam = net->fil.bTransGam; a.bRotGam = net->fil.bRotGam;
    a.bts = net->fil.brownTransScale; a.brs = net->fil.brownRotScale;
    a.forceSum = net->fil.forceSum; a.torqueSum = net->fil.torqueSum;
    a.e1Slot = net->fil.end1NbrSlot; a.e1Side = net->fil.end1NbrSide;
    a.e2Slot = net->fil.end2NbrSlot; a.e2Side = net->fil.end2NbrSide;
    a.offA = net->offA; a.idxA = net->idxA; a.offB = net->offB; a.idxB = net->idxB;
    a.linkState = net->xl.linkState; a.xlinkData = net->xl.xlinkData;
    a.hasMotors = net->hasMotors;
    a.activeLinkCount = net->xl.activeLinkCount;
    a.geom = net->geom;
    a.reps = net->reps;
    a.rodRep = net->rodRep;
    a.nextCell = (net->p.gridOn && !net->hasMotors) ? net->cellNext : nullptr;
    a.nextCounts = (a.nextCell && !skinOn) ? net->gridCounts[nxt] : nullptr;
    a.nextSlot = a.nextCounts ? net->gridSlot[nxt] : nullptr;
    a.linkPack = L > 0 ? net->linkPack : nullptr;
    a.skinRef = skinOn ? net->refC : nullptr;
    a.skin = skinOn ? net->skin : nullptr;
    a.rebuildFlag = skinOn ? net->skinFlags + SKF_EXTRA : nullptr;
    a.rsExtra = skinOn ? net->fsExtra : nullptr;
    a.rsCap = skinOn ? net->sp.extraCap : 0;
    a.minImage = skinOn && net->geom.periodic;
    if (net->rsMode) {
        a.skinRef = net->rsRef;
        a.skin = net->rsSkin;
        a.rebuildFlag = net->rsWords + RS_EXTRA;
        a.rsExtra = net->rsExtra;
        a.rsCap = net->rsp.extraCap;
        a.minImage = 0;
    }
    a.bodyPack = (net->p.gridOn && !net->hasMotors) ? net->bodyPack[nxt] : nullptr;
    a.boundR = net->boundR;
    a.segHeadCnt = net->segHeadCnt; a.segHeads = net->segHeads; a.boundSeg = net->mo.boundSeg; a.nM = net->mo.nM;
    a.bondData = net->mo.bondData;
    a.dStep = net->dStep; a.stepOff = stepOff;
    a.n = n; a.chainOn = net->p.chainOn; a.hasLinks = L > 0;
    a.confineOn = net->p.confineOn; a.checkInt = (int)net->p.boxParams[6];
    a.rng = net->p.rng; a.seed = net->p.seed; a.stream = net->p.stream;
    a.dt = net->p.params[0]; a.bfm = net->p.params[1];
    a.tau = net->p.boxParams[0];
    a.halfX = 0.5 * net->p.boxParams[1] - net->p.boxParams[4];
    a.halfY = 0.5 * net->p.boxParams[2] - net->p.boxParams[4];
    a.halfZ = 0.5 * net->p.boxParams[3] - net->p.boxParams[4];
    a.coeff = net->p.boxParams[5];
    {
        Scope sc(net, T_ROD);
        const long long waves = ((long long)n + CHAIN_SPAN - 1) / CHAIN_SPAN;
        if (net->rodThreads == RODS_PAIRED) {  // two rod groups per block of four waves
            auto kern = net->reps ? k_rod<true, true> : k_rod<false, true>;
            hipLaunchKernelGGL(kern, dim3((unsigned)((waves + 1) / 2)), dim3(256), 0, s, a, net->cp);
        } else {
            auto kern = net->reps ? k_rod<true, false> : k_rod<false, false>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(waves * 64, net->rodThreads)), dim3(net->rodThreads), 0, s,
                               a, net->cp);
        }
    }
    if (ovl && (r = ev_record(net, net->evRod[cb], s, "record evRod"))) return r;
    net->cur = nxt;
    net->lastCell = net->bodyCell;
    if (a.nextCell) {  // the next step's grid reads the cells k_rod just wrote
        int *t = net->bodyCell;
        net->bodyCell = net->cellNext;
        net->cellNext = t;
    }
    return check_launch(c, "k_rod");
}

int build_link_csr(sbx_net *net) {
    if (net->L == 0) return SBX_OK;
    int r = counting_sort(net->ctx, net->sort, net->xl.linkFilA, net->L, net->n, net->L, net->offA, net->idxA);
    if (r) return r;
    return counting_sort(net->ctx, net->sort, net->xl.linkFilB, net->L, net->n, net->L, net->offB, net->idxB);
}

// a captured step no longer matches the launch sequence; no event keeps a record made in a capture
// of the graphs destroyed
void destroy_graphs(sbx_net *net) {
    settle_events(net);
    if (net->graphExec) { hipGraphExecDestroy(net->graphExec); net->graphExec = nullptr; }
    if (net->graph) { hipGraphDestroy(net->graph); net->graph = nullptr; }
    for (int b = 0; b < 2; b++) {
        if (net->unitExec[b]) { hipGraphExecDestroy(net->unitExec[b]); net->unitExec[b] = nullptr; }
        if (net->unitGraph[b]) { hipGraphDestroy(net->unitGraph[b]); net->unitGraph[b] = nullptr; }
    }
}

void release(sbx_net *net) {
    if (!net) return;
    if (net->ctx) hipSetDevice(net->ctx->device);
    if (net->ctx) hipStreamSynchronize(net->ctx->stream);
    if (net->side) hipStreamSynchronize(net->side);
    if (net->sortS) hipStreamSynchronize(net->sortS);
    destroy_graphs(net);
    void *bufs[] = {net->ownPose,  net->boundR,    net->bodyCell, net->cellNext,    net->cellContents, net->linkPack,
                    net->cellOffsets,
                    net->candPartner, net->candCount, net->offA,  net->idxA,        net->offB,
                    net->idxB,     net->dStep,     net->ownerStore, net->ownerSlot,
                    net->segHeadCnt, net->segHeads, net->bbOff,  net->bbList, net->dimerOf,
                    net->cCell,    net->cOff,      net->cCont,    net->refC,    net->refU, net->skin,
                    net->cellSkin, net->cellHorizon, net->vlist,  net->vcount,  net->skinFlags, net->model,
                    net->cgDev,    net->cCellsDev, net->bodyPack[0], net->bodyPack[1], net->sortedPose[0],
                    net->sortedPose[1], net->sortedCell[0], net->sortedCell[1], net->csrOff[1], net->csrCont[1],
                    net->gridCounts[0], net->gridCounts[1], net->gridSlot[0], net->gridSlot[1],
                    net->reps,     net->rodRep,    net->linkRep,  net->formInts, net->formFloats,
                    net->rsRef,    net->rsRefU,    net->rsSkin,   net->rsCell,  net->rsOff,  net->rsCont,
                    net->rsExtra,  net->rsWords,   net->rsModel, net->fsExtra};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (net->rsHost) hipHostFree((void *)net->rsHost);
    for (hipEvent_t e : net->rsEv)
        if (e) hipEventDestroy(e);
    sort_scratch_free(net->sort);
    sort_scratch_free(net->formSort);
    turn_scratch_free(net->turnScratch);
    for (auto &pe : net->pending) {
        hipEventDestroy(pe.second.first);
        hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : net->eventPool) hipEventDestroy(e);
    for (hipEvent_t e : {net->evFork, net->evRod[0], net->evRod[1], net->evPlaced[0], net->evPlaced[1],
                         net->evJoin[0], net->evJoin[1]})
        if (e) hipEventDestroy(e);
    if (net->side) hipStreamDestroy(net->side);
    if (net->sortS) hipStreamDestroy(net->sortS);
    delete net;
}

template <typename P>
int alloc(sbx_ctx *c, P **p, size_t count) {
    return check_hip(c, hipMalloc((void **)p, sizeof(P) * (count ? count : 1)), "sbx_net alloc");
}

// replicas of a batch in the concatenated stores: one RepInfo per replica, the replica of every rod
// and every link
struct BatchLayout {
    std::vector<RepInfo> reps;
    std::vector<int> rodRep, linkRep;
};

// A network over caller-owned stores; batch != nullptr: the stores are the concatenated replicas of
// a batch (each replica's grid already checked), laid out as batch describes.
// buffers of the formation pass (FormPass): one int and one float allocation carved into arrays
int form_setup(sbx_net *net) {
    sbx_ctx *c = net->ctx;
    const sbx_net_params *p = &net->p;
    const long long n = net->n, L = net->L, K = p->formReqCap, cells = p->formGridDims[3];
    const long long ints = 2 * n + n + (cells + 1) + n + n + (n + 1) + 5 * K + (K + 1) + n + L + (L + 1) + L + 4;
    net->formOn = 1;
    int r = alloc(c, &net->formInts, (size_t)ints);
    if (!r) r = alloc(c, &net->formFloats, (size_t)(2 * K));
    if (r) return r;
    FormPass &f = net->form;
    int *q = net->formInts;
    auto take = [&](long long count) {
        int *b = q;
        q += count;
        return b;
    };
    f.filID = take(n); f.filIDScratch = take(n); f.bodyCell = take(n); f.gridOff = take(cells + 1);
    f.gridCont = take(n); f.candCount = take(n); f.candBase = take(n + 1);
    f.reqFilA = take(K); f.reqFilB = take(K); f.reqOrient = take(K); f.gatePass = take(K); f.acceptFlag = take(K);
    f.rankOffsets = take(K + 1); f.minCand = take(n);
    f.freeCount = take(L); f.freeOffsets = take(L + 1); f.freeList = take(L); f.stats = take(4);
    f.reqLoc1 = net->formFloats;
    f.reqLoc2 = net->formFloats + K;
    if ((r = fill_i32(c, f.stats, 0, 4))) return r;
    f.nSeg = net->n; f.C = net->L; f.reqCap = (int)K;
    f.end1 = net->fil.end1; f.end2 = net->fil.end2; f.segLength = net->fil.segLength;
    f.filState = nullptr;
    f.end2NbrSlot = net->fil.end2NbrSlot;
    f.linkFilA = net->xl.linkFilA; f.linkFilB = net->xl.linkFilB; f.linkState = net->xl.linkState;
    f.strainPlace = net->xl.strainPlace; f.linkOrientSame = net->xl.linkOrientSame;
    f.torqueMagPlace = net->xl.torqueMagPlace; f.activeLinkCount = net->xl.activeLinkCount;
    f.loc1 = net->xl.loc1; f.loc2 = net->xl.loc2; f.strainHist = net->xl.strainHist;
    f.torqueMagHist = net->xl.torqueMagHist;
    f.offA = net->offA; f.idxA = net->idxA; f.offB = net->offB; f.idxB = net->idxB;
    f.geom = grid_geom(p->formGridParams, p->formGridDims, p->periodic, p->periodic ? p->boxLen : nullptr);
    f.fp = form_params(p->formParams, p->formMode);
    f.rng = p->rng; f.seed = p->seed; f.stream = p->stream;
    long long keys = cells;
    for (long long v : {n + 1, K + 1, L + 1}) keys = keys > v ? keys : v;
    return sort_scratch_reserve(c, net->formSort, keys, n > L ? n : L);
}

__global__ void k_form_tally(int *stats) { stats[2] += stats[1]; }

// One formation pass before step `step` (V2OneXHarness.java:379), on the main stream after the previous
// step's mechanics: it reads the pose, its derived ends and the start-of-pass link state, writes new
// links into free slots, then rebuilds the link CSR; the next step recounts the active links.
int run_formation(sbx_net *net, int step) {
    sbx_ctx *c = net->ctx;
    int r;
    if (!net->countsValid) {  // start-of-pass active counts (countActiveLinks, V2OneXHarness.java:449)
        hipLaunchKernelGGL(k_link_counts, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream, net->offA,
                           net->idxA, net->offB, net->idxB, net->xl.linkState, net->xl.activeLinkCount, net->n);
        if ((r = check_launch(c, "k_link_counts"))) return r;
        net->countsValid = 1;
    }
    FormPass &f = net->form;
    f.coord = net->coordB[net->cur];
    f.uVec = net->uB[net->cur];
    f.filState = net->turnOn ? net->turn.filState : nullptr;
    if ((r = form_pass(c, net->formSort, f, step, net->formFilIDValid != 0, c->stream))) return r;
    net->formFilIDValid = 1;
    hipLaunchKernelGGL(k_form_tally, dim3(1), dim3(1), 0, c->stream, f.stats);
    if ((r = counting_sort(c, net->formSort, net->xl.linkFilA, net->L, net->n, net->L, net->offA, net->idxA))) return r;
    if ((r = counting_sort(c, net->formSort, net->xl.linkFilB, net->L, net->n, net->L, net->offB, net->idxB))) return r;
    net->countsValid = 0;
    net->formPasses++;
    return check_launch(c, "formation");
}

// One biochem cadence on the current pose buffers (FullSystemDemoHarness.java:650-684), then the step's
// prepared state that follows the pose, segment lengths and chain topology is marked stale: bounding radii
// are recomputed, the next step re-bins and re-packs, the formation pass recomputes filament identities.
int run_turnover(sbx_net *net, int step) {
    sbx_ctx *c = net->ctx;
    TurnArrays a{};
    a.C = net->n;
    a.coord = net->coordB[net->cur]; a.uVec = net->uB[net->cur]; a.yVec = net->yB[net->cur];
    a.segLength = net->fil.segLength; a.bTransGam = net->fil.bTransGam; a.bRotGam = net->fil.bRotGam;
    a.bts = net->fil.brownTransScale; a.brs = net->fil.brownRotScale;
    a.e1s = net->fil.end1NbrSlot; a.e1d = net->fil.end1NbrSide; a.e2s = net->fil.end2NbrSlot; a.e2d = net->fil.end2NbrSide;
    a.filState = net->turn.filState; a.monomerCount = net->turn.monomerCount; a.seedNode = net->turn.seedNode;
    a.nucFrac = net->turn.nucFrac; a.cofFrac = net->turn.cofFrac;
    a.bTransDiff = net->turn.bTransDiff; a.bRotDiff = net->turn.bRotDiff;
    a.pool = net->turn.pool; a.ledger = net->turn.ledger; a.stats = net->turn.stats;
    int r = turnover_cadence(c, a, net->tp, net->turnScratch, step, net->p.rng, net->p.stream, c->stream);
    if (r) return r;
    if (net->boundR) {
        hipLaunchKernelGGL(k_bounding_radius, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream,
                           net->fil.segLength, net->p.actinRadius, net->boundR, net->n);
        if ((r = check_launch(c, "k_bounding_radius"))) return r;
    }
    net->derivedValid = 0;
    net->formFilIDValid = 0;
    net->rsValid = 0;  // segment lengths, pose and slots changed
    return SBX_OK;
}

int net_create(sbx_ctx *c, const sbx_filaments *fil, const sbx_crosslinks *xl, const sbx_net_params *p,
               const BatchLayout *batch, sbx_net **out) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, fil && p && out, "net_create: null argument");
    *out = nullptr;
    const int n = fil->n;
    SBX_REQUIRE(c, n > 0, "net_create: need n > 0 rods");
    SBX_REQUIRE(c, fil->coord && fil->uVec && fil->yVec && fil->zVec && fil->end1 && fil->end2 && fil->segLength &&
                       fil->bTransGam && fil->bRotGam && fil->brownTransScale && fil->brownRotScale,
                "net_create: null filament array");
    SBX_REQUIRE(c, (fil->forceSum == nullptr) == (fil->torqueSum == nullptr),
                "net_create: forceSum and torqueSum must be both set or both null");
    SBX_REQUIRE(c, p->rng == SBX_RNG_WANG || p->rng == SBX_RNG_PHILOX, "net_create: unknown rng %d", p->rng);
    if (p->chainOn)
        SBX_REQUIRE(c, fil->end1NbrSlot && fil->end1NbrSide && fil->end2NbrSlot && fil->end2NbrSide,
                    "net_create: chain on but topology arrays null");
    const int L = xl ? xl->nLinks : 0;
    SBX_REQUIRE(c, L >= 0, "net_create: nLinks < 0");
    if (L > 0) {
        SBX_REQUIRE(c, xl->linkFilA && xl->linkFilB && xl->loc1 && xl->loc2 && xl->linkState && xl->strainHist &&
                           xl->strainPlace && xl->linkOrientSame && xl->torqueMagHist && xl->torqueMagPlace &&
                           xl->xlinkData,
                    "net_create: null crosslink array");
        SBX_REQUIRE(c, p->xlParams[2] != 0.0f, "net_create: xlParams dt must be non-zero");
        if (p->unbindOn) SBX_REQUIRE(c, p->offParams[4] != 0.0f, "net_create: offParams restLength must be non-zero");
    }
    if (p->confineOn) SBX_REQUIRE(c, p->boxParams[0] != 0.0f, "net_create: boxParams tau must be non-zero");
    if (p->formOn) {
        SBX_REQUIRE(c, !batch, "net_create: formation is not batched");
        SBX_REQUIRE(c, L > 0, "net_create: formation needs a crosslink pool (nLinks > 0)");
        SBX_REQUIRE(c, p->formInt > 0 && p->formReqCap > 0, "net_create: formInt and formReqCap must be > 0");
        SBX_REQUIRE(c, fil->end2NbrSlot, "net_create: formation needs end2NbrSlot (filament identity)");
        SBX_REQUIRE(c, xl->activeLinkCount, "net_create: formation needs activeLinkCount");
        int r = check_grid(c, p->formGridParams, p->formGridDims, p->periodic);
        if (r) return r;
        if (p->periodic)
            SBX_REQUIRE(c, p->formGridDims[0] >= 3 && p->formGridDims[1] >= 3 && p->formGridDims[2] >= 3,
                        "net_create: a periodic formation grid needs >= 3 cells per axis");
    }
    if (p->gridOn) {
        int r = batch ? SBX_OK : check_grid(c, p->gridParams, p->gridDims, p->periodic);
        if (r) return r;
        SBX_REQUIRE(c, p->maxCand > 0 && p->maxCand <= SBX_MAX_CAND, "net_create: maxCand must be in [1, %d]",
                    SBX_MAX_CAND);
        SBX_REQUIRE(c, (long long)n * p->maxCand < (1LL << 33), "net_create: candidate slices too large");
    }
    int r = set_device(c);
    if (r) return r;
    sbx_net *net = new (std::nothrow) sbx_net();
    if (!net) return fail(c, SBX_ERR_NOMEM, "net_create: host allocation");
    net->ctx = c;
    net->fil = *fil;
    if (xl) net->xl = *xl;
    net->xl.nLinks = L;
    net->p = *p;
    net->n = n;
    net->L = L;
    net->timers.resize(T_COUNT_TIMERS);
    for (int k = 0; k < T_COUNT_TIMERS; k++) net->timers[k].name = kTimerNames[k];
    net->cp = chain_params(p->chainParams);
    net->lp = link_params(p->xlParams, p->offParams, p->torsionParams);
    do {
        if ((r = alloc(c, &net->ownPose, (size_t)9 * n))) break;
        net->coordB[0] = fil->coord; net->uB[0] = fil->uVec; net->yB[0] = fil->yVec;
        net->coordB[1] = net->ownPose; net->uB[1] = net->ownPose + 3 * (size_t)n;
        net->yB[1] = net->ownPose + 6 * (size_t)n;
        if ((r = alloc(c, &net->dStep, 1))) break;
        if (p->gridOn) {
            net->geom = batch ? batch->reps[0].g
                              : grid_geom(p->gridParams, p->gridDims, p->periodic, p->periodic ? p->boxLen : nullptr);
            if ((r = alloc(c, &net->boundR, n))) break;
            if ((r = alloc(c, &net->bodyCell, n))) break;
            if ((r = alloc(c, &net->cellNext, n))) break;
            if ((r = fill_i32(c, net->bodyCell, -1, n))) break;  // no cells yet (the first step bins)
            if ((r = alloc(c, &net->cellOffsets, (size_t)p->gridDims[3] + 1))) break;
            if ((r = alloc(c, &net->cellContents, n))) break;
            if ((r = alloc(c, &net->candPartner, (size_t)n * p->maxCand))) break;
            if ((r = alloc(c, &net->candCount, n))) break;
            if ((r = alloc(c, &net->bodyPack[0], n))) break;
            if ((r = alloc(c, &net->bodyPack[1], n))) break;
            // default priority, like the context's stream
            if ((r = check_hip(c, hipStreamCreateWithFlags(&net->side, hipStreamNonBlocking), "side stream"))) break;
            if ((r = check_hip(c, hipStreamCreateWithFlags(&net->sortS, hipStreamNonBlocking), "sort stream"))) break;
            for (hipEvent_t *e : {&net->evFork, &net->evRod[0], &net->evRod[1], &net->evPlaced[0],
                                  &net->evPlaced[1], &net->evJoin[0], &net->evJoin[1]})
                if (!r) r = check_hip(c, hipEventCreateWithFlags(e, hipEventDisableTiming), "event");
            if (r) break;
            net->csrOff[0] = net->cellOffsets;
            net->csrCont[0] = net->cellContents;
            if ((r = alloc(c, &net->csrOff[1], (size_t)p->gridDims[3] + 1))) break;
            if ((r = alloc(c, &net->csrCont[1], n))) break;
            for (int b = 0; b < 2 && !r; b++) {
                if ((r = alloc(c, &net->sortedPose[b], n))) break;
                r = alloc(c, &net->sortedCell[b], n);
            }
            if (r) break;
            for (int b = 0; b < 2 && !r; b++) {
                if ((r = alloc(c, &net->gridCounts[b], (size_t)p->gridDims[3] + 1))) break;
                if ((r = alloc(c, &net->gridSlot[b], n))) break;
                r = fill_i32(c, net->gridCounts[b], 0, (long long)p->gridDims[3] + 1);
            }
            if (r) break;
            hipLaunchKernelGGL(k_bounding_radius, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, fil->segLength,
                               p->actinRadius, net->boundR, n);
            if ((r = check_launch(c, "k_bounding_radius"))) break;
        }
        if (L > 0) {
            if ((r = alloc(c, &net->offA, (size_t)n + 1))) break;
            if ((r = alloc(c, &net->linkPack, (size_t)LINK_PACK * n))) break;
            if ((r = alloc(c, &net->offB, (size_t)n + 1))) break;
            if ((r = alloc(c, &net->idxA, L))) break;
            if ((r = alloc(c, &net->idxB, L))) break;
            if ((r = build_link_csr(net))) break;
        }
        if (p->formOn && (r = form_setup(net))) break;
        if (batch) {
            net->R = (int)batch->reps.size();
            if ((r = alloc(c, &net->reps, batch->reps.size()))) break;
            if ((r = alloc(c, &net->rodRep, (size_t)n))) break;
            if ((r = alloc(c, &net->linkRep, (size_t)(L > 0 ? L : 1)))) break;
            r = check_hip(c, hipMemcpy(net->reps, batch->reps.data(), sizeof(RepInfo) * batch->reps.size(),
                                       hipMemcpyHostToDevice), "upload replicas");
            if (!r) r = check_hip(c, hipMemcpy(net->rodRep, batch->rodRep.data(), sizeof(int) * (size_t)n,
                                               hipMemcpyHostToDevice), "upload rod replicas");
            if (!r && L > 0)
                r = check_hip(c, hipMemcpy(net->linkRep, batch->linkRep.data(), sizeof(int) * (size_t)L,
                                           hipMemcpyHostToDevice), "upload link replicas");
            if (r) break;
        }
        // reserve the sort scratch now: nothing may allocate inside a captured step
        long long keys = p->gridOn ? (long long)p->gridDims[3] : 1;
        if (keys < n) keys = n;
        if ((r = sort_scratch_reserve(c, net->sort, keys, n > L ? n : L))) break;
        r = sbx_sync(c);
    } while (0);
    if (r) {
        release(net);
        return r;
    }
    *out = net;
    return SBX_OK;
}

}  // namespace

extern "C" {

int sbx_net_create(sbx_ctx *c, const sbx_filaments *fil, const sbx_crosslinks *xl, const sbx_net_params *p,
                   sbx_net **out) {
    return net_create(c, fil, xl, p, nullptr, out);
}

void sbx_net_destroy(sbx_net *net) { release(net); }

int sbx_net_attach_motors(sbx_net *net, const sbx_motors *m, const sbx_motor_params *p) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, m && p, "net_attach_motors: null argument");
    SBX_REQUIRE(c, !net->hasMotors, "net_attach_motors: motors already attached");
    SBX_REQUIRE(c, !net->reps, "net_attach_motors: a batch steps filament networks only");
    SBX_REQUIRE(c, net->p.gridOn && !net->p.periodic,
                "net_attach_motors: the binding grid needs gridOn with the clamped (non-periodic) grid");
    const int n = net->n, nM = m->nM;
    SBX_REQUIRE(c, nM > 0, "net_attach_motors: need nM > 0");
    SBX_REQUIRE(c, (long long)n + nM < (1LL << 31), "net_attach_motors: too many bodies");
    SBX_REQUIRE(c, m->coord && m->uVec && m->yVec && m->zVec && m->end1 && m->end2 && m->segLength && m->bTransGam &&
                       m->bRotGam && m->brownTransScale && m->brownRotScale && m->forceSum && m->torqueSum,
                "net_attach_motors: null motor body array");
    SBX_REQUIRE(c, m->head && m->hUVec && m->rodUVec && m->boundSeg && m->cooldown && m->stats && m->capStats &&
                       m->nucleotideState && m->forceDotPlace && m->bindArc && m->forceDotFil && m->forceMag &&
                       m->forceDotHist && m->bondData && m->reachSeg && m->reachCount,
                "net_attach_motors: null motor kinetic array");
    SBX_REQUIRE(c, m->nD >= 0 && m->nBb >= 0 && m->nMD >= 0, "net_attach_motors: negative counts");
    if (m->nD > 0) SBX_REQUIRE(c, m->motorA && m->motorB && m->parallel, "net_attach_motors: null dimer array");
    SBX_REQUIRE(c, p->nodeMode >= 0 && p->nodeMode <= 2, "net_attach_motors: nodeMode must be 0, 1 or 2");
    if (m->nMD > 0) {
        SBX_REQUIRE(c, m->nBb > 0, "net_attach_motors: tethered items without backbones or nodes");
        SBX_REQUIRE(c, m->headBackboneSlot && m->tetherMotor && m->miniData, "net_attach_motors: null tether array");
        if (p->nodeMode)
            SBX_REQUIRE(c, m->radial && m->attachCoeffK, "net_attach_motors: node shell without radial/attachCoeffK");
        else
            SBX_REQUIRE(c, m->attachAxial, "net_attach_motors: null attachAxial");
    }
    if (m->nBb > 0)
        SBX_REQUIRE(c, m->bbCoord && m->bbUVec && m->bbYVec && m->bbZVec && m->bbEnd1 && m->bbEnd2 &&
                           m->bbSegLength && m->bbTransGam && m->bbRotGam && m->bbBrownTransScale &&
                           m->bbBrownRotScale && m->bbInvDragY,
                    "net_attach_motors: null backbone array");
    SBX_REQUIRE(c, (m->bbForceSum == nullptr) == (m->bbTorqueSum == nullptr),
                "net_attach_motors: bbForceSum and bbTorqueSum must be both set or both null");
    SBX_REQUIRE(c, p->jointParams[0] != 0.0f && p->xbParams[3] != 0.0f && p->kinParams[5] != 0.0f,
                "net_attach_motors: zero dt or kT in the motor parameters");
    if (m->nD > 0) SBX_REQUIRE(c, p->dimerParams[0] != 0.0f, "net_attach_motors: zero dimer dt");
    if (m->nMD > 0) SBX_REQUIRE(c, p->miniParams[0] != 0.0f, "net_attach_motors: zero tether dt");
    int r = set_device(c);
    if (r) return r;
    const int cap = n + nM;
    // the grid now bins filaments + heads: regrow the per-body grid arrays
    int *bodyCell = nullptr, *cellContents = nullptr, *ownerStore = nullptr, *ownerSlot = nullptr,
        *segHeadCnt = nullptr, *segHeads = nullptr, *bbOff = nullptr, *bbList = nullptr, *dimerOf = nullptr;
    do {
        if ((r = alloc(c, &bodyCell, cap))) break;
        if ((r = alloc(c, &cellContents, cap))) break;
        if ((r = alloc(c, &ownerStore, cap))) break;
        if ((r = alloc(c, &ownerSlot, cap))) break;
        if ((r = alloc(c, &segHeadCnt, (size_t)n))) break;
        if ((r = alloc(c, &segHeads, (size_t)n * SEG_HEADS))) break;
        if ((r = fill_i32(c, segHeadCnt, 0, n))) break;  // k_rod keeps it zero between steps
        if ((r = alloc(c, &bbOff, (size_t)m->nBb + 1))) break;
        if ((r = alloc(c, &bbList, m->nMD))) break;
        if ((r = alloc(c, &dimerOf, nM))) break;
        if ((r = fill_i32(c, dimerOf, -1, nM))) break;
        if (m->nD > 0) {
            hipLaunchKernelGGL(k_dimer_of, dim3(blocks_for(m->nD, 256)), dim3(256), 0, c->stream, m->motorA,
                               m->motorB, m->nD, dimerOf);
            if ((r = check_launch(c, "k_dimer_of"))) break;
        }
        long long keys = net->p.gridDims[3];
        if (keys < n) keys = n;
        if (keys < m->nBb) keys = m->nBb;
        if ((r = sort_scratch_reserve(c, net->sort, keys, cap))) break;
    } while (0);
    if (r) {
        void *bufs[] = {bodyCell, cellContents, ownerStore, ownerSlot, segHeadCnt, segHeads, bbOff, bbList, dimerOf};
        for (void *b : bufs)
            if (b) hipFree(b);
        return r;
    }
    hipStreamSynchronize(c->stream);
    hipFree(net->bodyCell);
    hipFree(net->cellContents);
    net->bodyCell = bodyCell;
    net->lastCell = nullptr;
    net->cellContents = cellContents;
    net->csrCont[0] = cellContents;
    net->ownerStore = ownerStore;
    net->ownerSlot = ownerSlot;
    net->segHeadCnt = segHeadCnt;
    net->segHeads = segHeads;
    net->bbOff = bbOff;
    net->bbList = bbList;
    net->dimerOf = dimerOf;
    net->mo = *m;
    net->mp = *p;
    net->kp = kin_params(p->kinParams);
    net->np = nuc_params(p->nucParams);
    net->xp = xb_params(p->xbParams);
    net->jp = joint_params(p->jointParams);
    net->dp = dimer_params(p->dimerParams);
    net->mip = mini_params(p->miniParams);
    const float *gp = net->p.gridParams;
    net->rg = {gp[0], gp[1], gp[2], gp[4], net->p.gridDims[0], net->p.gridDims[1], net->p.gridDims[2]};
    net->hasMotors = 1;
    net->rsValid = 0;  // (reach skin) new motor parameters: rebuild before the next search
    hipLaunchKernelGGL(k_owner_view, dim3(blocks_for(cap, 256)), dim3(256), 0, c->stream, ownerStore, ownerSlot, n,
                       nM);
    if ((r = check_launch(c, "k_owner_view"))) return r;
    // tethers are static (MiniFilamentSystem.csrBuild keys headBackboneSlot): build the backbone CSR once
    if (m->nBb > 0 && (r = counting_sort(c, net->sort, m->headBackboneSlot, m->nMD, m->nBb, m->nMD, bbOff, bbList)))
        return r;
    destroy_graphs(net);  // a captured step no longer matches the launch sequence
    return sbx_sync(c);
}

}  // extern "C"

namespace {

// steps step0 .. step0 + nSteps - 1 with no formation pass among them
int step_range(sbx_net *net, int step0, int nSteps) {
    sbx_ctx *c = net->ctx;
    if (nSteps == 0) return SBX_OK;
    int r = set_device(c);
    if (r) return r;
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, c->stream, net->dStep, step0);
    if ((r = check_launch(c, "k_set_step"))) return r;
    // Eager launches use stepOff = done - advanced, where `advanced` counts the steps already added
    // to the device counter by graph replays (each replay ends with k_add_step(G)).
    int done = 0, advanced = 0;
    if (!net->derivedValid) {  // bin and pack the caller's pose once, eagerly (k_rod keeps them after)
        if (net->p.gridOn && !net->hasMotors) {
            const int nk = net->p.gridDims[3];
            // both histograms start over: the one k_rod took may sit under the other buffer index
            // (sbx_net_publish moves the pose to buffer 0), and the caller may have moved bodies
            for (int b = 0; b < 2; b++)
                if ((r = fill_i32(c, net->gridCounts[b], 0, (long long)nk + 1))) return r;
            if (net->reps) {
                hipLaunchKernelGGL(k_body_cell_batch, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream,
                                   net->coordB[net->cur], net->reps, net->rodRep, net->bodyCell, net->n);
                if ((r = check_launch(c, "k_body_cell_batch"))) return r;
            } else if ((r = launch_body_cell(c, net->coordB[net->cur], net->n, net->n, net->geom, net->bodyCell))) {
                return r;
            }
            if (net->skinMode == SBX_SKIN_REBUILD &&
                (r = count_keys(c, net->bodyCell, net->n, nk, net->gridCounts[net->cur], net->gridSlot[net->cur])))
                return r;
        }
        if (net->p.gridOn && !net->hasMotors) {
            hipLaunchKernelGGL(k_pack_bodies, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream,
                               net->coordB[net->cur], net->boundR, net->bodyPack[net->cur], net->n);
            if ((r = check_launch(c, "k_pack_bodies"))) return r;
        }
        if (net->L > 0) {
            hipLaunchKernelGGL(k_pack_links, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream,
                               net->coordB[net->cur], net->uB[net->cur], net->fil.end1, net->fil.bTransGam,
                               net->linkPack, net->n);
            if ((r = check_launch(c, "k_pack_links"))) return r;
        }
        net->derivedValid = 1;
    }
    if (net->L > 0 && !net->countsValid) {  // creation, publish or a topology change: count afresh
        hipLaunchKernelGGL(k_link_counts, dim3(blocks_for(net->n, 256)), dim3(256), 0, c->stream, net->offA,
                           net->idxA, net->offB, net->idxB, net->xl.linkState, net->xl.activeLinkCount, net->n);
        if ((r = check_launch(c, "k_link_counts"))) return r;
        net->countsValid = 1;
    }
    const bool units = net->useGraph && !net->profiling && overlapped(net) && net->skinMode == SBX_SKIN_REBUILD &&
                       !net->hasMotors;
    if (units) {
        // one 1-step graph per pose buffer: [broad phase of the previous step | sort | mechanics] on
        // three streams with no wait between the chains (each hazard's first half is in the previous
        // graph, complete before this one starts), then the step advance
        // replays read the other buffer's CSR: it must have been built
        while (nSteps - done > 0 && !net->csrValid[1 - net->cur]) {
            if ((r = enqueue_step(net, done - advanced))) return r;
            done++;
        }
        while (nSteps - done > 0) {
            const int p = net->cur;
            if (!net->unitExec[p]) {
                int *savedCell = net->bodyCell, *savedNext = net->cellNext, *savedLast = net->lastCell;
                const int savedLastCsr = net->lastCsr, savedPending = net->bpPending;
                if ((r = join_side(net))) return r;
                cap_begin(net);
                if ((r = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal),
                                   "begin capture")))
                    return r;
                net->graphSteps = 1;
                for (int b = 0; b < 2; b++) net->joinLive[b] = net->placedLive[b] = 0;
                net->bpPending = 1;
                net->bpPendingBuf = 1 - p;
                int er = enqueue_step(net, 0, 0, true);
                if (!er) er = join_side(net);
                if (!er) hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, 1);
                hipGraph_t g = nullptr;
                hipError_t ee = hipStreamEndCapture(c->stream, &g);
                net->graphSteps = 0;
                net->evCaptured = 1;
                if (!er) er = cap_verdict(net);
                for (int b = 0; b < 2; b++) net->joinLive[b] = net->placedLive[b] = 0;
                net->unitLastCell[p] = net->lastCell;
                net->cur = p;
                net->bodyCell = savedCell;
                net->cellNext = savedNext;
                net->lastCell = savedLast;
                net->lastCsr = savedLastCsr;
                net->bpPending = savedPending;
                if (er) {
                    if (g) hipGraphDestroy(g);
                    return er;
                }
                if ((r = check_hip(c, ee, "end capture"))) return r;
                net->unitGraph[p] = g;
                if ((r = check_hip(c, hipGraphInstantiate(&net->unitExec[p], g, nullptr, nullptr, 0),
                                   "graph instantiate")))
                    return r;
            }
            if (done != advanced) {  // a replay reads its step from the device counter: bring it up to date
                hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, done - advanced);
                if ((r = check_launch(c, "k_add_step"))) return r;
                advanced = done;
            }
            if ((r = join_side(net))) return r;
            if ((r = check_hip(c, hipGraphLaunch(net->unitExec[p], c->stream), "graph launch"))) return r;
            net->needFork = 1;
            net->evCaptured = 1;
            // the replay ran the pending broad phase (or re-ran the last one), sorted and stepped pose
            // buffer p and left its broad phase; the host mirrors its pointer swaps
            net->csrValid[0] = net->csrValid[1] = 1;
            net->bpPending = 1;
            net->bpPendingBuf = p;
            net->lastCell = net->unitLastCell[p];
            int *t = net->bodyCell;
            net->bodyCell = net->cellNext;
            net->cellNext = t;
            net->cur = 1 - p;
            done++;
            advanced++;
        }
    } else if (net->useGraph && !net->profiling) {
        // one captured graph = graphLen steps (even: the pose ping-pong returns to its start buffers) +
        // step advance; a longer graph drains the stream pipeline less often
        const int G = net->graphLen;
        if (!net->graphExec) {
            int savedCur = net->cur;
            // side-stream work of earlier eager steps is joined before the graph (which joins its own)
            if ((r = join_side(net))) return r;
            cap_begin(net);
            if ((r = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
                return r;
            net->graphSteps = G;
            int er = SBX_OK;
            for (int k = 0; k < G && !er; k++) er = enqueue_step(net, k, k);
            if (!er) er = join_side(net);
            net->graphLastCsr = net->lastCsr;
            net->graphLastCell = net->lastCell;
            if (!er) hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, G);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(c->stream, &g);
            net->graphSteps = 0;
            net->evCaptured = 1;
            if (!er) er = cap_verdict(net);
            net->cur = savedCur;
            net->graphCur = savedCur;
            if (er) {
                if (g) hipGraphDestroy(g);
                return er;
            }
            if ((r = check_hip(c, ee, "end capture"))) return r;
            net->graph = g;
            if ((r = check_hip(c, hipGraphInstantiate(&net->graphExec, g, nullptr, nullptr, 0), "graph instantiate")))
                return r;
        }
        if (net->cur != net->graphCur && nSteps - done > 0) {
            if ((r = rs_maybe_rebuild(net)) || (r = fs_maybe_rebuild(net))) return r;
            if ((r = enqueue_step(net, done - advanced))) return r;
            if ((r = rs_advance(net, 1))) return r;
            done++;
        }
        while (nSteps - done >= G) {
            if (done != advanced) {  // a replay reads its first step from the device counter: bring it up to date
                hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, done - advanced);
                if ((r = check_launch(c, "k_add_step"))) return r;
                advanced = done;
            }
            if ((r = join_side(net))) return r;
            if ((r = rs_maybe_rebuild(net)) || (r = fs_maybe_rebuild(net))) return r;
            if ((r = check_hip(c, hipGraphLaunch(net->graphExec, c->stream), "graph launch"))) return r;
            if ((r = rs_advance(net, G))) return r;
            net->needFork = 1;  // the replay's side work is joined into it; later side work forks anew
            net->evCaptured = 1;
            net->lastCsr = net->graphLastCsr;
            net->lastCell = net->graphLastCell;
            if (net->p.gridOn && !net->hasMotors && net->skinMode == SBX_SKIN_REBUILD)
                net->csrValid[0] = net->csrValid[1] = 1;  // its two steps sorted both CSR buffers
            done += G;
            advanced += G;  // G is even: net->cur is unchanged by a replay
        }
    }
    for (; done < nSteps; done++) {
        if ((r = rs_maybe_rebuild(net)) || (r = fs_maybe_rebuild(net))) return r;
        if ((r = enqueue_step(net, done - advanced))) return r;
        if ((r = rs_advance(net, 1))) return r;
    }
    if (net->bpPending) {  // the last replay's broad phase: the slices of the call's last pose
        const bool ovl = overlapped(net);
        // a replay records none of its events outside the graph: the side stream follows the replay
        // (which sorted this CSR) through a fork from the main stream, not through evPlaced
        if ((r = settle_events(net))) return r;
        if (ovl && net->needFork) {
            if ((r = ev_record(net, net->evFork, c->stream, "record evFork"))) return r;
            if ((r = ev_wait(net, net->side, net->evFork, "wait evFork"))) return r;
            if ((r = ev_wait(net, net->sortS, net->evFork, "wait evFork"))) return r;
            net->needFork = 0;
        }
        if ((r = enqueue_bp(net, net->bpPendingBuf, ovl ? net->side : c->stream, ovl, true))) return r;
        net->bpPending = 0;
    }
    // the caller's stream covers the whole call: join the side streams
    if ((r = join_side(net))) return r;
    if (net->profiling) return drain_timers(net);
    return SBX_OK;
}

}  // namespace

extern "C" {

int sbx_net_step(sbx_net *net, int step0, int nSteps) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSteps >= 0, "net_step: nSteps < 0");
    if (!net->formOn && !net->turnOn) return step_range(net, step0, nSteps);
    SBX_REQUIRE(c, step0 >= 0, "net_step: cadence steps count from 0 (step0 < 0)");
    const int end = step0 + nSteps;
    int r = set_device(c);
    if (r) return r;
    // the cadence passes run between captured/eager step ranges: turnover first, then formation, then
    // the step (FullSystemDemoHarness.cpuStep order)
    for (int t = step0; t < end;) {
        int next = end;
        if (net->turnOn) {
            const int every = net->tp.interval;
            if (t % every == 0 && (r = run_turnover(net, t))) return r;
            const int nt = (t / every + 1) * every;
            if (nt < next) next = nt;
        }
        if (net->formOn) {
            const int every = net->p.formInt;
            if (t % every == 0 && (r = run_formation(net, t))) return r;
            const int nt = (t / every + 1) * every;
            if (nt < next) next = nt;
        }
        if ((r = step_range(net, t, next - t))) return r;
        t = next;
    }
    return SBX_OK;
}

int sbx_net_set_turnover(sbx_net *net, const sbx_turnover *t, const sbx_turnover_params *p) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    if (!t) {
        net->turnOn = 0;
        net->form.filState = nullptr;
        return SBX_OK;
    }
    SBX_REQUIRE(c, p, "net_set_turnover: null params");
    SBX_REQUIRE(c, !net->reps, "net_set_turnover: turnover is not batched");
    SBX_REQUIRE(c, net->skinMode == SBX_SKIN_REBUILD, "net_set_turnover: not with the skin broad phase");
    SBX_REQUIRE(c, t->C == net->n, "net_set_turnover: C %d != rods %d", t->C, net->n);
    SBX_REQUIRE(c, t->filState && t->monomerCount && t->seedNode && t->nucFrac && t->cofFrac && t->pool && t->ledger &&
                       t->stats,
                "net_set_turnover: null turnover array");
    SBX_REQUIRE(c, net->fil.end1NbrSlot && net->fil.end1NbrSide && net->fil.end2NbrSlot && net->fil.end2NbrSide,
                "net_set_turnover: the network has no chain topology arrays");
    SBX_REQUIRE(c, p->interval > 0, "net_set_turnover: interval must be > 0");
    int r = set_device(c);
    if (!r) r = turn_scratch_reserve(c, net->turnScratch, net->n);
    if (r) return r;
    net->turn = *t;
    net->tp = *p;
    net->turnOn = 1;
    return SBX_OK;
}

int sbx_net_form_stats(sbx_net *net, long long *out, int cap) {
    if (!net || !out || cap < 0) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    int st[4] = {0, 0, 0, 0};
    if (net->formOn) {
        int r = check_hip(c, hipMemcpyAsync(st, net->form.stats, sizeof(st), hipMemcpyDeviceToHost, c->stream),
                          "form stats");
        if (!r) r = check_hip(c, hipStreamSynchronize(c->stream), "form stats sync");
        if (r) return r;
    }
    long long v[4] = {net->formPasses, st[0], st[1], st[2]};
    int k = cap < 4 ? cap : 4;
    for (int i = 0; i < k; i++) out[i] = v[i];
    return k;
}

int sbx_net_publish(sbx_net *net) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    int r = SBX_OK;
    if (net->cur != 0) {
        size_t bytes = sizeof(float) * 3 * (size_t)net->n;
        r = check_hip(c, hipMemcpyAsync(net->coordB[0], net->coordB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                      "publish coord");
        if (!r) r = check_hip(c, hipMemcpyAsync(net->uB[0], net->uB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                              "publish uVec");
        if (!r) r = check_hip(c, hipMemcpyAsync(net->yB[0], net->yB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                              "publish yVec");
        if (!r) {
            net->cur = 0;
            // each pose buffer keeps its cell array (a captured step sorts the cells of pose buffer b
            // from the array it was captured with): the pose prep bins the moved pose into bodyCell
            if (net->p.gridOn && !net->hasMotors) {
                int *t = net->bodyCell;
                net->bodyCell = net->cellNext;
                net->cellNext = t;
            }
        }
    }
    // the caller may now edit its pose and link arrays: re-bin, re-pack, recount the active links and
    // (skin modes) rebuild next step
    net->derivedValid = 0;
    net->countsValid = 0;
    net->rsValid = 0;
    net->fsValid = 0;
    return r;
}

int sbx_net_topology_changed(sbx_net *net) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    net->countsValid = 0;
    return build_link_csr(net);
}

int sbx_net_candidates(sbx_net *net, int **candPartner, int **candCount, int **gridCellOffsets,
                       int **gridCellContents, int **bodyCell) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    if (candPartner) *candPartner = net->candPartner;
    if (candCount) *candCount = net->candCount;
    const int b = (net->p.gridOn && !net->hasMotors) ? net->lastCsr : 0;
    if (gridCellOffsets) *gridCellOffsets = b ? net->csrOff[1] : net->cellOffsets;
    if (gridCellContents) *gridCellContents = b ? net->csrCont[1] : net->cellContents;
    if (bodyCell) *bodyCell = net->lastCell ? net->lastCell : net->bodyCell;
    return SBX_OK;
}

int sbx_net_set_skin(sbx_net *net, int mode, const sbx_skin_params *sp, const float *modelWeights) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, mode == SBX_SKIN_REBUILD || mode == SBX_SKIN_FIXED || mode == SBX_SKIN_LEARNED,
                "net_set_skin: unknown mode %d", mode);
    SBX_REQUIRE(c, !net->hasMotors, "net_set_skin: the skin broad phase covers the filament-only step");
    SBX_REQUIRE(c, mode == SBX_SKIN_REBUILD || !net->reps, "net_set_skin: a batch runs the per-step rebuild");
    SBX_REQUIRE(c, mode == SBX_SKIN_REBUILD || !net->turnOn,
                "net_set_skin: turnover changes segment lengths (the skin's reach); not with turnover");
    if (mode != SBX_SKIN_REBUILD) {
        SBX_REQUIRE(c, net->p.gridOn, "net_set_skin: needs the grid (gridOn)");
        SBX_REQUIRE(c, sp, "net_set_skin: null params");
        SBX_REQUIRE(c, sp->skinDefault > 0.0f && sp->skinMin > 0.0f && sp->skinMax >= sp->skinMin &&
                           sp->skinDefault <= sp->skinMax && sp->margin > 0.0f && sp->listCap > 0,
                    "net_set_skin: need 0 < skinMin <= skinDefault <= skinMax, margin > 0, listCap > 0");
        SBX_REQUIRE(c, sp->horizon >= 0 && sp->extraTrigger >= 0 && sp->extraCap > 0,
                    "net_set_skin: need horizon >= 0, extraTrigger >= 0, extraCap > 0");
        SBX_REQUIRE(c, mode != SBX_SKIN_LEARNED || modelWeights, "net_set_skin: learned mode needs model weights");
        SBX_REQUIRE(c, net->n < (1 << 26), "net_set_skin: more than 2^26 bodies");
    }
    int r = set_device(c);
    if (r) return r;
    if ((r = check_hip(c, hipStreamSynchronize(c->stream), "net_set_skin sync"))) return r;
    // a captured step no longer matches the launch sequence
    destroy_graphs(net);
    net->skinMode = mode;
    net->derivedValid = 0;  // the next step prepares the packed pose the mode reads
    if (mode == SBX_SKIN_REBUILD) return SBX_OK;
    net->sp = *sp;
    const int n = net->n;
    // largest bounding radius -> coarse cell >= the largest list reach
    int *dmax = nullptr;
    if ((r = check_hip(c, hipMalloc(&dmax, sizeof(int)), "hipMalloc"))) return r;
    int hmax = 0;
    r = fill_i32(c, dmax, 0, 1);
    if (!r) {
        hipLaunchKernelGGL(k_max_radius, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, net->boundR, n, dmax);
        r = check_launch(c, "k_max_radius");
    }
    if (!r) r = check_hip(c, hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r) r = check_hip(c, hipStreamSynchronize(c->stream), "sync");
    hipFree(dmax);
    if (r) return r;
    float maxR;
    memcpy(&maxR, &hmax, sizeof(float));
    const GridGeom &f = net->geom;
    float fineCell = 1.0f / f.invCell;
    // the coarse cell of a rebuild is 2 maxR + cutoff + 2 maxSkin (+ slack), at least the fine cell;
    // allocate for the smallest (skinMin), check the largest (skinMax) still fits the box
    float floorCell = fmaxf(fineCell, 2.0f * maxR + f.cutoff + 2.0f * sp->skinMin + 2.0f * SKIN_SLACK);
    float topCell = fmaxf(fineCell, 2.0f * maxR + f.cutoff + 2.0f * sp->skinMax + 2.0f * SKIN_SLACK);
    long long cc;
    if (f.periodic) {
        SBX_REQUIRE(c, f.Lx == f.Ly && f.Ly == f.Lz, "net_set_skin: the periodic box must be cubic");
        SBX_REQUIRE(c, (int)floorf(f.Lx / topCell) >= 3,
                    "net_set_skin: box %g um too small for skinMax %g um (needs 3 coarse cells of %g um)", f.Lx,
                    sp->skinMax, topCell);
        long long nb = (long long)floorf(f.Lx / floorCell);
        cc = nb * nb * nb;
    } else {
        cc = (long long)ceilf(fineCell * f.nX / floorCell) * (long long)ceilf(fineCell * f.nY / floorCell) *
             (long long)ceilf(fineCell * f.nZ / floorCell);
    }
    SBX_REQUIRE(c, cc >= 1 && cc < (1LL << 30), "net_set_skin: coarse grid too large");
    net->maxR = maxR;
    net->cellFloor = floorCell;
    net->cCellsMax = (int)cc;
    // (re)allocate the skin state
    void *old[] = {net->cCell, net->cOff,   net->cCont,     net->refC,  net->refU,      net->skin,
                   net->cellSkin, net->cellHorizon, net->vlist, net->vcount, net->skinFlags, net->model,
                   net->cgDev,  net->cCellsDev, net->fsExtra};
    for (void *b : old)
        if (b) hipFree(b);
    net->cCell = net->cOff = net->cCont = net->vlist = net->vcount = net->skinFlags = net->cCellsDev = nullptr;
    net->fsExtra = nullptr;
    net->refC = net->refU = net->skin = net->cellSkin = net->cellHorizon = net->model = nullptr;
    net->cgDev = nullptr;
    do {
        if ((r = alloc(c, &net->cCell, n))) break;
        if ((r = alloc(c, &net->cOff, (size_t)cc + 1))) break;
        if ((r = alloc(c, &net->cCont, n))) break;
        if ((r = alloc(c, &net->refC, (size_t)3 * n))) break;
        if ((r = alloc(c, &net->refU, (size_t)3 * n))) break;
        if ((r = alloc(c, &net->skin, n))) break;
        if ((r = alloc(c, &net->cellSkin, (size_t)cc))) break;
        if ((r = alloc(c, &net->cellHorizon, (size_t)cc))) break;
        if ((r = alloc(c, &net->vlist, (size_t)n * sp->listCap))) break;
        if ((r = alloc(c, &net->vcount, n))) break;
        if ((r = alloc(c, &net->skinFlags, SKF_WORDS))) break;
        if ((r = alloc(c, &net->cgDev, 1))) break;
        if ((r = alloc(c, &net->cCellsDev, 1))) break;
        if ((r = alloc(c, &net->fsExtra, (size_t)sp->extraCap))) break;
        if ((r = host_words(net))) break;
        if ((r = fill_i32(c, net->cCellsDev, 0, 1))) break;
        if (mode == SBX_SKIN_LEARNED) {
            if ((r = alloc(c, &net->model, MLP_NPARAM))) break;
            if ((r = check_hip(c, hipMemcpyAsync(net->model, modelWeights, sizeof(float) * MLP_NPARAM,
                                                 hipMemcpyDeviceToDevice, c->stream), "copy model")))
                break;
        }
        long long keys = net->p.gridDims[3] > cc ? (long long)net->p.gridDims[3] : cc;
        if ((r = sort_scratch_reserve(c, net->sort, keys, n))) break;
        if ((r = fill_i32(c, net->skinFlags, 0, SKF_WORDS))) break;
        r = sbx_sync(c);
    } while (0);
    if (r) {
        net->skinMode = SBX_SKIN_REBUILD;
        return r;
    }
    net->rsHost[2] = -1;
    net->rsHost[3] = 0;
    net->fsValid = 0;  // the first step builds
    net->fsGen = 0;
    net->fsSince = 0;
    net->fsWhy[0] = net->fsWhy[1] = net->fsWhy[2] = 0;
    return SBX_OK;
}

int sbx_net_skin_stats(sbx_net *net, long long *out, int cap) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, out && cap >= 0, "net_skin_stats: null out");
    if (!net->skinFlags) return 0;
    int f[SKF_WORDS];
    int r = check_hip(c, hipMemcpyAsync(f, net->skinFlags, sizeof(f), hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r) r = check_hip(c, hipStreamSynchronize(c->stream), "sync");
    if (r) return r;
    const long long v[] = {f[SKF_REBUILDS], f[SKF_EXTRA],    f[SKF_BRUTE], f[SKF_SELECT],
                           f[SKF_ENTRIES],  f[SKF_OVERFLOW], f[SKF_SINCE], f[SKF_STEPS]};
    int k = 0;
    for (; k < (int)(sizeof(v) / sizeof(v[0])) && k < cap; k++) out[k] = v[k];
    return k;
}

int sbx_net_set_reach_skin(sbx_net *net, int mode, const sbx_reach_skin_params *sp, const float *modelWeights) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, mode == SBX_SKIN_REBUILD || mode == SBX_SKIN_FIXED || mode == SBX_SKIN_LEARNED,
                "net_set_reach_skin: unknown mode %d", mode);
    SBX_REQUIRE(c, net->hasMotors, "net_set_reach_skin: needs motors attached (the binding grid's search)");
    if (mode != SBX_SKIN_REBUILD) {
        SBX_REQUIRE(c, sp, "net_set_reach_skin: null params");
        SBX_REQUIRE(c, sp->skinDefault > 0.0f && sp->skinMin > 0.0f && sp->skinMax >= sp->skinMin &&
                           sp->skinDefault <= sp->skinMax && sp->margin > 0.0f && sp->cellFloor >= 0.0f,
                    "net_set_reach_skin: need 0 < skinMin <= skinDefault <= skinMax, margin > 0, cellFloor >= 0");
        SBX_REQUIRE(c, sp->horizon >= 0 && sp->extraTrigger >= 0 && sp->extraCap > 0,
                    "net_set_reach_skin: need horizon >= 0, extraTrigger >= 0, extraCap > 0");
        SBX_REQUIRE(c, mode != SBX_SKIN_LEARNED || modelWeights, "net_set_reach_skin: learned mode needs model weights");
        SBX_REQUIRE(c, net->n < (1 << RS_KEY_BITS), "net_set_reach_skin: more than 2^26 segments");
        SBX_REQUIRE(c, !net->reps, "net_set_reach_skin: a batch runs the per-step rebuild");
    }
    int r = set_device(c);
    if (!r) r = check_hip(c, hipStreamSynchronize(c->stream), "net_set_reach_skin sync");
    if (r) return r;
    destroy_graphs(net);  // a captured step no longer matches the launch sequence
    net->rsMode = 0;
    if (mode == SBX_SKIN_REBUILD) return SBX_OK;
    const int n = net->n;
    // the reach of the longest segment now (a segment grown longer later gets a smaller skin, or the list)
    std::vector<float> len((size_t)n);
    r = check_hip(c, hipMemcpy(len.data(), net->fil.segLength, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost),
                  "net_set_reach_skin lengths");
    if (r) return r;
    float maxL = 0.0f;
    for (float v : len)
        if (v > maxL) maxL = v;
    const float fine = 1.0f / net->rg.invCell;
    const float floorCell = sp->cellFloor > 0.0f ? sp->cellFloor : fine;
    const float spanX = fine * (float)net->rg.nX, spanY = fine * (float)net->rg.nY, spanZ = fine * (float)net->rg.nZ;
    // the smallest cell a rebuild can take (k_rs_geom), which sizes the CSR
    const float R0 = 0.5f * maxL + net->kp.myoColTol;
    const float minSkin = mode == SBX_SKIN_LEARNED ? std::min(sp->skinMin, sp->skinDefault) : sp->skinDefault;
    const float minCell = std::max(floorCell, R0 + minSkin + 2.0f * RS_SLACK);
    const long long cells = (long long)std::max(1, (int)ceilf(spanX / minCell)) *
                            std::max(1, (int)ceilf(spanY / minCell)) * std::max(1, (int)ceilf(spanZ / minCell));
    SBX_REQUIRE(c, cells < (1LL << 30), "net_set_reach_skin: %lld coarse cells at the smallest cell %g um "
                "(raise cellFloor or the skins)", cells, minCell);
    // (re)allocate
    void *old[] = {net->rsRef, net->rsRefU, net->rsSkin, net->rsCell, net->rsOff, net->rsCont, net->rsExtra,
                   net->rsWords, net->rsModel};
    for (void *p : old)
        if (p) hipFree(p);
    net->rsRef = net->rsRefU = net->rsSkin = net->rsModel = nullptr;
    net->rsCell = net->rsOff = net->rsCont = net->rsExtra = net->rsWords = nullptr;
    do {
        if ((r = alloc(c, &net->rsRef, 3 * (size_t)n))) break;
        if ((r = alloc(c, &net->rsRefU, 3 * (size_t)n))) break;
        if ((r = alloc(c, &net->rsSkin, (size_t)n))) break;
        if ((r = alloc(c, &net->rsCell, (size_t)n))) break;
        if ((r = alloc(c, &net->rsOff, (size_t)cells + 1))) break;
        if ((r = alloc(c, &net->rsCont, (size_t)n))) break;
        if ((r = alloc(c, &net->rsExtra, (size_t)sp->extraCap))) break;
        if ((r = alloc(c, &net->rsWords, RS_WORDS))) break;
        if (mode == SBX_SKIN_LEARNED) {
            if ((r = alloc(c, &net->rsModel, MLP_NPARAM))) break;
            r = check_hip(c, hipMemcpyAsync(net->rsModel, modelWeights, sizeof(float) * MLP_NPARAM,
                                            hipMemcpyDeviceToDevice, c->stream), "net_set_reach_skin model");
            if (r) break;
        }
        if ((r = host_words(net))) break;
        if ((r = fill_i32(c, net->rsWords, 0, RS_WORDS))) break;
        r = check_hip(c, hipStreamSynchronize(c->stream), "net_set_reach_skin sync");
    } while (0);
    if (r) return r;
    net->rsHost[0] = -1;
    net->rsHost[1] = 0;
    net->rsp = *sp;
    net->rsR0 = R0;
    net->rsFloor = floorCell;
    net->rsCellsMax = cells;
    net->rsMode = mode;
    net->rsValid = 0;  // the next step rebuilds
    net->rsGen = 0;
    net->rsSince = 0;
    net->rsWhy[0] = net->rsWhy[1] = net->rsWhy[2] = 0;
    net->rsSteps = 0;
    return SBX_OK;
}

int sbx_net_reach_skin_stats(sbx_net *net, long long *out, int cap) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, out && cap >= 0, "net_reach_skin_stats: null out");
    if (!net->rsWords) return 0;
    int w[RS_WORDS];
    int r = check_hip(c, hipMemcpyAsync(w, net->rsWords, sizeof(w), hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r) r = check_hip(c, hipStreamSynchronize(c->stream), "sync");
    if (r) return r;
    float cell = 0.0f;
    memcpy(&cell, &w[RS_GEOM + 3], sizeof(cell));
    const long long v[] = {w[RS_REBUILDS],   w[RS_EXTRA],     w[RS_BRUTE],     w[RS_SELECT],
                           llroundf(1000.0f * cell),
                           net->rsWhy[0],    net->rsWhy[1],   net->rsWhy[2],   net->rsSteps};
    int k = 0;
    for (; k < (int)(sizeof(v) / sizeof(v[0])) && k < cap; k++) out[k] = v[k];
    return k;
}

int sbx_net_skin_state(sbx_net *net, float **refCoord, float **skin, int **coarseCell, int **coarseOffsets,
                       int **coarseContents, int *nCoarse, float *coarseGrid) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    if (refCoord) *refCoord = net->refC;
    if (skin) *skin = net->skin;
    if (coarseCell) *coarseCell = net->cCell;
    if (coarseOffsets) *coarseOffsets = net->cOff;
    if (coarseContents) *coarseContents = net->cCont;
    GridGeom g{};
    int nc = 0;
    if (net->cgDev) {
        int r = check_hip(net->ctx, hipMemcpyAsync(&g, net->cgDev, sizeof(g), hipMemcpyDeviceToHost, net->ctx->stream),
                          "copy coarse grid");
        if (!r) r = check_hip(net->ctx, hipMemcpyAsync(&nc, net->cCellsDev, sizeof(int), hipMemcpyDeviceToHost,
                                                       net->ctx->stream), "copy coarse count");
        if (!r) r = check_hip(net->ctx, hipStreamSynchronize(net->ctx->stream), "sync");
        if (r) return r;
    }
    if (nCoarse) *nCoarse = nc;
    if (coarseGrid) {
        float v[8] = {g.xMin, g.yMin, g.zMin, g.invCell > 0 ? 1.0f / g.invCell : 0.0f, g.invCell,
                      (float)g.nX, (float)g.nY, (float)g.nZ};
        memcpy(coarseGrid, v, sizeof(v));
    }
    return SBX_OK;
}

int sbx_net_set_profiling(sbx_net *net, int on) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    int r = drain_timers(net);
    net->profiling = on ? 1 : 0;
    if (on)
        for (auto &t : net->timers) { t.ms = 0.0; t.launches = 0; }
    return r;
}

int sbx_net_time_broad_phase(sbx_net *net, int reps, float *usPerLaunch) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, usPerLaunch && reps > 0 && reps <= 10000, "net_time_broad_phase: need 0 < reps <= 10000 and an out");
    SBX_REQUIRE(c, net->p.gridOn && !net->hasMotors && net->skinMode == SBX_SKIN_REBUILD,
                "net_time_broad_phase: the per-step-rebuild broad phase only");
    const int b = net->lastCsr;
    SBX_REQUIRE(c, net->csrValid[b], "net_time_broad_phase: no grid built yet (step first)");
    int r = set_device(c);
    if (r) return r;
    if (net->bpPending) return fail(c, SBX_ERR_STATE, "net_time_broad_phase: a step is still in flight");
    // every stream idle: the launches below re-run the last step's broad phase on its grid, so the
    // candidate slices are rewritten with the same values
    if ((r = check_hip(c, hipStreamSynchronize(c->stream), "time bp sync"))) return r;
    if (net->side && (r = check_hip(c, hipStreamSynchronize(net->side), "time bp sync"))) return r;
    if (net->sortS && (r = check_hip(c, hipStreamSynchronize(net->sortS), "time bp sync"))) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((r = check_hip(c, hipEventCreate(&e0), "time bp event"))) return r;
    if ((r = check_hip(c, hipEventCreate(&e1), "time bp event"))) {
        hipEventDestroy(e0);
        return r;
    }
    r = ev_record(net, e0, c->stream, "time bp record");
    for (int k = 0; k < reps && !r; k++)
        r = launch_broad_phase_sorted(c, net->sortedPose[b], net->sortedCell[b], net->csrOff[b], net->csrCont[b],
                                      net->geom, net->reps, net->rodRep, net->bpGroup, net->bpStaged, net->n, net->n,
                                      net->candPartner, net->candCount, net->p.maxCand, c->stream);
    if (!r) r = ev_record(net, e1, c->stream, "time bp record");
    if (!r) r = check_hip(c, hipEventSynchronize(e1), "time bp sync");
    float ms = 0.0f;
    if (!r) r = check_hip(c, hipEventElapsedTime(&ms, e0, e1), "time bp elapsed");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!r) *usPerLaunch = 1.0e3f * ms / (float)reps;
    return r;
}

int sbx_net_kernel_times(sbx_net *net, const char **names, double *total_ms, long long *launches, int cap) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    int r = drain_timers(net);
    if (r) return r;
    int k = 0;
    for (; k < (int)net->timers.size() && k < cap; k++) {
        if (names) names[k] = net->timers[k].name;
        if (total_ms) total_ms[k] = net->timers[k].ms;
        if (launches) launches[k] = net->timers[k].launches;
    }
    return k;
}

int sbx_net_use_graph(sbx_net *net, int on) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    // captures longer than 2 steps are switched off: an 8-step capture ended the host process
    // (segmentation fault inside sbx_net_step) with no cause found in this library
    SBX_REQUIRE(net->ctx, on == 0 || on == 1 || on == 2, "net_use_graph: on = 0, 1 or 2 (steps per graph)");
    if (on) SBX_REQUIRE(net->ctx, net->ctx->stream != nullptr, "net_use_graph: needs a non-default stream");
    const int len = 2;
    if (on && len != net->graphLen) {  // a captured graph of another length no longer applies
        int r = check_hip(net->ctx, hipStreamSynchronize(net->ctx->stream), "net_use_graph sync");
        if (r) return r;
        destroy_graphs(net);
        net->graphLen = len;
    }
    net->useGraph = on ? 1 : 0;
    return SBX_OK;
}

int sbx_net_set_overlap(sbx_net *net, int on) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    int r = set_device(c);
    if (r) return r;
    if ((r = check_hip(c, hipStreamSynchronize(c->stream), "net_set_overlap sync"))) return r;
    if (net->side && (r = check_hip(c, hipStreamSynchronize(net->side), "net_set_overlap sync"))) return r;
    // a captured step no longer matches the launch sequence
    destroy_graphs(net);
    net->overlap = on ? 1 : 0;
    return SBX_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ launch shape and autotuner
namespace {

// a captured step no longer matches the launch sequence
int drop_graph(sbx_net *net) {
    sbx_ctx *c = net->ctx;
    int r = check_hip(c, hipStreamSynchronize(c->stream), "net shape sync");
    if (!r && net->side) r = check_hip(c, hipStreamSynchronize(net->side), "net shape sync");
    if (!r && net->sortS) r = check_hip(c, hipStreamSynchronize(net->sortS), "net shape sync");
    destroy_graphs(net);
    return r;
}

}  // namespace

extern "C" {

int sbx_capture_check(const int *ops, int nOps) {
    if (!ops || nOps < 0) return nOps == 0 ? -1 : SBX_ERR_ARG;
    return capture_check(ops, nOps);
}

int sbx_net_capture_log(const sbx_net *net, int *out, int cap) {
    if (!net || cap < 0 || (cap > 0 && !out)) return SBX_ERR_ARG;
    const int n = (int)net->capLog.size();
    for (int k = 0; k < n && k < cap; k++) out[k] = net->capLog[k];
    return n;
}

int sbx_net_set_motor_group(sbx_net *net, int lanes) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64,
                "net_set_motor_group: 4, 8, 16, 32 or 64 lanes per motor");
    int r = set_device(c);
    if (!r) r = drop_graph(net);
    if (r) return r;
    net->mbGroup = lanes;
    return SBX_OK;
}

int sbx_net_motor_group(const sbx_net *net) { return net ? net->mbGroup : SBX_ERR_ARG; }

int sbx_net_set_shape(sbx_net *net, int bpGroup, int bpStaged, int scanItems, int rodThreads) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, bpGroup == 2 || bpGroup == 4 || bpGroup == 8, "net_set_shape: bpGroup 2, 4 or 8");
    SBX_REQUIRE(c, bpStaged == 0 || ((bpStaged == 1 || bpStaged == 2) && bpGroup != 2),
                "net_set_shape: bpStaged 0, or 1 (LDS-staged) or 2 (row-split) with 4 or 8 lanes per body");
    SBX_REQUIRE(c, scanItems == 4 || scanItems == 8 || scanItems == 16, "net_set_shape: scanItems 4, 8 or 16");
    SBX_REQUIRE(c, rodThreads == 64 || rodThreads == 128 || rodThreads == 256 || rodThreads == RODS_PAIRED,
                "net_set_shape: rodThreads 64, 128, 256 or 512 (the noise-paired form)");
    int r = set_device(c);
    if (!r) r = drop_graph(net);
    if (r) return r;
    net->bpGroup = bpGroup;
    net->bpStaged = bpStaged;
    net->sort.scanItems = scanItems;
    net->rodThreads = rodThreads;
    return SBX_OK;
}

int sbx_net_tune(sbx_net *net, int step0, int trialSteps, int *advanced, int *shape, float *msPerStep) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, trialSteps >= 1 && advanced, "net_tune: trialSteps >= 1 and advanced required");
    int r = set_device(c);
    if (r) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((r = check_hip(c, hipEventCreate(&e0), "tune event"))) return r;
    if ((r = check_hip(c, hipEventCreate(&e1), "tune event"))) { hipEventDestroy(e0); return r; }
    const int groups[3] = {4, 2, 8}, items[3] = {8, 4, 16}, threads[3] = {256, 128, RODS_PAIRED};
    const int overlaps = net->side ? 2 : 1;  // with streams: the overlapped and the one-stream step
    int t = step0, best[5] = {net->bpGroup, net->bpStaged, net->sort.scanItems, net->rodThreads, net->overlap};
    float bestMs = -1.0f;
    for (int oi = 0; oi < overlaps && !r; oi++)
    for (int gi = 0; gi < 3 && !r; gi++)
    for (int si = 0; si < 3 && !r; si++)
        for (int ii = 0; ii < 3 && !r; ii++)
            for (int ti = 0; ti < 3 && !r; ti++) {
                if (si >= 1 && groups[gi] == 2) continue;  // staged and row-split forms take 4 or 8 lanes
                if ((r = sbx_net_set_shape(net, groups[gi], si, items[ii], threads[ti]))) break;
                net->overlap = oi == 0 ? 1 : 0;
                if ((r = sbx_net_step(net, t, 2))) break;  // warm (and capture, with graphs on)
                t += 2;
                if ((r = join_side(net))) break;
                if ((r = ev_record(net, e0, c->stream, "tune record"))) break;
                if ((r = sbx_net_step(net, t, trialSteps))) break;
                t += trialSteps;
                if ((r = join_side(net))) break;
                if ((r = ev_record(net, e1, c->stream, "tune record"))) break;
                if ((r = check_hip(c, hipEventSynchronize(e1), "tune sync"))) break;
                float ms = 0.0f;
                if ((r = check_hip(c, hipEventElapsedTime(&ms, e0, e1), "tune elapsed"))) break;
                ms /= (float)trialSteps;
                if (bestMs < 0.0f || ms < bestMs) {
                    bestMs = ms;
                    best[0] = groups[gi]; best[1] = si; best[2] = items[ii]; best[3] = threads[ti];
                    best[4] = net->overlap;
                }
            }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *advanced = t - step0;
    if (r) return r;
    if ((r = sbx_net_set_shape(net, best[0], best[1], best[2], best[3]))) return r;
    net->overlap = best[4];
    if (net->hasMotors) {  // then the motor stage's group size, at the chosen shape
        // the reach skin's search also tries narrow groups (few candidates per head); the per-step grid walk
        // runs 4 and 8 as 16, so every network takes the same number of tuning steps
        const int gs[5] = {4, 8, 16, 32, 64};
        const int nGs = 5;
        int bestG = net->mbGroup;
        float bestGm = -1.0f;
        hipEvent_t g0 = nullptr, g1 = nullptr;
        if ((r = check_hip(c, hipEventCreate(&g0), "tune event"))) return r;
        if ((r = check_hip(c, hipEventCreate(&g1), "tune event"))) { hipEventDestroy(g0); return r; }
        for (int gi = 0; gi < nGs && !r; gi++) {
            if ((r = sbx_net_set_motor_group(net, gs[gi]))) break;
            if ((r = sbx_net_step(net, t, 2))) break;
            t += 2;
            if ((r = join_side(net))) break;
            if ((r = ev_record(net, g0, c->stream, "tune record"))) break;
            if ((r = sbx_net_step(net, t, trialSteps))) break;
            t += trialSteps;
            if ((r = join_side(net))) break;
            if ((r = ev_record(net, g1, c->stream, "tune record"))) break;
            if ((r = check_hip(c, hipEventSynchronize(g1), "tune sync"))) break;
            float ms = 0.0f;
            if ((r = check_hip(c, hipEventElapsedTime(&ms, g0, g1), "tune elapsed"))) break;
            ms /= (float)trialSteps;
            if (bestGm < 0.0f || ms < bestGm) { bestGm = ms; bestG = gs[gi]; }
        }
        hipEventDestroy(g0);
        hipEventDestroy(g1);
        *advanced = t - step0;
        if (r) return r;
        if ((r = sbx_net_set_motor_group(net, bestG))) return r;
        if (bestGm < bestMs) bestMs = bestGm;
    }
    if (shape)
        for (int k = 0; k < 5; k++) shape[k] = best[k];
    if (msPerStep) *msPerStep = bestMs;
    return SBX_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batched step
struct sbx_batch {
    sbx_ctx *ctx = nullptr;
    std::vector<sbx_filaments> fil;  // the caller's replicas
    std::vector<sbx_crosslinks> xl;
    BatchLayout lay;
    sbx_filaments cf{};               // the concatenated stores (library-owned)
    sbx_crosslinks cx{};
    std::vector<void *> owned;
    sbx_net *net = nullptr;
    int n = 0, L = 0, cells = 0;
    bool forces = false;
};

namespace {

// rod and link ids of a replica's topology become ids in the concatenated stores (-1 stays -1)
__global__ void __launch_bounds__(256) k_offset_ids(int *__restrict__ v, int count, int base) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k < count && v[k] >= 0) v[k] += base;
}

template <typename T>
int batch_alloc(sbx_batch *b, T **p, size_t count) {
    int r = check_hip(b->ctx, hipMalloc((void **)p, sizeof(T) * (count ? count : 1)), "sbx_batch alloc");
    if (!r) b->owned.push_back((void *)*p);
    return r;
}

// rows [base, base + count) of a stride-`total` planar array with `comps` planes <-> a replica's
// planar array (stride count); toBatch picks the direction
template <typename T>
int copy_planes(sbx_ctx *c, T *batchArr, long long total, long long base, T *repArr, long long count, int comps,
                bool toBatch) {
    for (int k = 0; k < comps; k++) {
        T *bp = batchArr + k * total + base;
        T *rp = repArr + k * count;
        int r = check_hip(c, hipMemcpyAsync(toBatch ? (void *)bp : (void *)rp, toBatch ? (void *)rp : (void *)bp,
                                            sizeof(T) * (size_t)count, hipMemcpyDeviceToDevice, c->stream),
                          "batch copy");
        if (r) return r;
    }
    return SBX_OK;
}

bool same_bytes(const void *x, const void *y, size_t bytes) { return std::memcmp(x, y, bytes) == 0; }

// the rod physics every replica of a batch shares
bool same_physics(const sbx_net_params &a, const sbx_net_params &b) {
    return same_bytes(a.params, b.params, sizeof(a.params)) &&
           same_bytes(a.chainParams, b.chainParams, sizeof(a.chainParams)) && a.chainOn == b.chainOn &&
           a.maxCand == b.maxCand && a.actinRadius == b.actinRadius && a.rng == b.rng && a.stream == b.stream;
}

// the crosslink physics every replica with crosslinks shares
bool same_link_physics(const sbx_net_params &a, const sbx_net_params &b) {
    return same_bytes(a.xlParams, b.xlParams, sizeof(a.xlParams)) &&
           same_bytes(a.offParams, b.offParams, sizeof(a.offParams)) &&
           same_bytes(a.torsionParams, b.torsionParams, sizeof(a.torsionParams)) && a.unbindOn == b.unbindOn;
}

void batch_release(sbx_batch *b) {
    if (!b) return;
    if (b->net) release(b->net);
    if (b->ctx) hipSetDevice(b->ctx->device);
    for (void *q : b->owned) hipFree(q);
    delete b;
}

// the state a step writes, between the concatenated stores and the replicas' own
int batch_copy_state(sbx_batch *b, bool toBatch, bool everything) {
    sbx_ctx *c = b->ctx;
    const int R = (int)b->fil.size();
    int r = SBX_OK;
    for (int q = 0; q < R && !r; q++) {
        const sbx_filaments &f = b->fil[q];
        const RepInfo &ri = b->lay.reps[q];
        const long long N = b->n, base = ri.rodBase, nr = f.n;
        float *const pairs3[] = {b->cf.coord, f.coord, b->cf.uVec, f.uVec, b->cf.yVec, f.yVec,
                                 b->cf.zVec, f.zVec, b->cf.end1, f.end1, b->cf.end2, f.end2};
        for (int k = 0; k < 6 && !r; k++) r = copy_planes(c, pairs3[2 * k], N, base, pairs3[2 * k + 1], nr, 3, toBatch);
        if (!r && b->forces) {
            r = copy_planes(c, b->cf.forceSum, N, base, f.forceSum, nr, 3, toBatch);
            if (!r) r = copy_planes(c, b->cf.torqueSum, N, base, f.torqueSum, nr, 3, toBatch);
        }
        if (!r && everything) {
            float *const pf[] = {b->cf.segLength, f.segLength, b->cf.brownTransScale, f.brownTransScale,
                                 b->cf.brownRotScale, f.brownRotScale};
            for (int k = 0; k < 3 && !r; k++) r = copy_planes(c, pf[2 * k], N, base, pf[2 * k + 1], nr, 1, toBatch);
            if (!r) r = copy_planes(c, b->cf.bTransGam, N, base, f.bTransGam, nr, 3, toBatch);
            if (!r) r = copy_planes(c, b->cf.bRotGam, N, base, f.bRotGam, nr, 3, toBatch);
            int *const pi[] = {b->cf.end1NbrSlot, f.end1NbrSlot, b->cf.end1NbrSide, f.end1NbrSide,
                               b->cf.end2NbrSlot, f.end2NbrSlot, b->cf.end2NbrSide, f.end2NbrSide};
            for (int k = 0; k < 4 && !r; k++)
                if (pi[2 * k + 1]) r = copy_planes(c, pi[2 * k], N, base, pi[2 * k + 1], nr, 1, toBatch);
        }
        if (r || b->L == 0) continue;
        const sbx_crosslinks &x = b->xl[q];
        const long long Lt = b->L, lb = ri.linkBase, lr = x.nLinks;
        if (!r && x.activeLinkCount) r = copy_planes(c, b->cx.activeLinkCount, N, base, x.activeLinkCount, nr, 1, toBatch);
        if (lr == 0) continue;
        if (!r) r = copy_planes(c, b->cx.linkState, Lt, lb, x.linkState, lr, 1, toBatch);
        if (!r) r = copy_planes(c, b->cx.strainPlace, Lt, lb, x.strainPlace, lr, 1, toBatch);
        if (!r) r = copy_planes(c, b->cx.torqueMagPlace, Lt, lb, x.torqueMagPlace, lr, 1, toBatch);
        if (!r) r = copy_planes(c, b->cx.strainHist, Lt * SBX_STRAIN_WIN, lb * SBX_STRAIN_WIN, x.strainHist,
                                lr * SBX_STRAIN_WIN, 1, toBatch);
        if (!r) r = copy_planes(c, b->cx.torqueMagHist, Lt * SBX_TORQUE_WIN, lb * SBX_TORQUE_WIN, x.torqueMagHist,
                                lr * SBX_TORQUE_WIN, 1, toBatch);
        if (!r) r = copy_planes(c, b->cx.xlinkData, Lt * SBX_XL_STRIDE, lb * SBX_XL_STRIDE, x.xlinkData,
                                lr * SBX_XL_STRIDE, 1, toBatch);
        if (!r && everything) {
            float *const pf[] = {b->cx.loc1, x.loc1, b->cx.loc2, x.loc2};
            for (int k = 0; k < 2 && !r; k++) r = copy_planes(c, pf[2 * k], Lt, lb, pf[2 * k + 1], lr, 1, toBatch);
            int *const pi[] = {b->cx.linkFilA, x.linkFilA, b->cx.linkFilB, x.linkFilB, b->cx.linkOrientSame,
                               x.linkOrientSame};
            for (int k = 0; k < 3 && !r; k++) r = copy_planes(c, pi[2 * k], Lt, lb, pi[2 * k + 1], lr, 1, toBatch);
        }
    }
    return r;
}

}  // namespace

extern "C" {

int sbx_batch_create(sbx_ctx *c, int nRep, const sbx_filaments *fil, const sbx_crosslinks *xl,
                     const sbx_net_params *p, sbx_batch **out) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, out && fil && p && nRep >= 1, "batch_create: null argument or nRep < 1");
    *out = nullptr;
    long long n = 0, L = 0, cells = 0;
    const bool forces = fil[0].forceSum != nullptr;
    int linkRef = -1;  // first replica with crosslinks: the link physics of the batch
    for (int q = 0; q < nRep; q++) {
        const sbx_net_params &pq = p[q];
        SBX_REQUIRE(c, fil[q].n > 0, "batch_create: replica %d has no rods", q);
        SBX_REQUIRE(c, pq.gridOn && !pq.confineOn, "batch_create: replica %d: grid on, containment off", q);
        SBX_REQUIRE(c, same_physics(pq, p[0]), "batch_create: replica %d: physics differs from replica 0 (only "
                       "grid, box and seed may differ)", q);
        if (xl && xl[q].nLinks > 0) {
            if (linkRef < 0) linkRef = q;
            SBX_REQUIRE(c, same_link_physics(pq, p[linkRef]), "batch_create: replica %d: crosslink physics "
                           "differs from replica %d", q, linkRef);
        }
        SBX_REQUIRE(c, (fil[q].forceSum != nullptr) == forces, "batch_create: forceSum set for some replicas only");
        if (pq.chainOn)
            SBX_REQUIRE(c, fil[q].end1NbrSlot && fil[q].end1NbrSide && fil[q].end2NbrSlot && fil[q].end2NbrSide,
                        "batch_create: replica %d: chain on but topology arrays null", q);
        int r = check_grid(c, pq.gridParams, pq.gridDims, pq.periodic);
        if (r) return r;
        n += fil[q].n;
        L += xl ? xl[q].nLinks : 0;
        cells += pq.gridDims[3];
    }
    SBX_REQUIRE(c, n < (1LL << 30) && L < (1LL << 30) && cells < (1LL << 30), "batch_create: batch too large");
    int r = set_device(c);
    if (r) return r;
    sbx_batch *b = new (std::nothrow) sbx_batch();
    if (!b) return fail(c, SBX_ERR_NOMEM, "batch_create: host allocation");
    b->ctx = c;
    b->fil.assign(fil, fil + nRep);
    if (xl) b->xl.assign(xl, xl + nRep);
    else b->xl.assign(nRep, sbx_crosslinks{});
    b->n = (int)n;
    b->L = (int)L;
    b->cells = (int)cells;
    b->forces = forces;
    b->lay.rodRep.resize((size_t)n);
    b->lay.linkRep.resize((size_t)L);
    long long rb = 0, lb = 0, cb = 0;
    for (int q = 0; q < nRep; q++) {
        const sbx_net_params &pq = p[q];
        RepInfo ri;
        ri.g = grid_geom(pq.gridParams, pq.gridDims, pq.periodic, pq.periodic ? pq.boxLen : nullptr);
        ri.cellBase = (int)cb;
        ri.rodBase = (int)rb;
        ri.linkBase = (int)lb;
        ri.seed = pq.seed;
        b->lay.reps.push_back(ri);
        for (int i = 0; i < fil[q].n; i++) b->lay.rodRep[(size_t)(rb + i)] = q;
        for (int k = 0; k < b->xl[q].nLinks; k++) b->lay.linkRep[(size_t)(lb + k)] = q;
        rb += fil[q].n;
        lb += b->xl[q].nLinks;
        cb += pq.gridDims[3];
    }
    do {
        sbx_filaments &f = b->cf;
        f.n = (int)n;
        for (float **a : {&f.coord, &f.uVec, &f.yVec, &f.zVec, &f.end1, &f.end2, &f.bTransGam, &f.bRotGam})
            if (!r) r = batch_alloc(b, a, 3 * (size_t)n);
        for (float **a : {&f.segLength, &f.brownTransScale, &f.brownRotScale})
            if (!r) r = batch_alloc(b, a, (size_t)n);
        if (!r && forces) r = batch_alloc(b, &f.forceSum, 3 * (size_t)n);
        if (!r && forces) r = batch_alloc(b, &f.torqueSum, 3 * (size_t)n);
        for (int **a : {&f.end1NbrSlot, &f.end1NbrSide, &f.end2NbrSlot, &f.end2NbrSide})
            if (!r) r = batch_alloc(b, a, (size_t)n);
        if (r) break;
        for (int *a : {f.end1NbrSlot, f.end1NbrSide, f.end2NbrSlot, f.end2NbrSide})  // replicas without topology
            if ((r = fill_i32(c, a, -1, n))) break;
        if (r) break;
        sbx_crosslinks &x = b->cx;
        x.nLinks = (int)L;
        if (L > 0) {
            for (int **a : {&x.linkFilA, &x.linkFilB, &x.linkState, &x.strainPlace, &x.linkOrientSame,
                            &x.torqueMagPlace})
                if (!r) r = batch_alloc(b, a, (size_t)L);
            for (float **a : {&x.loc1, &x.loc2})
                if (!r) r = batch_alloc(b, a, (size_t)L);
            if (!r) r = batch_alloc(b, &x.strainHist, (size_t)L * SBX_STRAIN_WIN);
            if (!r) r = batch_alloc(b, &x.torqueMagHist, (size_t)L * SBX_TORQUE_WIN);
            if (!r) r = batch_alloc(b, &x.xlinkData, (size_t)L * SBX_XL_STRIDE);
            if (!r) r = batch_alloc(b, &x.activeLinkCount, (size_t)n);
            if (r) break;
            if ((r = fill_i32(c, x.activeLinkCount, 0, n))) break;
        }
        if ((r = batch_copy_state(b, true, true))) break;
        for (int q = 0; q < nRep && !r; q++) {  // ids into the concatenated stores
            const RepInfo &ri = b->lay.reps[q];
            const int nr = fil[q].n, lr = b->xl[q].nLinks;
            for (int *a : {f.end1NbrSlot, f.end2NbrSlot}) {
                hipLaunchKernelGGL(k_offset_ids, dim3(blocks_for(nr, 256)), dim3(256), 0, c->stream, a + ri.rodBase,
                                   nr, ri.rodBase);
            }
            if (lr > 0) {
                for (int *a : {x.linkFilA, x.linkFilB}) {
                    hipLaunchKernelGGL(k_offset_ids, dim3(blocks_for(lr, 256)), dim3(256), 0, c->stream,
                                       a + ri.linkBase, lr, ri.rodBase);
                }
            }
            r = check_launch(c, "k_offset_ids");
        }
        if (r) break;
        sbx_net_params pc = p[0];
        pc.gridDims[3] = (int)cells;
        if (linkRef >= 0) {
            std::memcpy(pc.xlParams, p[linkRef].xlParams, sizeof(pc.xlParams));
            std::memcpy(pc.offParams, p[linkRef].offParams, sizeof(pc.offParams));
            std::memcpy(pc.torsionParams, p[linkRef].torsionParams, sizeof(pc.torsionParams));
            pc.unbindOn = p[linkRef].unbindOn;
        }
        r = net_create(c, &b->cf, L > 0 ? &b->cx : nullptr, &pc, &b->lay, &b->net);
    } while (0);
    if (r) {
        batch_release(b);
        return r;
    }
    *out = b;
    return SBX_OK;
}

int sbx_batch_net(sbx_batch *b, sbx_net **net) {
    if (!b || !net) return SBX_ERR_ARG;
    *net = b->net;
    return SBX_OK;
}

int sbx_batch_step(sbx_batch *b, int step0, int nSteps) {
    if (!b) return SBX_ERR_ARG;
    return sbx_net_step(b->net, step0, nSteps);
}

int sbx_batch_publish(sbx_batch *b) {
    if (!b) return SBX_ERR_ARG;
    SBX_CHECK_CTX(b->ctx);
    int r = sbx_net_publish(b->net);
    if (!r) r = set_device(b->ctx);
    if (!r) r = batch_copy_state(b, false, false);
    return r;
}

int sbx_batch_layout(sbx_batch *b, int rep, int *rodBase, int *linkBase, int *cellBase) {
    if (!b) return SBX_ERR_ARG;
    SBX_CHECK_CTX(b->ctx);
    SBX_REQUIRE(b->ctx, rep >= 0 && rep < (int)b->lay.reps.size(), "batch_layout: replica %d out of range", rep);
    const RepInfo &ri = b->lay.reps[rep];
    if (rodBase) *rodBase = ri.rodBase;
    if (linkBase) *linkBase = ri.linkBase;
    if (cellBase) *cellBase = ri.cellBase;
    return SBX_OK;
}

void sbx_batch_destroy(sbx_batch *b) { batch_release(b); }

}  // extern "C"

