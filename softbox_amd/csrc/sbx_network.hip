// sbx_network.hip -- the whole filament-network mechanics step as a short fixed kernel sequence.
//
// Replaces the filament block of DenseContractileHarness.buildPlan/cpuStep
// (DenseContractileHarness.java:350-429, 460-606), which the reference runs as ~30 TaskGraph tasks:
//   [grid]  k_body_cell -> counting sort (fill, count, 3-pass scan, scatter, place) -> k_broad_phase
//   [links] k_net_unbind -> k_net_count -> k_net_link (force + torsion payload)
//   [rods]  k_rod: zero + Brownian + chain + gather A + gather B + confine + integrate + derive
// k_rod keeps forceSum/torqueSum and the Brownian draw in registers and ping-pongs the pose buffers
// (the chain reads neighbours' pre-step pose while each rod writes its own new pose).  The step
// counter lives in device memory so a run of steps can be captured once as a hipGraph and replayed.
// The crosslink CSR (keyed by linkFilA / linkFilB) depends only on the link topology, so it is built
// at creation and on sbx_net_topology_changed instead of every step; linkState (which unbinding
// changes every step) is applied by the gather guard, as in CrosslinkerSystem.segGatherA/B.
#include <new>
#include <string>
#include <vector>

#include "sbx_device.h"
#include "sbx_grid.h"
#include "sbx_internal.h"
#include "sbx_links.h"

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
    // links
    int *offA = nullptr, *idxA = nullptr, *offB = nullptr, *idxB = nullptr, *ownCount = nullptr;
    LinkParams lp{};
    ChainParams cp{};
    // step counter
    int *dStep = nullptr;
    // profiling
    int profiling = 0;
    std::vector<Timer> timers;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> eventPool;
    // graph
    int useGraph = 0;
    int graphSteps = 0;  // non-zero while a capture is being recorded
    int graphCur = 0;    // pose buffer index the captured graph starts from
    hipGraphExec_t graphExec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

enum { T_BIN = 0, T_SORT, T_PAIRS, T_UNBIND, T_COUNT, T_LINK, T_ROD, T_COUNT_TIMERS };
const char *kTimerNames[T_COUNT_TIMERS] = {"k_body_cell", "grid_sort", "k_broad_phase", "k_net_unbind",
                                           "k_net_count",  "k_net_link", "k_rod"};

struct RodArgs {
    const float *coord, *uVec, *yVec;
    float *coordOut, *uOut, *yOut, *zVec, *end1, *end2;
    const float *segLength, *bTransGam, *bRotGam, *bts, *brs;
    float *forceSum, *torqueSum;
    const int *e1Slot, *e1Side, *e2Slot, *e2Side;
    const int *offA, *idxA, *offB, *idxB, *linkState;
    const float *xlinkData;
    const int *dStep;
    int stepOff;
    int n, chainOn, hasLinks, confineOn, checkInt, rng, seed;
    uint32_t stream;
    float dt, bfm;
    double tau, halfX, halfY, halfZ, coeff;
};

// Per-rod fused step.  The force accumulation keeps the reference's exact float rounding sequence:
// forceSum = 0 -> (float)(f + chain) -> (float)(f + sumA) -> (float)(f + sumB) -> confine -> integrate.
__global__ void __launch_bounds__(256) k_rod(RodArgs a, ChainParams cp) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int n = a.n;
    if (i >= n) return;
    int step = *a.dStep + a.stepOff;
    float f[3] = {0.0f, 0.0f, 0.0f}, t[3] = {0.0f, 0.0f, 0.0f};
    if (a.chainOn) {
        double fd[3] = {0.0, 0.0, 0.0}, td[3] = {0.0, 0.0, 0.0};
        chain_body(n, i, a.coord, a.uVec, a.segLength, a.e2Slot[i], a.e2Side[i], a.e1Slot[i], a.e1Side[i],
                   a.bTransGam, a.bRotGam, cp, fd, td);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            f[k] = (float)(f[k] + fd[k]);
            t[k] = (float)(t[k] + td[k]);
        }
    }
    if (a.hasLinks) {
        double fa[3] = {0, 0, 0}, ta[3] = {0, 0, 0};
        gather_side(i, a.offA, a.idxA, a.xlinkData, a.linkState, 0, fa, ta);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            f[k] = (float)(f[k] + fa[k]);
            t[k] = (float)(t[k] + ta[k]);
        }
        double fb[3] = {0, 0, 0}, tb[3] = {0, 0, 0};
        gather_side(i, a.offB, a.idxB, a.xlinkData, a.linkState, 6, fb, tb);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            f[k] = (float)(f[k] + fb[k]);
            t[k] = (float)(t[k] + tb[k]);
        }
    }
    Pose p;
    float gT[3], gR[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p.c[k] = a.coord[k * n + i];
        p.u[k] = a.uVec[k * n + i];
        p.y[k] = a.yVec[k * n + i];
        gT[k] = a.bTransGam[k * n + i];
        gR[k] = a.bRotGam[k * n + i];
    }
    float segLen = a.segLength[i];
    if (a.confineOn && !(a.checkInt > 1 && (step % a.checkInt) != 0))
        confine_body(p.c, p.u, segLen, gT[0], a.tau, a.halfX, a.halfY, a.halfZ, a.coeff, f, t);
    if (a.forceSum) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            a.forceSum[k * n + i] = f[k];
            a.torqueSum[k * n + i] = t[k];
        }
    }
    float gf[3], gt[3], rf[3], rt[3];
    brownian_noise(a.rng, (uint32_t)i, step, a.seed, a.stream, gf, gt);
    float tS = a.bts[i], rS = a.brs[i];
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
}

__global__ void __launch_bounds__(256) k_net_unbind(const float *__restrict__ uVec, const float *__restrict__ end1,
                                                    const int *__restrict__ linkFilA,
                                                    const int *__restrict__ linkFilB, const float *__restrict__ loc1,
                                                    const float *__restrict__ loc2, int *__restrict__ linkState,
                                                    float *__restrict__ strainHist, int *__restrict__ strainPlace,
                                                    LinkParams lp, const int *__restrict__ dStep, int stepOff,
                                                    int seed, int rng, uint32_t stream, int nSeg, int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    int st = linkState[k];
    if (st < 0) return;
    int step = *dStep + stepOff;
    int ns = link_unbind_one(k, nSeg, linkFilA[k], linkFilB[k], loc1[k], loc2[k], uVec, end1, strainHist,
                             strainPlace, lp, step, seed, rng, stream, st);
    if (ns != st) linkState[k] = ns;
}

// activeLinkCount[s] = ACTIVE links keyed to s on either side (CrosslinkerSystem.countActiveLinks,
// CrosslinkerSystem.java:694-706), read from the two CSR ranges: no atomics.
__global__ void __launch_bounds__(256) k_net_count(const int *__restrict__ offA, const int *__restrict__ idxA,
                                                   const int *__restrict__ offB, const int *__restrict__ idxB,
                                                   const int *__restrict__ linkState,
                                                   int *__restrict__ activeLinkCount, int nSeg) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nSeg) return;
    int cnt = 0;
    for (int k = offA[s]; k < offA[s + 1]; k++) cnt += linkState[idxA[k]] >= 0 ? 1 : 0;
    for (int k = offB[s]; k < offB[s + 1]; k++) cnt += linkState[idxB[k]] >= 0 ? 1 : 0;
    activeLinkCount[s] = cnt;
}

// linkForces + linkTorsion for one link (CrosslinkerSystem.java:62-123, 96-128)
__global__ void __launch_bounds__(256) k_net_link(const float *__restrict__ coord, const float *__restrict__ uVec,
                                                  const float *__restrict__ end1, const float *__restrict__ bTransGam,
                                                  const int *__restrict__ linkFilA, const int *__restrict__ linkFilB,
                                                  const float *__restrict__ loc1, const float *__restrict__ loc2,
                                                  const int *__restrict__ activeLinkCount,
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
        link_force_one(nSeg, a, b, loc1[k], loc2[k], coord, uVec, end1, bTransGam, activeLinkCount[a],
                       activeLinkCount[b], lp, pay);
        if (lp.torsionOn && linkState[k] >= 0)
            link_torsion_one(k, nSeg, a, b, linkOrientSame[k] != 0, uVec, torqueMagHist, torqueMagPlace, lp, pay);
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

hipEvent_t take_event(sbx_net *net) {
    if (!net->eventPool.empty()) {
        hipEvent_t e = net->eventPool.back();
        net->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
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
            hipEventRecord(a, net->ctx->stream);
        }
    }
    ~Scope() {
        if (a) {
            hipEventRecord(b, net->ctx->stream);
            net->pending.push_back({id, {a, b}});
        }
    }
};

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

// one step's launches; stepOff is added to the device step counter
int enqueue_step(sbx_net *net, int stepOff) {
    sbx_ctx *c = net->ctx;
    hipStream_t s = c->stream;
    const int n = net->n, L = net->L;
    const float *coord = net->coordB[net->cur], *uVec = net->uB[net->cur], *yVec = net->yB[net->cur];
    int r;
    if (net->p.gridOn) {
        {
            Scope sc(net, T_BIN);
            r = launch_body_cell(c, coord, n, n, net->geom, net->bodyCell);
            if (r) return r;
        }
        {
            Scope sc(net, T_SORT);
            r = counting_sort(c, net->sort, net->bodyCell, n, net->p.gridDims[3], n, net->cellOffsets,
                              net->cellContents);
            if (r) return r;
        }
        {
            Scope sc(net, T_PAIRS);
            r = launch_broad_phase(c, coord, net->boundR, net->bodyCell, net->cellOffsets, net->cellContents,
                                   net->geom, n, n, net->candPartner, net->candCount, net->p.maxCand);
            if (r) return r;
        }
    }
    int *activeCount = net->xl.activeLinkCount ? net->xl.activeLinkCount : net->ownCount;
    if (L > 0) {
        if (net->p.unbindOn) {
            Scope sc(net, T_UNBIND);
            hipLaunchKernelGGL(k_net_unbind, dim3(blocks_for(L, 256)), dim3(256), 0, s, uVec, net->fil.end1,
                               net->xl.linkFilA, net->xl.linkFilB, net->xl.loc1, net->xl.loc2, net->xl.linkState,
                               net->xl.strainHist, net->xl.strainPlace, net->lp, net->dStep, stepOff, net->p.seed,
                               net->p.rng, net->p.stream, n, L);
        }
        {
            Scope sc(net, T_COUNT);
            hipLaunchKernelGGL(k_net_count, dim3(blocks_for(n, 256)), dim3(256), 0, s, net->offA, net->idxA,
                               net->offB, net->idxB, net->xl.linkState, activeCount, n);
        }
        {
            Scope sc(net, T_LINK);
            hipLaunchKernelGGL(k_net_link, dim3(blocks_for(L, 256)), dim3(256), 0, s, coord, uVec, net->fil.end1,
                               net->fil.bTransGam, net->xl.linkFilA, net->xl.linkFilB, net->xl.loc1, net->xl.loc2,
                               activeCount, net->xl.linkState, net->xl.linkOrientSame, net->xl.torqueMagHist,
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
        hipLaunchKernelGGL(k_rod, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, net->cp);
    }
    net->cur = nxt;
    return check_launch(c, "k_rod");
}

int build_link_csr(sbx_net *net) {
    if (net->L == 0) return SBX_OK;
    int r = counting_sort(net->ctx, net->sort, net->xl.linkFilA, net->L, net->n, net->L, net->offA, net->idxA);
    if (r) return r;
    return counting_sort(net->ctx, net->sort, net->xl.linkFilB, net->L, net->n, net->L, net->offB, net->idxB);
}

void release(sbx_net *net) {
    if (!net) return;
    if (net->ctx) hipSetDevice(net->ctx->device);
    if (net->ctx) hipStreamSynchronize(net->ctx->stream);
    if (net->graphExec) hipGraphExecDestroy(net->graphExec);
    if (net->graph) hipGraphDestroy(net->graph);
    void *bufs[] = {net->ownPose, net->boundR, net->bodyCell, net->cellOffsets, net->cellContents, net->candPartner,
                    net->candCount, net->offA, net->idxA, net->offB, net->idxB, net->ownCount, net->dStep};
    for (void *b : bufs)
        if (b) hipFree(b);
    sort_scratch_free(net->sort);
    for (auto &pe : net->pending) {
        hipEventDestroy(pe.second.first);
        hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : net->eventPool) hipEventDestroy(e);
    delete net;
}

template <typename P>
int alloc(sbx_ctx *c, P **p, size_t count) {
    return check_hip(c, hipMalloc((void **)p, sizeof(P) * (count ? count : 1)), "sbx_net alloc");
}

}  // namespace

extern "C" {

int sbx_net_create(sbx_ctx *c, const sbx_filaments *fil, const sbx_crosslinks *xl, const sbx_net_params *p,
                   sbx_net **out) {
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
    if (p->gridOn) {
        int r = check_grid(c, p->gridParams, p->gridDims, p->periodic);
        if (r) return r;
        SBX_REQUIRE(c, p->maxCand > 0, "net_create: maxCand must be > 0");
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
            net->geom = grid_geom(p->gridParams, p->gridDims, p->periodic, p->periodic ? p->boxLen : nullptr);
            if ((r = alloc(c, &net->boundR, n))) break;
            if ((r = alloc(c, &net->bodyCell, n))) break;
            if ((r = alloc(c, &net->cellOffsets, (size_t)p->gridDims[3] + 1))) break;
            if ((r = alloc(c, &net->cellContents, n))) break;
            if ((r = alloc(c, &net->candPartner, (size_t)n * p->maxCand))) break;
            if ((r = alloc(c, &net->candCount, n))) break;
            hipLaunchKernelGGL(k_bounding_radius, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, fil->segLength,
                               p->actinRadius, net->boundR, n);
            if ((r = check_launch(c, "k_bounding_radius"))) break;
        }
        if (L > 0) {
            if ((r = alloc(c, &net->offA, (size_t)n + 1))) break;
            if ((r = alloc(c, &net->offB, (size_t)n + 1))) break;
            if ((r = alloc(c, &net->idxA, L))) break;
            if ((r = alloc(c, &net->idxB, L))) break;
            if (!xl->activeLinkCount && (r = alloc(c, &net->ownCount, n))) break;
            if ((r = build_link_csr(net))) break;
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

void sbx_net_destroy(sbx_net *net) { release(net); }

int sbx_net_step(sbx_net *net, int step0, int nSteps) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSteps >= 0, "net_step: nSteps < 0");
    if (nSteps == 0) return SBX_OK;
    int r = set_device(c);
    if (r) return r;
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, c->stream, net->dStep, step0);
    if ((r = check_launch(c, "k_set_step"))) return r;
    // Eager launches use stepOff = done - advanced, where `advanced` counts the steps already added
    // to the device counter by graph replays (each replay ends with k_add_step(G)).
    int done = 0, advanced = 0;
    if (net->useGraph && !net->profiling) {
        // one captured graph = 2 steps (the pose ping-pong returns to its start buffers) + step advance
        const int G = 2;
        if (!net->graphExec) {
            int savedCur = net->cur;
            if ((r = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
                return r;
            net->graphSteps = G;
            int er = SBX_OK;
            for (int k = 0; k < G && !er; k++) er = enqueue_step(net, k);
            if (!er) hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, G);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(c->stream, &g);
            net->graphSteps = 0;
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
            if ((r = enqueue_step(net, done - advanced))) return r;
            done++;
        }
        while (nSteps - done >= G) {
            if (done != advanced) {  // a replay reads its first step from the device counter: bring it up to date
                hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, c->stream, net->dStep, done - advanced);
                if ((r = check_launch(c, "k_add_step"))) return r;
                advanced = done;
            }
            if ((r = check_hip(c, hipGraphLaunch(net->graphExec, c->stream), "graph launch"))) return r;
            done += G;
            advanced += G;  // G is even: net->cur is unchanged by a replay
        }
    }
    for (; done < nSteps; done++) {
        if ((r = enqueue_step(net, done - advanced))) return r;
    }
    if (net->profiling) return drain_timers(net);
    return SBX_OK;
}

int sbx_net_publish(sbx_net *net) {
    if (!net) return SBX_ERR_ARG;
    sbx_ctx *c = net->ctx;
    SBX_CHECK_CTX(c);
    if (net->cur == 0) return SBX_OK;
    size_t bytes = sizeof(float) * 3 * (size_t)net->n;
    int r = check_hip(c, hipMemcpyAsync(net->coordB[0], net->coordB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                      "publish coord");
    if (!r) r = check_hip(c, hipMemcpyAsync(net->uB[0], net->uB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                          "publish uVec");
    if (!r) r = check_hip(c, hipMemcpyAsync(net->yB[0], net->yB[1], bytes, hipMemcpyDeviceToDevice, c->stream),
                          "publish yVec");
    if (!r) net->cur = 0;
    return r;
}

int sbx_net_topology_changed(sbx_net *net) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    return build_link_csr(net);
}

int sbx_net_candidates(sbx_net *net, int **candPartner, int **candCount, int **gridCellOffsets,
                       int **gridCellContents, int **bodyCell) {
    if (!net) return SBX_ERR_ARG;
    SBX_CHECK_CTX(net->ctx);
    if (candPartner) *candPartner = net->candPartner;
    if (candCount) *candCount = net->candCount;
    if (gridCellOffsets) *gridCellOffsets = net->cellOffsets;
    if (gridCellContents) *gridCellContents = net->cellContents;
    if (bodyCell) *bodyCell = net->bodyCell;
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
    if (on) SBX_REQUIRE(net->ctx, net->ctx->stream != nullptr, "net_use_graph: needs a non-default stream");
    net->useGraph = on ? 1 : 0;
    return SBX_OK;
}

}  // extern "C"
