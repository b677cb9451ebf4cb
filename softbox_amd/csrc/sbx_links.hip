// sbx_links.hip -- crosslink Systems (CrosslinkerSystem.java) behind the C ABI: one thread per
// link for the force laws (self-written 12-float payload rows, no atomics) and one thread per
// segment for the CSR gathers (FP64 sums in ascending link order).
#include "sbx_internal.h"
#include "sbx_links.h"

namespace sbx {

LinkParams link_params(const float *xlParams, const float *offParams, const float *torsionParams) {
    LinkParams lp = {};
    if (xlParams) {
        lp.restLength = xlParams[0];
        lp.fracMoveBase = xlParams[1];
        lp.dt = xlParams[2];
    }
    if (offParams) {
        lp.offConst = offParams[0];
        lp.offCoeff = offParams[1];
        lp.offExp = offParams[2];
        lp.offDt = offParams[3];
        lp.offRest = offParams[4];
    }
    if (torsionParams) {
        lp.torsionSpring = torsionParams[0];
        lp.torsionOn = torsionParams[1] < 0.5f ? 0 : 1;
    }
    return lp;
}

}  // namespace sbx

using namespace sbx;

namespace {

__global__ void __launch_bounds__(256) k_count_links(const int *__restrict__ linkState,
                                                     const int *__restrict__ linkFilA,
                                                     const int *__restrict__ linkFilB,
                                                     int *__restrict__ activeLinkCount, int C) {
    int L = blockIdx.x * 256 + threadIdx.x;
    if (L >= C) return;
    if (linkState[L] < 0) return;
    atomicAdd(&activeLinkCount[linkFilA[L]], 1);  // integer totals: order-independent
    atomicAdd(&activeLinkCount[linkFilB[L]], 1);
}

__global__ void __launch_bounds__(256) k_link_unbind(const float *__restrict__ uVec, const float *__restrict__ end1,
                                                     const int *__restrict__ linkFilA,
                                                     const int *__restrict__ linkFilB, const float *__restrict__ loc1,
                                                     const float *__restrict__ loc2, int *__restrict__ linkState,
                                                     float *__restrict__ strainHist, int *__restrict__ strainPlace,
                                                     LinkParams lp, int step, int seed, int rng, uint32_t stream,
                                                     int nSeg, int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    int st = linkState[k];
    if (st < 0) return;
    LinkEnd A = link_end_planar(linkFilA[k], nSeg, nullptr, uVec, end1, nullptr);
    LinkEnd B = link_end_planar(linkFilB[k], nSeg, nullptr, uVec, end1, nullptr);
    int ns = link_unbind_one(k, (uint32_t)k, A, B, loc1[k], loc2[k], strainHist, strainPlace, lp, step, seed, rng,
                             stream, st);
    if (ns != st) linkState[k] = ns;
}

__global__ void __launch_bounds__(256) k_link_forces(const float *__restrict__ coord, const float *__restrict__ uVec,
                                                     const float *__restrict__ end1,
                                                     const float *__restrict__ bTransGam,
                                                     const int *__restrict__ linkFilA,
                                                     const int *__restrict__ linkFilB, const float *__restrict__ loc1,
                                                     const float *__restrict__ loc2, const int *__restrict__ filLinkCt,
                                                     float *__restrict__ xlinkData, LinkParams lp, int nSeg,
                                                     int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    float pay[12];
#pragma unroll
    for (int j = 0; j < 12; j++) pay[j] = 0.0f;
    int a = linkFilA[k], b = linkFilB[k];
    if (a >= 0 && b >= 0) {
        LinkEnd A = link_end_planar(a, nSeg, coord, uVec, end1, bTransGam);
        LinkEnd B = link_end_planar(b, nSeg, coord, uVec, end1, bTransGam);
        link_force_one(A, B, loc1[k], loc2[k], filLinkCt[a], filLinkCt[b], lp, pay);
    }
    float *d = xlinkData + (size_t)k * XL_STRIDE;
#pragma unroll
    for (int j = 0; j < 12; j++) d[j] = pay[j];
}

__global__ void __launch_bounds__(256) k_link_torsion(const float *__restrict__ uVec, const int *__restrict__ linkFilA,
                                                      const int *__restrict__ linkFilB,
                                                      const int *__restrict__ linkState,
                                                      const int *__restrict__ linkOrientSame,
                                                      float *__restrict__ torqueMagHist,
                                                      int *__restrict__ torqueMagPlace, float *__restrict__ xlinkData,
                                                      LinkParams lp, int nSeg, int nLinks) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nLinks) return;
    if (linkState[k] < 0) return;
    int a = linkFilA[k], b = linkFilB[k];
    if (a < 0 || b < 0) return;
    float *d = xlinkData + (size_t)k * XL_STRIDE;
    float pay[12];
#pragma unroll
    for (int j = 0; j < 12; j++) pay[j] = d[j];
    LinkEnd A = link_end_planar(a, nSeg, nullptr, uVec, uVec, nullptr);
    LinkEnd B = link_end_planar(b, nSeg, nullptr, uVec, uVec, nullptr);
    link_torsion_one(k, A, B, linkOrientSame[k] != 0, torqueMagHist, torqueMagPlace, lp, pay);
    d[3] = pay[3]; d[4] = pay[4]; d[5] = pay[5];
    d[9] = pay[9]; d[10] = pay[10]; d[11] = pay[11];
}

__global__ void __launch_bounds__(256) k_seg_gather(const int *__restrict__ offsets, const int *__restrict__ idx,
                                                    const float *__restrict__ xlinkData,
                                                    const int *__restrict__ linkState, float *__restrict__ forceSum,
                                                    float *__restrict__ torqueSum, int nSeg, int off) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nSeg) return;
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    int active = 0;
    gather_side(s, offsets, idx, xlinkData, linkState, off, f, t, active);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        forceSum[a * nSeg + s] = (float)(forceSum[a * nSeg + s] + f[a]);
        torqueSum[a * nSeg + s] = (float)(torqueSum[a * nSeg + s] + t[a]);
    }
}

}  // namespace

extern "C" {

int sbx_count_active_links(sbx_ctx *c, const int *linkState, const int *linkFilA, const int *linkFilB,
                           int *activeLinkCount, int C, int nSeg) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, C >= 0 && nSeg >= 0, "count_active_links: negative size");
    SBX_REQUIRE(c, nSeg == 0 || activeLinkCount, "count_active_links: null activeLinkCount");
    int r = fill_i32(c, activeLinkCount, 0, nSeg);
    if (r || C == 0) return r;
    SBX_REQUIRE(c, linkState && linkFilA && linkFilB, "count_active_links: null array");
    hipLaunchKernelGGL(k_count_links, dim3(blocks_for(C, 256)), dim3(256), 0, c->stream, linkState, linkFilA,
                       linkFilB, activeLinkCount, C);
    return check_launch(c, "k_count_links");
}

int sbx_link_unbind(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1,
                    const int *linkFilA, const int *linkFilB, const float *loc1, const float *loc2, int *linkState,
                    float *strainHist, int *strainPlace, const float *offParams, const int *counts, int nSeg,
                    int rng, uint32_t stream) {
    (void)filCoord;
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, offParams && counts && nSeg >= 0, "link_unbind: null params/counts");
    SBX_REQUIRE(c, rng == SBX_RNG_WANG || rng == SBX_RNG_PHILOX, "link_unbind: unknown rng %d", rng);
    int nLinks = counts[0];
    SBX_REQUIRE(c, nLinks >= 0, "link_unbind: nLinks < 0");
    if (nLinks == 0) return SBX_OK;
    SBX_REQUIRE(c, filUVec && filEnd1 && linkFilA && linkFilB && loc1 && loc2 && linkState && strainHist &&
                       strainPlace,
                "link_unbind: null array");
    SBX_REQUIRE(c, offParams[4] != 0.0f, "link_unbind: restLength must be non-zero");
    LinkParams lp = link_params(nullptr, offParams, nullptr);
    hipLaunchKernelGGL(k_link_unbind, dim3(blocks_for(nLinks, 256)), dim3(256), 0, c->stream, filUVec, filEnd1,
                       linkFilA, linkFilB, loc1, loc2, linkState, strainHist, strainPlace, lp, counts[1], counts[2],
                       rng, stream, nSeg, nLinks);
    return check_launch(c, "k_link_unbind");
}

int sbx_link_forces(sbx_ctx *c, const float *filCoord, const float *filUVec, const float *filEnd1,
                    const float *filBTransGam, const int *linkFilA, const int *linkFilB, const float *loc1,
                    const float *loc2, const int *filLinkCt, float *xlinkData, const float *xlParams, int nSeg,
                    int nLinks) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, xlParams && nSeg >= 0 && nLinks >= 0, "link_forces: bad sizes or null xlParams");
    if (nLinks == 0) return SBX_OK;
    SBX_REQUIRE(c, filCoord && filUVec && filEnd1 && filBTransGam && linkFilA && linkFilB && loc1 && loc2 &&
                       filLinkCt && xlinkData,
                "link_forces: null array");
    SBX_REQUIRE(c, xlParams[2] != 0.0f, "link_forces: dt must be non-zero");
    LinkParams lp = link_params(xlParams, nullptr, nullptr);
    hipLaunchKernelGGL(k_link_forces, dim3(blocks_for(nLinks, 256)), dim3(256), 0, c->stream, filCoord, filUVec,
                       filEnd1, filBTransGam, linkFilA, linkFilB, loc1, loc2, filLinkCt, xlinkData, lp, nSeg, nLinks);
    return check_launch(c, "k_link_forces");
}

int sbx_link_torsion(sbx_ctx *c, const float *filUVec, const int *linkFilA, const int *linkFilB,
                     const int *linkState, const int *linkOrientSame, float *torqueMagHist, int *torqueMagPlace,
                     float *xlinkData, const float *torsionParams, int nSeg, int nLinks) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, torsionParams && nSeg >= 0 && nLinks >= 0, "link_torsion: bad sizes or null params");
    if (nLinks == 0 || torsionParams[1] < 0.5f) return SBX_OK;
    SBX_REQUIRE(c, filUVec && linkFilA && linkFilB && linkState && linkOrientSame && torqueMagHist &&
                       torqueMagPlace && xlinkData,
                "link_torsion: null array");
    LinkParams lp = link_params(nullptr, nullptr, torsionParams);
    hipLaunchKernelGGL(k_link_torsion, dim3(blocks_for(nLinks, 256)), dim3(256), 0, c->stream, filUVec, linkFilA,
                       linkFilB, linkState, linkOrientSame, torqueMagHist, torqueMagPlace, xlinkData, lp, nSeg,
                       nLinks);
    return check_launch(c, "k_link_torsion");
}

int sbx_seg_gather(sbx_ctx *c, const int *segOffsets, const int *segLinkIdx, const float *xlinkData,
                   const int *linkState, float *filForceSum, float *filTorqueSum, int nSeg, int side) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nSeg >= 0 && (side == 0 || side == 1), "seg_gather: bad nSeg or side");
    if (nSeg == 0) return SBX_OK;
    SBX_REQUIRE(c, segOffsets && segLinkIdx && xlinkData && linkState && filForceSum && filTorqueSum,
                "seg_gather: null array");
    hipLaunchKernelGGL(k_seg_gather, dim3(blocks_for(nSeg, 256)), dim3(256), 0, c->stream, segOffsets, segLinkIdx,
                       xlinkData, linkState, filForceSum, filTorqueSum, nSeg, side == 0 ? 0 : 6);
    return check_launch(c, "k_seg_gather");
}

}  // extern "C"
