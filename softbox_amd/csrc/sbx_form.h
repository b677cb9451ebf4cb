// sbx_form.h -- crosslinker formation (CrosslinkerSystem.java:298-760, FormationGrid.java, FilIDSystem.java):
// host launchers shared by the per-System entry points and the fused network step.
#pragma once
#include "sbx_grid.h"
#include "sbx_internal.h"

namespace sbx {

constexpr int FORM_MAXC = 256;  // CrosslinkerSystem.FORM_MAXC (CrosslinkerSystem.java:457)

// CrosslinkerStore.setFormParams (CrosslinkerStore.java:196-206): maxAngle, grab^2, minSep,
// maxLinksOnSeg, pForm, minFilLinkSep (the location jitter), grab; mode 0 both / 1 parallel / -1 anti
struct FormParams {
    float maxAngle, grabSq, minSep, maxLinks, pForm, jit, grab;
    int mode;
};
FormParams form_params(const float *formParams, int mode);

// Everything one formation pass reads and writes (device arrays).  nSeg segments, C link slots,
// reqCap candidate slots.  The link CSR (offA/idxA by linkFilA, offB/idxB by linkFilB) is the
// start-of-pass one; activeLinkCount holds the start-of-pass active counts.
struct FormPass {
    int nSeg, C, reqCap;
    const float *coord, *uVec, *end1, *end2, *segLength;
    const int *filState, *end2NbrSlot;  // filState may be null (every slot ACTIVE)
    int *linkFilA, *linkFilB, *linkState, *strainPlace, *linkOrientSame, *torqueMagPlace;
    const int *activeLinkCount;
    float *loc1, *loc2, *strainHist, *torqueMagHist;
    const int *offA, *idxA, *offB, *idxB;
    GridGeom geom;  // the formation grid
    FormParams fp;
    int rng, seed;
    uint32_t stream;
    // scratch
    int *filID, *filIDScratch, *bodyCell, *gridOff, *gridCont, *candCount, *candBase;
    int *reqFilA, *reqFilB, *reqOrient, *gatePass, *minCand, *acceptFlag, *rankOffsets;
    int *freeCount, *freeOffsets, *freeList;
    float *reqLoc1, *reqLoc2;
    int *stats;  // device [4]: [0] candidates of the pass, [1] links formed
};

int form_filid(sbx_ctx *c, const int *filState, const int *end2NbrSlot, int *filID, int *scratch, int n,
               hipStream_t q);
int form_candidates(sbx_ctx *c, SortScratch &s, const float *coord, const float *segLength, const int *filID,
                    const int *gridOff, const int *gridCont, const GridGeom &g, double grab,
                    int *candCount, int *candBase, int *reqFilA, int *reqFilB, int *total, int nSeg, int reqCap,
                    hipStream_t q);
int form_gates(sbx_ctx *c, const float *coord, const float *uVec, const float *end1, const float *end2,
               const float *segLength, const int *reqFilA, const int *reqFilB, float *reqLoc1, float *reqLoc2,
               int *reqOrient, int *gatePass, const FormParams &fp, int step, int seed, int rng, uint32_t stream,
               const GridGeom &g, int nSeg, int reqCap, hipStream_t q);
int form_admit_reduce(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const int *gatePass, int *minCand,
                      int reqCap, int nSeg, hipStream_t q);
int form_admit(sbx_ctx *c, const int *reqFilA, const int *reqFilB, const float *reqLoc1, const float *reqLoc2,
               const int *gatePass, const int *minCand, const int *activeLinkCount, const int *linkState,
               const int *offA, const int *idxA, const int *offB, const int *idxB, const float *loc1,
               const float *loc2, int *acceptFlag, const FormParams &fp, int reqCap, hipStream_t q);
int form_allocate(sbx_ctx *c, SortScratch &s, const int *reqFilA, const int *reqFilB, const float *reqLoc1,
                  const float *reqLoc2, const int *reqOrient, const int *acceptFlag, int *freeCount, int *freeOffsets,
                  int *freeList, int *rankOffsets, int *linkFilA, int *linkFilB, float *loc1, float *loc2,
                  int *linkState, float *strainHist, int *strainPlace, int *linkOrientSame, float *torqueMagHist,
                  int *torqueMagPlace, int *formed, int K, int C, hipStream_t q);
// the whole pass (V2OneXHarness.formationCpu order): filID, formation grid, candidates, gates,
// admission, allocation.  The caller rebuilds its link CSR and recounts its active links afterwards.
int form_pass(sbx_ctx *c, SortScratch &s, const FormPass &f, int step, bool filIDValid, hipStream_t q);

}  // namespace sbx
