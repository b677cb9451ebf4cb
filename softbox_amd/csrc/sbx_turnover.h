// sbx_turnover.h -- actin turnover (AgingSystem, DepolySystem, GrowthSystem, SeveringSystem,
// FilamentBirthSystem, ActinPool): per-slot device functions and the cadence launcher shared by the
// per-System entry points and the network step.
#pragma once
#include "sbx_device.h"
#include "sbx_internal.h"

namespace sbx {

constexpr uint32_t SALT_GROW = 0x47524F57u;  // "GROW" (GrowthSystem.java:97)
constexpr uint32_t SALT_DEPO = 0x4445504Fu;  // "DEPO" (DepolySystem.java:89)
constexpr int FIL_ACTIVE = 0, FIL_FREE = -1;

// Everything one cadence reads and writes (device arrays) over C slots.
struct TurnArrays {
    int C;
    float *coord, *uVec, *yVec, *segLength, *bTransGam, *bRotGam, *bts, *brs;
    int *e1s, *e1d, *e2s, *e2d;
    int *filState, *monomerCount, *seedNode;
    float *nucFrac, *cofFrac, *bTransDiff, *bRotDiff;
    double *pool;
    long long *ledger;
    int *stats;
};

// The cadence's scratch: flags, the two scans, the free list, the split requests and the event tallies.
struct TurnScratch {
    int *deathFlag = nullptr, *severFlag = nullptr, *acceptFlag = nullptr, *freeCount = nullptr;
    int *freeOffsets = nullptr, *freeList = nullptr, *rankOffsets = nullptr, *tally = nullptr;
    float *reqCoord = nullptr, *reqUVec = nullptr, *reqYVec = nullptr;
    SortScratch scan;
    long long cap = 0;
};
int turn_scratch_reserve(sbx_ctx *c, TurnScratch &s, int C);
void turn_scratch_free(TurnScratch &s);

// One biochem cadence (FullSystemDemoHarness.java:650-684) on stream q.
int turnover_cadence(sbx_ctx *c, const TurnArrays &a, const sbx_turnover_params &p, TurnScratch &s, int step, int rng,
                     uint32_t stream, hipStream_t q);

}  // namespace sbx
