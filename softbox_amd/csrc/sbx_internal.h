// sbx_internal.h -- host-side context, error plumbing and launch helpers of libsoftbox_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/softbox_hip.h"

namespace sbx {
// Scratch of the stable counting sort (keys -> CSR offsets + ascending item list).
// counts[] is kept all-zero between sorts (`clean`), so no per-sort fill is needed.
// Segments (items of one key) longer than LONG_SEG are placed by a workgroup each (LONG_THREADS
// threads, LDS sort of up to LONG_LDS items), at most LONG_BLOCKS workgroups per sort.
constexpr int LONG_SEG = 32;
constexpr int LONG_LDS = 4096;
constexpr int LONG_THREADS = 256;
constexpr int LONG_BLOCKS = 64;
constexpr long long long_list_len(long long nItems) { return nItems / (LONG_SEG + 1) + 1; }
struct SortScratch {
    int *counts = nullptr, *tileSums = nullptr, *tilePrefix = nullptr, *tmp = nullptr, *slot = nullptr;
    int *longCount = nullptr, *longList = nullptr;  // long segments of the current sort (count, keys)
    long long keyCap = 0, itemCap = 0, tileCap = 0;
    bool clean = false;
    bool owned = true;  // false: the pointers are views of caller-provided memory (never freed here)
    int scanItems = 8;  // keys per scan thread: tiles of 256 * scanItems keys (4, 8 or 16; same result)
};
struct ChainParams;
struct TurnScratch;
}  // namespace sbx

struct sbx_ctx {
    int device;
    hipStream_t stream;
    int owns_stream;
    int sticky;            // first HIP failure code, kept until sbx_clear_error
    char err[512];
    int cu_count;
    sbx::SortScratch api_sort;  // scratch of the stand-alone sbx_grid_build / sbx_csr_build entry points
    int *form_csr;              // sbx_form_admit's link CSRs (by linkFilA, by linkFilB)
    long long form_csr_ints;
    sbx::TurnScratch *turn;     // scratch of the stand-alone sbx_turnover_cadence
};

namespace sbx {

int fail(sbx_ctx *c, int code, const char *fmt, ...);
int check_hip(sbx_ctx *c, hipError_t e, const char *what);
int check_launch(sbx_ctx *c, const char *what);
int set_device(sbx_ctx *c);

// gfx950 dispatches workgroup b to XCD b % 8, and each XCD has its own L2.  Kernels whose
// consecutive blocks touch neighbouring data renumber their blocks so that each XCD gets one
// contiguous run of logical blocks (a bijection of [0, nb)).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    int xcd = b % kXcds, k = b / kXcds;
    int per = nb / kXcds, rem = nb % kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

inline unsigned blocks_for(long long n, int threads) {
    long long b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

int sort_scratch_reserve(sbx_ctx *c, SortScratch &s, long long nKeys, long long nItems);
void sort_scratch_free(SortScratch &s);
// Stable counting sort of items by integer key (key < 0 skipped) into offsets[nKeys+1] and
// list[...]; offsets[nKeys] is the total clamped to contentsCap; positions >= contentsCap are dropped.
// gate (device int, may be null): when it reads 0 every kernel of the sort returns at once (a
// rebuild the device decides on, inside a captured step).  st: stream (null: the context's).
// pay: optional per-item data written in list order (out[pos] = in[list[pos]], keyOut[pos] = key).
struct SortPayload {
    const float4 *in = nullptr;
    float4 *out = nullptr;
    int *keyOut = nullptr;
};
int counting_sort(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap, int *offsets,
                  int *list, const int *gate = nullptr, hipStream_t st = nullptr, SortPayload pay = SortPayload());
// The same sort with the histogram already taken: counts[key] (all-zero outside the keys present)
// and slot[m] = the atomicAdd return of item m (k_rod takes them when it bins the new pose).  The
// sort leaves counts all-zero again.
int counting_sort_counted(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap,
                          int *offsets, int *list, int *counts, const int *slot, hipStream_t st = nullptr,
                          SortPayload pay = SortPayload());
// the sort on scratch that is already large enough, with counts[0, nKeys] and tileSums all zero
int counting_sort_core(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap,
                       int *offsets, int *list, const int *gate = nullptr, hipStream_t st = nullptr,
                       SortPayload pay = SortPayload());
// the allocation-free entry point's set-up on c's device (sbx_noalloc.hip), called by sbx_create
int noalloc_setup(sbx_ctx *c);
// smallest scan tile (cells): sizes the tile-sum arrays of a sort
constexpr int SCAN_MIN_TILE_KEYS = 1024;
// out[0..n] = exclusive prefix sum of in[0..n) (out[n] = the total), with s's tile scratch
int exclusive_scan(sbx_ctx *c, SortScratch &s, const int *in, int n, int *out, hipStream_t st = nullptr);
// the histogram pass alone: slot[m] = atomicAdd(counts[key[m]], 1)
int count_keys(sbx_ctx *c, const int *key, int nItems, int nKeys, int *counts, int *slot, hipStream_t st = nullptr);
int fill_i32(sbx_ctx *c, int *dst, int value, long long count);
// frees and deletes a turnover scratch (sbx_turnover.hip); null is a no-op
void turn_scratch_destroy(TurnScratch *s);

ChainParams chain_params(const float *cp);

}  // namespace sbx

#define SBX_CHECK_CTX(c)                         \
    do {                                         \
        if (!(c)) return SBX_ERR_ARG;            \
        if ((c)->sticky) return (c)->sticky;     \
    } while (0)

#define SBX_REQUIRE(c, cond, ...)                                     \
    do {                                                              \
        if (!(cond)) return sbx::fail((c), SBX_ERR_ARG, __VA_ARGS__); \
    } while (0)
