// sbx_tiles.h -- block-level building blocks of the tiled candidate kernels (the per-step broad
// phase and the skin broad phase).
//
// A tiled candidate kernel gives one block TILE_BODIES bodies and flattens their candidate tests
// into one index range that all TILE_THREADS threads sweep, so a body with a long neighbourhood
// does not hold back the wave it happens to sit in.  Passing candidates are appended to a per-body
// key buffer in LDS and each key is placed at its rank among its body's keys: a key orders
// candidates exactly as SpatialGrid.broadPhase emits them (stencil position << 26 | index), keys
// of one body are distinct, so the slice comes out in the reference's order without a sort.
#pragma once
#include <hip/hip_runtime.h>

#include "sbx_internal.h"

namespace sbx {

constexpr int TILE_BODIES = 64;    // bodies per block (one wave's prefix sum)
constexpr int TILE_THREADS = 256;  // threads per block
constexpr int TILE_KEYS = 64;      // keys buffered per body; a body with more is done block-wide
constexpr int TILE_CHUNK = 8;      // consecutive candidate tests per thread (in flight together)
constexpr int KEY_INDEX_BITS = 26;
constexpr int KEY_INDEX_MASK = (1 << KEY_INDEX_BITS) - 1;
constexpr int KEEP_MAX = 256;      // keys the chunked reduction keeps (maxCand <= SBX_MAX_CAND)
static_assert(KEEP_MAX >= TILE_THREADS && (KEEP_MAX & (KEEP_MAX - 1)) == 0 &&
                  2 * KEEP_MAX <= TILE_BODIES * TILE_KEYS,
              "bitonic halves fit the key buffer");

// Inclusive prefix sum over the 64 lanes of one wave.
__device__ __forceinline__ int wave_inclusive_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Largest b in [0, TILE_BODIES) with start[b] <= e (start non-decreasing, start[0] = 0 <= e).
__device__ __forceinline__ int tile_body_of(const int *start, int e) {
    int lo = 0;
#pragma unroll
    for (int step = TILE_BODIES / 2; step > 0; step >>= 1)
        if (start[lo + step] <= e) lo += step;
    return lo;
}

// Places keys[0, m) (distinct) at their ranks: emit(rank, key) for rank < limit.  Called by the
// whole block; each thread ranks keys tid, tid + TILE_THREADS, ...
template <typename Emit>
__device__ __forceinline__ void block_rank_place(const int *keys, int m, int limit, Emit emit) {
    for (int p = threadIdx.x; p < m; p += TILE_THREADS) {
        const int key = keys[p];
        int rank = 0;
        for (int x = 0; x < m; x++) rank += keys[x] < key ? 1 : 0;
        if (rank < limit) emit(rank, key);
    }
}

// The KEEP_MAX smallest of the keys keyAt(k), k in [0, total) (keyAt returns -1 for a rejected
// candidate), ascending in buf[0, KEEP_MAX) padded with INT_MAX: candidates are taken in chunks
// of TILE_THREADS and (kept | chunk) is bitonic-sorted after each chunk.  For a body with more
// candidates than the block buffers; called by the whole block; buf holds 2 * KEEP_MAX ints.
template <typename KeyAt>
__device__ void block_keep_smallest(int total, KeyAt keyAt, int *buf, int *chunkN) {
    const int tid = threadIdx.x;
    for (int x = tid; x < KEEP_MAX; x += TILE_THREADS) buf[x] = 0x7fffffff;
    for (int base = 0; base < total; base += TILE_THREADS) {
        __syncthreads();  // the previous chunk's sort and every read of *chunkN are done
        if (tid == 0) *chunkN = 0;
        for (int x = tid; x < KEEP_MAX; x += TILE_THREADS) buf[KEEP_MAX + x] = 0x7fffffff;
        __syncthreads();
        const int k = base + tid;
        if (k < total) {
            const int key = keyAt(k);
            if (key >= 0) buf[KEEP_MAX + atomicAdd(chunkN, 1)] = key;  // <= TILE_THREADS per chunk
        }
        __syncthreads();
        if (*chunkN == 0) continue;
        for (int size = 2; size <= 2 * KEEP_MAX; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int p = tid; p < KEEP_MAX; p += TILE_THREADS) {
                    int lo = ((p & ~(stride - 1)) << 1) | (p & (stride - 1));
                    int hi = lo + stride;
                    int x = buf[lo], y = buf[hi];
                    if ((x > y) == ((lo & size) == 0)) {
                        buf[lo] = y;
                        buf[hi] = x;
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
}

}  // namespace sbx
