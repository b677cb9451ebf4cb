// sbx_core.hip -- context, device memory, error plumbing, scan and stable counting sort.
//
// The counting sort is the one primitive behind SpatialGrid's cell CSR (SpatialGrid.java:99-331)
// and CrossBridgeSystem's CSR inverse (CrossBridgeSystem.java:190-298).  The reference builds it
// with per-chunk private histogram rows so that the within-key order is ascending item index.  On
// gfx950 the same unique result is produced by: atomic key histogram (which also hands each item a
// provisional slot) -> one-pass tiled scan -> scatter to the provisional slots -> stable placement
// (ascending item index within each key's segment): a segment of up to LONG_SEG items is placed by
// per-item rank (the number of same-key items with a smaller index, read back from the unordered
// segment), a longer one by one workgroup sorting it in LDS, so no segment costs more than
// O(L log^2 L) work.  The totals and the final order do not depend on atomic ordering, so the CSR is
// bit-identical to the serial reference build.
#include "sbx_internal.h"

#include <new>

namespace sbx {

int fail(sbx_ctx *c, int code, const char *fmt, ...) {
    if (!c) return code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
    if (code != SBX_ERR_ARG && !c->sticky) c->sticky = code;
    return code;
}

int check_hip(sbx_ctx *c, hipError_t e, const char *what) {
    if (e == hipSuccess) return SBX_OK;
    int code = (e == hipErrorOutOfMemory) ? SBX_ERR_NOMEM : SBX_ERR_HIP;
    return fail(c, code, "%s: %s", what, hipGetErrorString(e));
}

int check_launch(sbx_ctx *c, const char *what) { return check_hip(c, hipGetLastError(), what); }

int set_device(sbx_ctx *c) { return check_hip(c, hipSetDevice(c->device), "hipSetDevice"); }

// ------------------------------------------------------------------ fill
__global__ void __launch_bounds__(256) k_fill_i32(int *__restrict__ dst, int value, long long count) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = value;
}

int fill_i32(sbx_ctx *c, int *dst, int value, long long count) {
    if (count <= 0) return SBX_OK;
    long long b = (count + 255) / 256;
    unsigned blocks = (unsigned)(b > 65535 * 4 ? 65535 * 4 : b);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, c->stream, dst, value, count);
    return check_launch(c, "k_fill_i32");
}

// ------------------------------------------------------------------ scan
// Tile = 256 threads x 8 items.  Wave64 inclusive scan by __shfl_up, then wave totals in LDS.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_MIN_TILE = SCAN_MIN_TILE_KEYS;  // smallest tile (scanItems 4): sizes the tile sums
static_assert(SCAN_MIN_TILE == SCAN_THREADS * 4, "scan tile sizes");

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// block-wide exclusive scan of one value per thread; returns exclusive prefix, *total = block sum
__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
    __shared__ int waveSums[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    int waveOff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        int s = waveSums[w];
        if (w < wave) waveOff += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return waveOff + inc - v;
}

// Exclusive scan of tileSums[nTiles] into tilePrefix[nTiles + 1] (tilePrefix[nTiles] = the grand
// total): one block, each thread a contiguous run of tiles.
__global__ void __launch_bounds__(1024) k_tile_prefix(const int *__restrict__ tileSums, int nTiles,
                                                      int *__restrict__ tilePrefix, const int *__restrict__ gate,
                                                      int *__restrict__ longCount) {
    if (gate && *gate == 0) return;
    if (longCount && threadIdx.x == 0) *longCount = 0;  // the sort's long-segment list starts empty
    __shared__ int waveSums[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int per = (nTiles + 1023) / 1024;
    const int b = threadIdx.x * per;
    int sum = 0;
#pragma unroll 8
    for (int k = 0; k < per; k++) sum += b + k < nTiles ? tileSums[b + k] : 0;
    const int inc = wave_inclusive_scan(sum);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    int run = inc - sum;
    for (int w = 0; w < wave; w++) run += waveSums[w];
#pragma unroll 8
    for (int k = 0; k < per; k++) {
        const int v = b + k < nTiles ? tileSums[b + k] : 0;
        if (b + k < nTiles) tilePrefix[b + k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) {
        int grand = 0;
        for (int w = 0; w < 16; w++) grand += waveSums[w];
        tilePrefix[nTiles] = grand;
    }
}

// Exclusive scan of counts[n] into out[n+1] given tilePrefix[t] (the sum of the tiles before tile t,
// tilePrefix[nTiles] = the grand total), so no tile waits on another.  out[n] = grand total, clamped
// to clampTotal when >= 0.
template <int SCAN_ITEMS>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int *__restrict__ in, int n,
                                                             const int *__restrict__ tilePrefix, int nTiles,
                                                             int *__restrict__ out, int clampTotal,
                                                             const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
    const int tile = blockIdx.x;
    // this thread's SCAN_ITEMS consecutive items (16-byte loads when the tile is full)
    long long base = (long long)tile * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    if ((long long)(tile + 1) * SCAN_TILE <= n) {
        const int4 *p = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; q++) {
            const int4 a = p[q];
            v[4 * q] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) v[k] = (base + k < n) ? in[base + k] : 0;
    }
    const int prefix = tilePrefix[tile];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) sum += v[k];
    int total;
    int pre = block_exclusive_scan(sum, &total) + prefix;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        long long idx = base + k;
        if (idx < n) out[idx] = pre;
        pre += v[k];
    }
    if (tile == nTiles - 1 && threadIdx.x == 0) {
        const int grand = tilePrefix[nTiles];
        out[n] = (clampTotal >= 0 && grand > clampTotal) ? clampTotal : grand;
    }
}

int sort_scratch_reserve(sbx_ctx *c, SortScratch &s, long long nKeys, long long nItems) {
    long long nt = (nKeys + SCAN_MIN_TILE - 1) / SCAN_MIN_TILE + 1;
    if (nKeys > s.keyCap || nt > s.tileCap) {
        if (s.counts) hipFree(s.counts);
        if (s.tileSums) hipFree(s.tileSums);
        if (s.tilePrefix) hipFree(s.tilePrefix);
        s.counts = s.tileSums = s.tilePrefix = nullptr;
        s.keyCap = s.tileCap = 0;
        int r = check_hip(c, hipMalloc(&s.counts, sizeof(int) * (size_t)(nKeys + 1)), "hipMalloc counts");
        if (r) return r;
        r = check_hip(c, hipMalloc(&s.tileSums, sizeof(int) * (size_t)(nt + 1)), "hipMalloc tileSums");
        if (r) return r;
        r = check_hip(c, hipMalloc(&s.tilePrefix, sizeof(int) * (size_t)(nt + 1)), "hipMalloc tilePrefix");
        if (r) return r;
        s.keyCap = nKeys;
        s.tileCap = nt;
        s.clean = false;  // zeroed by the next sort
    }
    if (nItems > s.itemCap || !s.longCount) {
        if (s.tmp) hipFree(s.tmp);
        if (s.slot) hipFree(s.slot);
        if (s.longCount) hipFree(s.longCount);
        s.tmp = s.slot = s.longList = s.longCount = nullptr;
        s.itemCap = 0;
        int r = check_hip(c, hipMalloc(&s.tmp, sizeof(int) * (size_t)(nItems + 1)), "hipMalloc tmp");
        if (r) return r;
        r = check_hip(c, hipMalloc(&s.slot, sizeof(int) * (size_t)(nItems + 1)), "hipMalloc slot");
        if (r) return r;
        // [0] = count of long segments, then their keys (at most nItems / (LONG_SEG + 1))
        r = check_hip(c, hipMalloc(&s.longList, sizeof(int) * (size_t)(long_list_len(nItems) + 1)), "hipMalloc long");
        if (r) return r;
        s.longCount = s.longList;
        s.longList = s.longCount + 1;
        s.itemCap = nItems;
    }
    if (!s.clean) {
        int r = fill_i32(c, s.counts, 0, s.keyCap + 1);
        if (!r) r = fill_i32(c, s.tileSums, 0, s.tileCap + 1);
        if (r) return r;
        s.clean = true;
    }
    return SBX_OK;
}

void sort_scratch_free(SortScratch &s) {
    if (!s.owned) {
        s = SortScratch();
        return;
    }
    if (s.counts) hipFree(s.counts);
    if (s.tileSums) hipFree(s.tileSums);
    if (s.tilePrefix) hipFree(s.tilePrefix);
    if (s.tmp) hipFree(s.tmp);
    if (s.slot) hipFree(s.slot);
    if (s.longCount) hipFree(s.longCount);
    s = SortScratch();
}

// ------------------------------------------------------------------ counting sort
// Scratch invariant: counts[] is all zero between sorts (zeroed once at allocation; each sort
// re-zeroes exactly the entries it dirtied), so a sort issues no fill over the key range -- which
// for a sparse grid is ~10x larger than the item count.
//
// 1. k_key_count:   slot[m] = atomicAdd(counts[key]) (a provisional, unordered position);
// 2. k_tile_reduce: tileSums = per-tile sums of counts; k_tile_prefix: their exclusive scan (one block);
// 3. k_scan_tiles:  offsets = exclusive scan of counts (each tile from its prefix);
// 4. k_key_scatter: tmp[offsets[key] + slot[m]] = m; counts[key] = 0;
// 5. k_key_place:   rank of m among its key's segment (items with a smaller index) -> list.
// The result is independent of the atomic order: ascending item index within each key.
__global__ void __launch_bounds__(256) k_key_count(const int *__restrict__ key, int nItems, int nKeys,
                                                   int *__restrict__ counts, int *__restrict__ slot,
                                                   const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k < 0 || k >= nKeys) return;
    slot[m] = atomicAdd(&counts[k], 1);
}

// tileSums[t] = sum of tile t of counts (one block per tile; no atomics: device-scope atomics on a
// few hot addresses serialise at the memory side)
template <int SCAN_ITEMS>
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_reduce(const int *__restrict__ in, int n,
                                                              int *__restrict__ tileSums,
                                                              const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
    __shared__ int red[SCAN_THREADS / 64];
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int sum = 0;
    if ((long long)(blockIdx.x + 1) * SCAN_TILE <= n) {
        const int4 *p = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS / 4; q++) {
            const int4 a = p[q];
            sum += a.x + a.y + a.z + a.w;
        }
    } else {
        for (int k = 0; k < SCAN_ITEMS; k++) sum += (base + k < n) ? in[base + k] : 0;
    }
    sum = wave_sum(sum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_THREADS / 64; w++) t += red[w];
        tileSums[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(256) k_key_scatter(const int *__restrict__ key, int nItems, int nKeys,
                                                     const int *__restrict__ offsets, const int *__restrict__ slot,
                                                     int *__restrict__ counts, int *__restrict__ tmp, int contentsCap,
                                                     const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k < 0 || k >= nKeys) return;
    counts[k] = 0;
    int pos = offsets[k] + slot[m];
    if (pos < contentsCap) tmp[pos] = m;
}

__global__ void __launch_bounds__(256) k_key_place(const int *__restrict__ key, int nItems, int nKeys,
                                                   const int *__restrict__ offsets, const int *__restrict__ tmp,
                                                   int *__restrict__ list, int contentsCap,
                                                   const int *__restrict__ gate, SortPayload pay,
                                                   int *__restrict__ longCount, int *__restrict__ longList) {
    if (gate && *gate == 0) return;
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k < 0 || k >= nKeys) return;
    int start = offsets[k];
    int end = offsets[k + 1];
    if (end > contentsCap) end = contentsCap;
    if (end - start > LONG_SEG) {
        // a long segment is placed by k_sort_long; the item at its first position lists it once
        if (tmp[start] == m) longList[atomicAdd(longCount, 1)] = k;
        return;
    }
    int rank = 0;
    for (int p = start; p < end; p++) rank += (tmp[p] < m) ? 1 : 0;
    int pos = start + rank;
    if (pos < contentsCap) {
        list[pos] = m;
        if (pay.in) pay.out[pos] = pay.in[m];
        if (pay.keyOut) pay.keyOut[pos] = k;
    }
}

// Segments longer than LONG_SEG: one workgroup per segment (grid-stride over the listed keys).  Up to
// LONG_LDS items are sorted ascending in LDS by a bitonic network (padded with INT_MAX); a longer
// segment is placed by per-item rank over the whole segment, strided over the workgroup.
__global__ void __launch_bounds__(LONG_THREADS) k_sort_long(const int *__restrict__ offsets,
                                                            const int *__restrict__ tmp, int *__restrict__ list,
                                                            int contentsCap, const int *__restrict__ gate,
                                                            SortPayload pay, const int *__restrict__ longCount,
                                                            const int *__restrict__ longList) {
    if (gate && *gate == 0) return;
    __shared__ int v[LONG_LDS];
    const int nLong = *longCount;
    for (int q = blockIdx.x; q < nLong; q += gridDim.x) {
        const int k = longList[q];
        const int start = offsets[k];
        int end = offsets[k + 1];
        if (end > contentsCap) end = contentsCap;
        const int len = end - start;
        if (len <= LONG_LDS) {
            int pw = 1;
            while (pw < len) pw <<= 1;
            for (int i = threadIdx.x; i < pw; i += LONG_THREADS) v[i] = i < len ? tmp[start + i] : 0x7fffffff;
            __syncthreads();
            for (int size = 2; size <= pw; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int i = threadIdx.x; i < (pw >> 1); i += LONG_THREADS) {
                        const int lo = 2 * stride * (i / stride) + (i % stride);
                        const int hi = lo + stride;
                        const bool up = (lo & size) == 0;
                        const int a = v[lo], b = v[hi];
                        if ((a > b) == up) {
                            v[lo] = b;
                            v[hi] = a;
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = threadIdx.x; i < len; i += LONG_THREADS) {
                const int m = v[i];
                list[start + i] = m;
                if (pay.in) pay.out[start + i] = pay.in[m];
                if (pay.keyOut) pay.keyOut[start + i] = k;
            }
            __syncthreads();
        } else {
            for (int p = start + threadIdx.x; p < end; p += LONG_THREADS) {
                const int m = tmp[p];
                int rank = 0;
                for (int r = start; r < end; r++) rank += (tmp[r] < m) ? 1 : 0;
                list[start + rank] = m;
                if (pay.in) pay.out[start + rank] = pay.in[m];
                if (pay.keyOut) pay.keyOut[start + rank] = k;
            }
        }
    }
}

// offsets = exclusive scan of counts[nKeys]: tile sums, their prefix (one block), then the tiles
// (scanItems keys per thread)
template <int ITEMS>
void launch_scan_items(const SortScratch &s, const int *counts, int nKeys, int *offsets, int clampTotal,
                       const int *gate, hipStream_t q) {
    int nt = (nKeys + SCAN_THREADS * ITEMS - 1) / (SCAN_THREADS * ITEMS);
    if (nt < 1) nt = 1;
    hipLaunchKernelGGL(k_tile_reduce<ITEMS>, dim3(nt), dim3(SCAN_THREADS), 0, q, counts, nKeys, s.tileSums, gate);
    hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(1024), 0, q, s.tileSums, nt, s.tilePrefix, gate, s.longCount);
    hipLaunchKernelGGL(k_scan_tiles<ITEMS>, dim3(nt), dim3(SCAN_THREADS), 0, q, counts, nKeys, s.tilePrefix, nt,
                       offsets, clampTotal, gate);
}

void launch_scan(const SortScratch &s, const int *counts, int nKeys, int *offsets, int clampTotal, const int *gate,
                 hipStream_t q) {
    if (s.scanItems == 4) launch_scan_items<4>(s, counts, nKeys, offsets, clampTotal, gate, q);
    else if (s.scanItems == 16) launch_scan_items<16>(s, counts, nKeys, offsets, clampTotal, gate, q);
    else launch_scan_items<8>(s, counts, nKeys, offsets, clampTotal, gate, q);
}



// stable placement: short segments by per-item rank, long ones by a workgroup each
void launch_place(const SortScratch &s, const int *key, int nItems, int nKeys, const int *offsets, int *list,
                  int contentsCap, const int *gate, SortPayload pay, hipStream_t q) {
    hipLaunchKernelGGL(k_key_place, dim3(blocks_for(nItems, 256)), dim3(256), 0, q, key, nItems, nKeys, offsets,
                       s.tmp, list, contentsCap, gate, pay, s.longCount, s.longList);
    long long maxLong = nItems / (LONG_SEG + 1);
    if (maxLong < 1) return;  // no segment of these items can be long
    unsigned nb = (unsigned)(maxLong < LONG_BLOCKS ? maxLong : LONG_BLOCKS);
    hipLaunchKernelGGL(k_sort_long, dim3(nb), dim3(LONG_THREADS), 0, q, offsets, s.tmp, list, contentsCap, gate, pay,
                       s.longCount, s.longList);
}

int exclusive_scan(sbx_ctx *c, SortScratch &s, const int *in, int n, int *out, hipStream_t st) {
    int r = sort_scratch_reserve(c, s, n, 0);
    if (r) return r;
    launch_scan(s, in, n, out, -1, nullptr, st ? st : c->stream);
    return check_launch(c, "exclusive_scan");
}

int count_keys(sbx_ctx *c, const int *key, int nItems, int nKeys, int *counts, int *slot, hipStream_t st) {
    if (nItems <= 0) return SBX_OK;
    hipLaunchKernelGGL(k_key_count, dim3(blocks_for(nItems, 256)), dim3(256), 0, st ? st : c->stream, key, nItems,
                       nKeys, counts, slot, nullptr);
    return check_launch(c, "k_key_count");
}

int counting_sort_counted(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap,
                          int *offsets, int *list, int *counts, const int *slot, hipStream_t st, SortPayload pay) {
    int r = sort_scratch_reserve(c, s, nKeys, nItems);
    if (r) return r;
    hipStream_t q = st ? st : c->stream;
    launch_scan(s, counts, nKeys, offsets, contentsCap, nullptr, q);
    if (nItems > 0) {
        hipLaunchKernelGGL(k_key_scatter, dim3(blocks_for(nItems, 256)), dim3(256), 0, q, key, nItems, nKeys,
                           offsets, slot, counts, s.tmp, contentsCap, nullptr);
        launch_place(s, key, nItems, nKeys, offsets, list, contentsCap, nullptr, pay, q);
    }
    return check_launch(c, "counting_sort_counted");
}

// the sort on scratch that is already large enough, with counts[0, nKeys] and tileSums all zero
int counting_sort_core(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap,
                       int *offsets, int *list, const int *gate, hipStream_t st, SortPayload pay) {
    int r;
    hipStream_t q = st ? st : c->stream;
    if (nItems > 0) {
        hipLaunchKernelGGL(k_key_count, dim3(blocks_for(nItems, 256)), dim3(256), 0, q, key, nItems, nKeys,
                           s.counts, s.slot, gate);
    }
    launch_scan(s, s.counts, nKeys, offsets, contentsCap, gate, q);
    if (nItems > 0) {
        hipLaunchKernelGGL(k_key_scatter, dim3(blocks_for(nItems, 256)), dim3(256), 0, q, key, nItems, nKeys,
                           offsets, s.slot, s.counts, s.tmp, contentsCap, gate);
        launch_place(s, key, nItems, nKeys, offsets, list, contentsCap, gate, pay, q);
    }
    r = check_launch(c, "counting_sort");
    if (r) s.clean = false;
    return r;
}

int counting_sort(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap, int *offsets,
                  int *list, const int *gate, hipStream_t st, SortPayload pay) {
    int r = sort_scratch_reserve(c, s, nKeys, nItems);
    if (r) return r;
    return counting_sort_core(c, s, key, nItems, nKeys, contentsCap, offsets, list, gate, st, pay);
}

}  // namespace sbx

using namespace sbx;

// ------------------------------------------------------------------ C ABI: context + memory
extern "C" {

const char *sbx_version(void) { return "softbox_hip 0.1 gfx950"; }

int sbx_create(sbx_ctx **out, int device, void *hip_stream) {
    if (!out) return SBX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SBX_ERR_HIP;
    if (device < 0 || device >= ndev) return SBX_ERR_ARG;
    sbx_ctx *c = new (std::nothrow) sbx_ctx();  // value-initialised: zero fields, empty scratch
    if (!c) return SBX_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return SBX_ERR_HIP; }
    if (hip_stream) {
        c->stream = (hipStream_t)hip_stream;
        c->owns_stream = 0;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SBX_ERR_HIP; }
        c->owns_stream = 1;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (noalloc_setup(c) != SBX_OK) {  // this device's address of sbx_csr_build_noalloc's scratch
        if (c->owns_stream) hipStreamDestroy(c->stream);
        delete c;
        return SBX_ERR_HIP;
    }
    *out = c;
    return SBX_OK;
}

void sbx_destroy(sbx_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    sort_scratch_free(c->api_sort);
    if (c->form_csr) hipFree(c->form_csr);
    turn_scratch_destroy(c->turn);
    if (c->owns_stream) {
        hipStreamSynchronize(c->stream);
        hipStreamDestroy(c->stream);
    }
    delete c;
}

const char *sbx_last_error(const sbx_ctx *c) { return c ? c->err : "null context"; }

void sbx_clear_error(sbx_ctx *c) {
    if (!c) return;
    c->sticky = 0;
    c->err[0] = 0;
}

int sbx_sync(sbx_ctx *c) {
    SBX_CHECK_CTX(c);
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int sbx_alloc(sbx_ctx *c, void **dptr, size_t bytes) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dptr != nullptr, "sbx_alloc: null out pointer");
    *dptr = nullptr;
    int r = set_device(c);
    if (r) return r;
    return check_hip(c, hipMalloc(dptr, bytes ? bytes : 4), "hipMalloc");
}

int sbx_free(sbx_ctx *c, void *dptr) {
    if (!c) return SBX_ERR_ARG;
    if (!dptr) return SBX_OK;
    return check_hip(c, hipFree(dptr), "hipFree");
}

int sbx_push(sbx_ctx *c, void *dst, const void *src, size_t bytes) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dst && src, "sbx_push: null pointer");
    return check_hip(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "sbx_push");
}

int sbx_pull(sbx_ctx *c, void *dst, const void *src, size_t bytes) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dst && src, "sbx_pull: null pointer");
    int r = check_hip(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "sbx_pull");
    if (r) return r;
    return check_hip(c, hipStreamSynchronize(c->stream), "sbx_pull sync");
}

int sbx_copy(sbx_ctx *c, void *dst, const void *src, size_t bytes) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dst && src, "sbx_copy: null pointer");
    return check_hip(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream), "sbx_copy");
}

int sbx_fill_i32(sbx_ctx *c, int *dst, int value, size_t count) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, dst != nullptr, "sbx_fill_i32: null pointer");
    return fill_i32(c, dst, value, (long long)count);
}

}  // extern "C"
