// sbx_core.hip -- context, device memory, error plumbing, scan and stable counting sort.
//
// The counting sort is the one primitive behind SpatialGrid's cell CSR (SpatialGrid.java:99-331)
// and CrossBridgeSystem's CSR inverse (CrossBridgeSystem.java:190-298).  The reference builds it
// with per-chunk private histogram rows so that the within-key order is ascending item index.  On
// gfx950 the same unique result is produced by: atomic key histogram -> wave64 shuffle scan ->
// atomic (unordered) scatter -> per-item rank placement (rank = number of same-key items with a
// smaller index, read back from the unordered segment).  The totals and the final order do not
// depend on atomic ordering, so the CSR is bit-identical to the serial reference build.
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
// Tile = 256 threads x 4 items.  Wave64 inclusive scan by __shfl_up, then wave totals in LDS.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
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

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const int *__restrict__ in, int n,
                                                              int *__restrict__ blockSums) {
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        long long idx = base + k;
        if (idx < n) s += in[idx];
    }
    int total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

// single block: exclusive scan of blockSums[nb] in place, blockSums[nb] = grand total
__global__ void __launch_bounds__(1024) k_scan_blocksums(int *__restrict__ blockSums, int nb) {
    __shared__ int part[1024];
    int per = (nb + 1023) / 1024;
    int start = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; k++) {
        int idx = start + k;
        if (idx < nb) s += blockSums[idx];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele over 1024 partials (runs once per sort; nb is small)
    for (int d = 1; d < 1024; d <<= 1) {
        int v = (threadIdx.x >= (unsigned)d) ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int acc = part[threadIdx.x] - s;  // exclusive prefix of this thread's chunk
    for (int k = 0; k < per; k++) {
        int idx = start + k;
        if (idx < nb) {
            int v = blockSums[idx];
            blockSums[idx] = acc;
            acc += v;
        }
    }
    if (threadIdx.x == 1023) blockSums[nb] = part[1023];
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const int *__restrict__ in, int n,
                                                             const int *__restrict__ blockSums,
                                                             int *__restrict__ out, int nb, int clampTotal) {
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        long long idx = base + k;
        v[k] = (idx < n) ? in[idx] : 0;
        s += v[k];
    }
    int total;
    int pre = block_exclusive_scan(s, &total) + blockSums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        long long idx = base + k;
        if (idx < n) out[idx] = pre;
        pre += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int grand = blockSums[nb];
        out[n] = (clampTotal >= 0 && grand > clampTotal) ? clampTotal : grand;
    }
}

int sort_scratch_reserve(sbx_ctx *c, SortScratch &s, long long nKeys, long long nItems) {
    long long nb = (nKeys + SCAN_TILE - 1) / SCAN_TILE + 1;
    if (nKeys > s.keyCap) {
        if (s.counts) hipFree(s.counts);
        if (s.cursor) hipFree(s.cursor);
        s.counts = s.cursor = nullptr;
        int r = check_hip(c, hipMalloc(&s.counts, sizeof(int) * (size_t)(nKeys + 1)), "hipMalloc counts");
        if (r) return r;
        r = check_hip(c, hipMalloc(&s.cursor, sizeof(int) * (size_t)(nKeys + 1)), "hipMalloc cursor");
        if (r) return r;
        s.keyCap = nKeys;
    }
    if (nItems > s.itemCap) {
        if (s.tmp) hipFree(s.tmp);
        s.tmp = nullptr;
        int r = check_hip(c, hipMalloc(&s.tmp, sizeof(int) * (size_t)(nItems + 1)), "hipMalloc tmp");
        if (r) return r;
        s.itemCap = nItems;
    }
    if (nb > s.blockCap) {
        if (s.blockSums) hipFree(s.blockSums);
        s.blockSums = nullptr;
        int r = check_hip(c, hipMalloc(&s.blockSums, sizeof(int) * (size_t)(nb + 1)), "hipMalloc blockSums");
        if (r) return r;
        s.blockCap = nb;
    }
    return SBX_OK;
}

void sort_scratch_free(SortScratch &s) {
    if (s.counts) hipFree(s.counts);
    if (s.cursor) hipFree(s.cursor);
    if (s.tmp) hipFree(s.tmp);
    if (s.blockSums) hipFree(s.blockSums);
    s = SortScratch();
}

int exclusive_scan(sbx_ctx *c, SortScratch &s, const int *in, int *out, int n, int clampTotal) {
    int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, c->stream, in, n, s.blockSums);
    hipLaunchKernelGGL(k_scan_blocksums, dim3(1), dim3(1024), 0, c->stream, s.blockSums, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, c->stream, in, n, s.blockSums, out, nb,
                       clampTotal);
    return check_launch(c, "exclusive_scan");
}

// ------------------------------------------------------------------ counting sort
__global__ void __launch_bounds__(256) k_key_count(const int *__restrict__ key, int nItems, int *__restrict__ counts) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k >= 0) atomicAdd(&counts[k], 1);
}

__global__ void __launch_bounds__(256) k_key_scatter(const int *__restrict__ key, int nItems,
                                                     const int *__restrict__ offsets, int *__restrict__ cursor,
                                                     int *__restrict__ tmp, int contentsCap) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k < 0) return;
    int pos = offsets[k] + atomicAdd(&cursor[k], 1);
    if (pos < contentsCap) tmp[pos] = m;
}

// rank placement: the key's segment in tmp holds exactly its items (any order); count smaller ids.
__global__ void __launch_bounds__(256) k_key_place(const int *__restrict__ key, int nItems,
                                                   const int *__restrict__ offsets, const int *__restrict__ counts,
                                                   const int *__restrict__ tmp, int *__restrict__ list,
                                                   int contentsCap) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nItems) return;
    int k = key[m];
    if (k < 0) return;
    int start = offsets[k];
    int cnt = counts[k];
    int end = start + cnt;
    if (end > contentsCap) end = contentsCap;
    int rank = 0;
    for (int p = start; p < end; p++) rank += (tmp[p] < m) ? 1 : 0;
    int pos = start + rank;
    if (pos < contentsCap) list[pos] = m;
}

int counting_sort(sbx_ctx *c, SortScratch &s, const int *key, int nItems, int nKeys, int contentsCap, int *offsets,
                  int *list) {
    int r = sort_scratch_reserve(c, s, nKeys, nItems);
    if (r) return r;
    r = fill_i32(c, s.counts, 0, nKeys);
    if (r) return r;
    r = fill_i32(c, s.cursor, 0, nKeys);
    if (r) return r;
    if (nItems > 0) {
        hipLaunchKernelGGL(k_key_count, dim3(blocks_for(nItems, 256)), dim3(256), 0, c->stream, key, nItems,
                           s.counts);
        r = check_launch(c, "k_key_count");
        if (r) return r;
    }
    r = exclusive_scan(c, s, s.counts, offsets, nKeys, contentsCap);
    if (r) return r;
    if (nItems > 0) {
        hipLaunchKernelGGL(k_key_scatter, dim3(blocks_for(nItems, 256)), dim3(256), 0, c->stream, key, nItems,
                           offsets, s.cursor, s.tmp, contentsCap);
        hipLaunchKernelGGL(k_key_place, dim3(blocks_for(nItems, 256)), dim3(256), 0, c->stream, key, nItems,
                           offsets, s.counts, s.tmp, list, contentsCap);
        r = check_launch(c, "counting_sort scatter/place");
    }
    return r;
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
    *out = c;
    return SBX_OK;
}

void sbx_destroy(sbx_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    sort_scratch_free(c->api_sort);
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
