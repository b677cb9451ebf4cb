// sbx_noalloc.hip -- sbx_csr_build_noalloc: the stable counting sort of sbx_csr_build with no device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It has no counterpart in the reference.
//
// Its scratch is one file-scope device array, `scratch`, of a fixed size (every GPU the library runs
// on holds its own copy until the process ends), laid out as
//   counts[keys + 1] | tileSums[tiles + 1] | tilePrefix[tiles + 1] | tmp[items + 1] | slot[items + 1] |
//   longCount[1] | longList[items / 33 + 1]
// for at most SBX_NOALLOC_MAX_KEYS keys and SBX_NOALLOC_MAX_ITEMS items.  In host code the name
// `scratch` is a host-side stand-in, not a device address: noalloc_setup (called by sbx_create, with
// the context's device current) resolves the array's address on that device once with
// hipGetSymbolAddress and keeps it per device, and the entry point passes only that address to kernels.
// Every call on one GPU shares that GPU's array and finds the previous call's contents in it, so each
// call clears what the sort needs at zero (with a fill kernel); calls that overlap on one GPU race.
#include <mutex>

#include "sbx_internal.h"

namespace sbx {
namespace {

constexpr long long NA_TILES = SBX_NOALLOC_MAX_KEYS / SCAN_MIN_TILE_KEYS + 2;
constexpr long long NA_COUNTS = 0;
constexpr long long NA_TILESUMS = NA_COUNTS + SBX_NOALLOC_MAX_KEYS + 1;
constexpr long long NA_TILEPREFIX = NA_TILESUMS + NA_TILES + 1;
constexpr long long NA_TMP = NA_TILEPREFIX + NA_TILES + 1;
constexpr long long NA_SLOT = NA_TMP + SBX_NOALLOC_MAX_ITEMS + 1;
constexpr long long NA_LONG = NA_SLOT + SBX_NOALLOC_MAX_ITEMS + 1;
constexpr long long NA_TOTAL = NA_LONG + 1 + long_list_len(SBX_NOALLOC_MAX_ITEMS);

static __device__ __align__(256) int scratch[NA_TOTAL];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
int *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (null: not resolved)

}  // namespace

int noalloc_setup(sbx_ctx *c) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, SBX_ERR_ARG, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[c->device]) return SBX_OK;
    void *ptr = nullptr;
    int r = check_hip(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (r) return r;
    if (!ptr) return fail(c, SBX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch[c->device] = static_cast<int *>(ptr);
    return SBX_OK;
}

}  // namespace sbx

using namespace sbx;

extern "C" {

int sbx_csr_build_noalloc(sbx_ctx *c, const int *key, int nItems, int nKeys, int *offsets, int *list) {
    SBX_CHECK_CTX(c);
    SBX_REQUIRE(c, nItems >= 0 && nKeys >= 0 && offsets, "csr_build_noalloc: bad sizes or null offsets");
    SBX_REQUIRE(c, nItems == 0 || (key && list), "csr_build_noalloc: null key/list");
    SBX_REQUIRE(c, nKeys <= SBX_NOALLOC_MAX_KEYS && nItems <= SBX_NOALLOC_MAX_ITEMS,
                "csr_build_noalloc: %d keys and %d items exceed the static scratch (at most %d keys and %d items)",
                nKeys, nItems, SBX_NOALLOC_MAX_KEYS, SBX_NOALLOC_MAX_ITEMS);
    int *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!base) return fail(c, SBX_ERR_STATE, "csr_build_noalloc: scratch of device %d not resolved", c->device);
    SortScratch s;
    s.owned = false;
    s.counts = base + NA_COUNTS;
    s.tileSums = base + NA_TILESUMS;
    s.tilePrefix = base + NA_TILEPREFIX;
    s.tmp = base + NA_TMP;
    s.slot = base + NA_SLOT;
    s.longCount = base + NA_LONG;
    s.longList = base + NA_LONG + 1;
    s.keyCap = nKeys;
    s.itemCap = nItems;
    s.tileCap = (nKeys + SCAN_MIN_TILE_KEYS - 1) / SCAN_MIN_TILE_KEYS + 1;
    // the previous call's contents are still there: zero the histogram and the tile sums
    int r = fill_i32(c, s.counts, 0, (long long)nKeys + 1);
    if (!r) r = fill_i32(c, s.tileSums, 0, s.tileCap + 1);
    if (r) return r;
    s.clean = true;
    return counting_sort_core(c, s, key, nItems, nKeys, nItems, offsets, list);
}

}  // extern "C"
