// pe_stage_broadphase_noalloc: the sweep-and-prune stage of pe_stage_broadphase for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its working memory is one file-scope device array, `scratch`, of a fixed
// size: each GPU holds its own copy for the life of the process.  Host code never passes the
// array's name to a kernel (on the host it only names the symbol): the array's device address is
// looked up once per device with hipGetSymbolAddress, with that device current, and that address
// is what the kernels get.  Every call on a GPU shares that GPU's array, so calls that overlap on
// one GPU race; calls on one stream, one after another, do not.
#include "pe_common.hpp"
#include "pe_sap.hpp"
#include "../../include/pe_amd_noalloc.h"

#include <mutex>

using namespace pe;

namespace {

constexpr int NA_MAX_BODIES = PE_NOALLOC_MAX_BODIES;
constexpr int NA_CHUNK_CAP = 1 << 20;   // sweep chunks: 32 Mi candidate tests

// Byte offsets of the working arrays inside the scratch array (each 256-byte aligned).
struct NoAllocLayout {
    size_t keys, vals, keys2, vals2, skey, cand, hist, red, chunk, chunk_hit, scan_part, cnt, end;
    constexpr NoAllocLayout()
        : keys(0), vals(0), keys2(0), vals2(0), skey(0), cand(0), hist(0), red(0), chunk(0), chunk_hit(0), scan_part(0),
          cnt(0), end(0) {
        size_t o = 0;
        auto take = [&o](size_t bytes) { const size_t at = o; o += (bytes + 255) & ~(size_t)255; return at; };
        keys = take(4 * (size_t)NA_MAX_BODIES);
        vals = take(4 * (size_t)NA_MAX_BODIES);
        keys2 = take(4 * (size_t)NA_MAX_BODIES);
        vals2 = take(4 * (size_t)NA_MAX_BODIES);
        skey = take(4 * (size_t)NA_MAX_BODIES);
        cand = take(4 * ((size_t)NA_MAX_BODIES + 1));
        hist = take(4 * 256 * (size_t)((NA_MAX_BODIES + RS_TILE - 1) / RS_TILE));
        red = take(8 * SAP_RED_VALUES * 256);
        chunk = take(4 * (size_t)NA_CHUNK_CAP);
        chunk_hit = take(4 * (size_t)NA_CHUNK_CAP);
        scan_part = take(4 * (size_t)(((NA_CHUNK_CAP > NA_MAX_BODIES + 1 ? NA_CHUNK_CAP : NA_MAX_BODIES + 1) + SCAN_TILE - 1) /
                                          SCAN_TILE + 1));
        cnt = take(sizeof(DevCounts));
        end = o;
    }
};
constexpr NoAllocLayout NA_LAYOUT{};
static_assert(NA_LAYOUT.end <= PE_NOALLOC_SCRATCH_BYTES, "the working arrays fit the scratch array");

// result[0] = pairs found, result[1] = next axis, result[2] = capacity-error bits (as pe_stage_broadphase)
__global__ void k_noalloc_result(SapView s, int32_t* result) {
    result[0] = (int32_t)s.cnt->nchunk_pairs;
    result[1] = s.cnt->axis;
    result[2] = s.cnt->overflow;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[PE_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[NA_MAX_DEVICES] = {};

// The scratch array's device address on the current device, looked up on the first call there.
int scratch_address(unsigned char** out) {
    int dev = 0;
    PE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) { pe_set_error("device index beyond the no-allocation table"); return PE_ERR_DEVICE; }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            pe_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return PE_ERR_DEVICE;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return PE_OK;
}

}  // namespace

extern "C" int pe_stage_broadphase_noalloc(int32_t n, const float* aabb, int32_t axis, int32_t* pair_a, int32_t* pair_b,
                                           int32_t cap, int32_t* result, pe_stream_t stream) {
    if (n > NA_MAX_BODIES) {
        pe_set_error("pe_stage_broadphase_noalloc: " + std::to_string(n) + " bodies exceed its fixed working memory (" +
                     std::to_string(NA_MAX_BODIES) + " bodies)");
        return PE_ERR_INVALID;
    }
    if (n < 0 || axis < 0 || axis > 2 || cap < 0 || !result || (n > 0 && !aabb) || (cap > 0 && (!pair_a || !pair_b))) {
        pe_set_error("pe_stage_broadphase_noalloc: invalid argument");
        return PE_ERR_INVALID;
    }
    pe_clear_stale_error();
    unsigned char* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    SapView s{};
    s.n = n;
    s.amin = aabb;
    s.amax = aabb ? aabb + 4 : nullptr;
    s.fs = 1;
    s.bs = 8;
    s.rec = aabb;
    s.keys = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.keys);
    s.vals = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.vals);
    s.keys2 = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.keys2);
    s.vals2 = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.vals2);
    s.skey = reinterpret_cast<float*>(base + NA_LAYOUT.skey);
    s.cand = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.cand);
    s.hist = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.hist);
    s.red = reinterpret_cast<double*>(base + NA_LAYOUT.red);
    s.chunk = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.chunk);
    s.chunk_hit = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.chunk_hit);
    s.scan_part = reinterpret_cast<uint32_t*>(base + NA_LAYOUT.scan_part);
    s.cnt = reinterpret_cast<DevCounts*>(base + NA_LAYOUT.cnt);
    s.pa = pair_a;
    s.pb = pair_b;
    s.pcap = cap;
    s.chunk_cap = NA_CHUNK_CAP;
    hipStream_t st = (hipStream_t)stream;
    // the counters the sweep accumulates into start from zero, whatever the last call left
    hipLaunchKernelGGL(k_sap_begin, dim3(1), dim3(1), 0, st, s, (int)axis);
    if (n >= 2) sap_launch(s, st);
    hipLaunchKernelGGL(k_noalloc_result, dim3(1), dim3(1), 0, st, s, result);
    PE_HIP(hipGetLastError());
    return PE_OK;
}
