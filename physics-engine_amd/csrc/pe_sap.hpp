// Sweep-and-prune broadphase over n AABBs on the whole GPU (BroadPhaseSweepAndPrune::
// calcCollisionPairs, broad_phase_sweep_and_prune.cpp:6-41), shared by the single-world pipeline
// (its body store) and the stage entry points (a caller's AABB records):
//   k_sap_keys + radix   stable LSD sort of (aabb_min[axis] + 0, index), 4 passes of 8 bits
//   k_sap_ranges/cand    per sorted body, the candidate run j > i with min_j <= max_i
//   k_sap_chunks         load-balanced sweep over all candidates, two passes (count, emit), so
//                        pairs come out in the reference's (i, j) sweep order
//   k_sap_axis_*         next axis = argmax variance of AABB centres (FP64, fixed blocks)
#pragma once
#include "pe_common.hpp"
#include "pe_prims.hpp"

namespace pe {

struct DevCounts {          // device-side counters of one step
    int32_t npairs;         // pairs emitted by the broadphase (clamped to capacity)
    int32_t prev_npairs;    // pairs of the previous narrowphase (contact slots that survive)
    int32_t axis;           // SAP axis for the next sort
    int32_t overflow;       // sticky capacity-error bits (OVF_*), cleared when the host reports them
    uint32_t ncand;         // total sweep candidates
    uint32_t nchunk_pairs;  // pairs counted by the sweep (unclamped)
    int32_t ncolors;
    int32_t nordered;       // pairs in the solve order
    int32_t npoints;        // contact points of the last narrowphase
    uint32_t nchunks;       // live sweep chunks
    uint32_t bar_count;     // grid barrier: arrivals (0 between barriers)
    uint32_t bar_gen;       // grid barrier: generation
    int32_t nshape;         // pairs of this narrowphase that involve a cylinder or a mesh ...
    int32_t nsat;           // ... of which these run the polyhedral separating-axis routine
    int32_t ncomp;          // ... and these involve a compound body
    int32_t nact;           // pairs the colouring orders (contact points, not both bodies kinematic)
    uint32_t jp_rem[3];     // colouring rounds: pairs left uncoloured, round r in slot r % 3 (a slot is
                            // cleared two rounds before it is counted, one round after it is read)
    uint32_t adj_total;     // colouring: adjacency entries
    int32_t jp_rounds;      // colouring: rounds the last colouring took
};
// Capacity errors, OR-ed into DevCounts::overflow by the step that hits them and kept until the
// host reads and reports them: a result that lost pairs or contact points is never returned as a
// success.
constexpr int OVF_PAIRS = 1;         // the broadphase found more pairs than the pair capacity
constexpr int OVF_CHUNKS = 4;        // the sweep had more candidates than its chunk capacity
constexpr int OVF_CLIP = 8;          // a clipped polygon outgrew its buffer (contact points lost)

constexpr int SAP_CHUNK = 32;        // candidates per thread in the sweep
static_assert(SAP_CHUNK <= 32, "a chunk's test outcomes are one 32-bit mask");
constexpr int SAP_RED_THREADS = 256;
constexpr int SAP_RED_VALUES = 7;   // axis reduction: centre sums (3), squared sums (3), bodies

// The AABBs and the working memory of one sweep.  AABB minimum k of body i is
// amin[k * fs + i * bs] (field-major SoA: fs = stride, bs = 1; records: fs = 1, bs = 8); rec holds
// one 8-float record per body: min xyz, kinematic flag, max xyz, ignore-collision flag (non-zero:
// the body is in no pair), read by the sweep's pair test.
struct SapView {
    int n;
    const float* amin;
    const float* amax;
    int fs, bs;
    const float* rec;
    uint32_t* keys;         // [n] x 2 radix buffers
    uint32_t* vals;
    uint32_t* keys2;
    uint32_t* vals2;
    float* skey;            // [n] sorted keys as floats
    float* srec;            // [n][8] the AABB records in sorted order (or null: read through vals)
    uint32_t* cand;         // [n + 1] candidate counts -> offsets
    uint32_t* chunk;        // [chunk_cap] pairs per chunk -> offsets
    uint32_t* chunk_hit;    // [chunk_cap] per chunk, bit k: its candidate k is a pair
    uint32_t* scan_part;    // [scan_parts(n, chunk_cap)] tile totals of the grid-wide scans
    uint32_t* hist;         // [256 * radix_blocks(n)] radix histograms
    double* red;            // [6 * red_blocks(n)] axis reduction partials
    DevCounts* cnt;
    int32_t* pa;            // [pcap] pair output: body A, body B (sorted position order)
    int32_t* pb;
    int pcap;
    int chunk_cap;
    const uint64_t* ignore; // [nignore] sorted keys (lo << 32 | hi) of body pairs that never collide
    int nignore;
    __device__ __forceinline__ float mn(int k, int i) const { return amin[(size_t)k * fs + (size_t)i * bs]; }
    __device__ __forceinline__ float mx(int k, int i) const { return amax[(size_t)k * fs + (size_t)i * bs]; }
};

__host__ inline int sap_radix_blocks(int n) { return (n + RS_TILE - 1) / RS_TILE; }
__host__ inline int sap_red_blocks(int n) { return n < 1 ? 1 : (n + SAP_RED_THREADS - 1) / SAP_RED_THREADS < 256 ? (n + SAP_RED_THREADS - 1) / SAP_RED_THREADS : 256; }
__host__ inline int sap_scan_parts(int n, int chunk_cap) { return ((chunk_cap > n + 1 ? chunk_cap : n + 1) + SCAN_TILE - 1) / SCAN_TILE + 1; }

// BroadPhaseBase::validateCollisionPair (broad_phase_base.cpp:5-19) on two AABB records: the
// ignore-collision flags, both kinematic, then closed-interval overlap on x, y, z.  Ignore-id pairs
// (broad_phase_base.cpp:12-13) are tested by sap_pair_allowed.
__device__ __forceinline__ bool sap_pair_ok(const float (&A)[8], const float (&B)[8]) {
    if (A[7] == COLL_IGNORE || B[7] == COLL_IGNORE) return false;
    if (A[3] != 0.0f && B[3] != 0.0f) return false;
    if (A[0] > B[4] || B[0] > A[4]) return false;
    if (A[1] > B[5] || B[1] > A[5]) return false;
    if (A[2] > B[6] || B[2] > A[6]) return false;
    return true;
}
// sap_pair_ok, then for two bodies that both have ignore partners, the ignore-pair list (binary
// search of the pair's key).
__device__ __forceinline__ bool sap_pair_allowed(const SapView& s, const float (&A)[8], const float (&B)[8], int i, int j) {
    if (!sap_pair_ok(A, B)) return false;
    if (A[7] != COLL_PARTNERS || B[7] != COLL_PARTNERS) return true;
    const uint64_t key = i < j ? ((uint64_t)i << 32 | (uint32_t)j) : ((uint64_t)j << 32 | (uint32_t)i);
    int lo = 0, hi = s.nignore;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s.ignore[mid] < key) lo = mid + 1; else hi = mid;
    }
    return !(lo < s.nignore && s.ignore[lo] == key);
}
__device__ __forceinline__ void sap_ld_rec(const float* src, float (&dst)[8]) {
    const float4* s = reinterpret_cast<const float4*>(src);
    const float4 a = s[0], b = s[1];
    dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
    dst[4] = b.x; dst[5] = b.y; dst[6] = b.z; dst[7] = b.w;
}
// The record of the body at sorted position q: from the sorted copy (contiguous loads along a run
// of candidates) when there is one.
__device__ __forceinline__ void sap_ld_sorted(const SapView& s, int q, float (&dst)[8]) {
    if (s.srec) sap_ld_rec(s.srec + (size_t)q * 8, dst);
    else sap_ld_rec(s.rec + (size_t)s.vals[q] * 8, dst);
}
// sap_pair_allowed for the bodies at sorted positions qa and qb (their indices are loaded only
// for the ignore-pair search).
__device__ __forceinline__ bool sap_pair_allowed_sorted(const SapView& s, const float (&A)[8], const float (&B)[8],
                                                        int qa, int qb) {
    if (!sap_pair_ok(A, B)) return false;
    if (A[7] != COLL_PARTNERS || B[7] != COLL_PARTNERS) return true;
    return sap_pair_allowed(s, A, B, (int)s.vals[qa], (int)s.vals[qb]);
}

// Sets the axis of this sweep and clears the overflow bits a previous stage call left (the
// single-world pipeline keeps both in its counters and does not launch this).
static __global__ void k_sap_begin(SapView s, int axis) {
    s.cnt->axis = axis;
    s.cnt->overflow = 0;
    s.cnt->npairs = 0;
    s.cnt->nchunk_pairs = 0;
}

static __global__ void k_sap_keys(SapView s) {
    const int ax = s.cnt->axis;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        s.keys[i] = float_key(s.mn(ax, i) + 0.0f);
        s.vals[i] = (uint32_t)i;
    }
}

// After the sort: sorted float keys and, per sorted position s, the number of candidates
// j > s with key_j <= max_i (the reference's sweep breaks at the first key_j > max_i).
static __global__ void k_sap_ranges(SapView s) {
    const int ax = s.cnt->axis;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < s.n; q += gridDim.x * blockDim.x) {
        const int i = (int)s.vals[q];
        s.skey[q] = s.mn(ax, i) + 0.0f;
        if (s.srec) {
            const float4* src = reinterpret_cast<const float4*>(s.rec + (size_t)i * 8);
            float4* dst = reinterpret_cast<float4*>(s.srec + (size_t)q * 8);
            dst[0] = src[0];
            dst[1] = src[1];
        }
    }
}
static __global__ void k_sap_cand(SapView s) {
    const int n = s.n;
    const int ax = s.cnt->axis;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i = (int)s.vals[q];
        const float m = s.mx(ax, i);
        int lo = q + 1, hi = n;  // first position in [q+1, n) with skey > m
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s.skey[mid] > m) hi = mid; else lo = mid + 1;
        }
        s.cand[q] = (uint32_t)(lo - q - 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) s.cand[n] = 0;
}

// Locates candidate t: sorted position q with cand[q] <= t < cand[q+1] (cand holds offsets).
__device__ __forceinline__ int sap_find(const uint32_t* off, int n, uint32_t t) {
    int lo = 0, hi = n;  // last q with off[q] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool EMIT>
static __global__ void k_sap_chunks(SapView s) {
    const int n = s.n;
    const uint32_t S = s.cnt->ncand;
    const uint32_t nchunks = min((S + SAP_CHUNK - 1) / SAP_CHUNK, (uint32_t)s.chunk_cap);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        uint32_t t = c * SAP_CHUNK;
        const uint32_t tend = min(S, t + SAP_CHUNK);
        int q = sap_find(s.cand, n, t);
        while (s.cand[q + 1] <= t) q++;
        if (EMIT) {
            // the counting pass left the outcome of each of the chunk's tests in chunk_hit[c]
            uint32_t out = s.chunk[c];
            uint32_t hit = s.chunk_hit[c];
            const uint32_t t0 = t;
            int cur = q;
            uint32_t cb = s.cand[cur], nb = s.cand[cur + 1];
            while (hit) {
                const uint32_t tt = t0 + (uint32_t)(__ffs(hit) - 1);
                hit &= hit - 1;
                while (nb <= tt) {
                    cur++;
                    cb = nb;
                    nb = s.cand[cur + 1];
                }
                if ((int)out < s.pcap) {
                    s.pa[out] = (int)s.vals[cur];
                    s.pb[out] = (int)s.vals[cur + 1 + (int)(tt - cb)];
                }
                out++;
            }
        } else {
            uint32_t cntv = 0, hit = 0;
            const uint32_t t0 = t;
            int cur = q;
            // the run's first candidate and the next run's, in registers: a candidate's record is then
            // addressed without a load of the offsets (only a change of run reads them)
            uint32_t cb = s.cand[cur], nb = s.cand[cur + 1];
            float A[8], B[8];
            sap_ld_sorted(s, cur, A);   // the run body's record stays in registers
            for (; t < tend; t++) {
                while (nb <= t) {
                    cur++;
                    cb = nb;
                    nb = s.cand[cur + 1];
                    sap_ld_sorted(s, cur, A);
                }
                const int qj = cur + 1 + (int)(t - cb);
                sap_ld_sorted(s, qj, B);
                if (sap_pair_allowed_sorted(s, A, B, cur, qj)) {
                    cntv++;
                    hit |= 1u << (t - t0);
                }
            }
            s.chunk[c] = cntv;
            s.chunk_hit[c] = hit;
        }
    }
}

// Number of live sweep chunks, for the chunk scan.
static __global__ void k_sap_nchunks(SapView s) {
    const uint32_t want = (s.cnt->ncand + SAP_CHUNK - 1) / SAP_CHUNK;
    s.cnt->nchunks = min(want, (uint32_t)s.chunk_cap);
    if (want > (uint32_t)s.chunk_cap) s.cnt->overflow |= OVF_CHUNKS;
}

static __global__ void k_sap_finish(SapView s) {
    const uint32_t total = s.cnt->nchunk_pairs;
    if (total > (uint32_t)s.pcap) s.cnt->overflow |= OVF_PAIRS;
    s.cnt->npairs = total > (uint32_t)s.pcap ? s.pcap : (int32_t)total;
}

// Next SAP axis (broad_phase_sweep_and_prune.cpp:37-40): variance of AABB centres, reduced in
// FP64 over fixed blocks (deterministic), argmax with the reference's tie rule.
// A body whose AABB is empty (min > max) has left the world (World::removeRigidBody, an ended
// lifetime): it sorts last with an infinite key, has no candidates and is not in the variance; with
// fewer than two bodies left the axis stays (the reference's sweep returns early).
static __global__ void __launch_bounds__(SAP_RED_THREADS) k_sap_axis_partial(SapView s) {
    __shared__ double sh[SAP_RED_VALUES][SAP_RED_THREADS];
    double acc[SAP_RED_VALUES] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        if (!(s.mn(0, i) <= s.mx(0, i))) continue;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float c = (s.mn(k, i) + s.mx(k, i)) * 0.5f;
            acc[k] += (double)c;
            acc[3 + k] += (double)(c * c);
        }
        acc[6] += 1.0;
    }
#pragma unroll
    for (int k = 0; k < SAP_RED_VALUES; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = SAP_RED_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < SAP_RED_VALUES; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < SAP_RED_VALUES; k++) s.red[k * gridDim.x + blockIdx.x] = sh[k][0];
}
static __global__ void k_sap_axis_final(SapView s, int nblocks) {
    if (threadIdx.x != 0) return;
    double t[SAP_RED_VALUES] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblocks; b++)
        for (int k = 0; k < SAP_RED_VALUES; k++) t[k] += s.red[k * nblocks + b];
    if (t[6] < 2.0) return;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = t[3 + k] - t[k] * t[k] / t[6];
    int axis = 0;
    if (v[1] > v[0]) axis = 1;
    if (v[2] > v[axis]) axis = 2;
    s.cnt->axis = axis;
}

// Grid-wide exclusive scan of n elements (of *count when count is given; at most cap).
inline void sap_scan(const SapView& s, uint32_t* data, int cap, int n, const uint32_t* count, uint32_t* total, hipStream_t st) {
    const int nb = (cap + SCAN_TILE - 1) / SCAN_TILE < 1 ? 1 : (cap + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(1024), 0, st, data, n, count, s.scan_part);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, st, s.scan_part, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, st, data, n, count, s.scan_part);
}

// The whole sweep on stream st (no host synchronisation): pairs to s.pa / s.pb, their count and
// the next axis to s.cnt.  n >= 1.
inline void sap_launch(const SapView& s, hipStream_t st) {
    const int n = s.n;
    auto grid = [](int work, int block) { const int g = (work + block - 1) / block; return g < 1 ? 1 : (g > 4096 ? 4096 : g); };
    const int rsb = sap_radix_blocks(n);
    hipLaunchKernelGGL(k_sap_keys, dim3(grid(n, 256)), dim3(256), 0, st, s);
    uint32_t *ki = s.keys, *vi = s.vals, *ko = s.keys2, *vo = s.vals2;
    for (int pass = 0; pass < 4; pass++) {
        hipLaunchKernelGGL(k_radix_hist, dim3(rsb), dim3(RS_THREADS), 0, st, ki, n, pass * 8, s.hist, rsb);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, st, s.hist, 256 * rsb, (uint32_t*)nullptr);
        hipLaunchKernelGGL(k_radix_scatter, dim3(rsb), dim3(RS_THREADS), 0, st, ki, vi, ko, vo, n, pass * 8, s.hist, rsb);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    // after 4 passes the sorted data is back in keys / vals
    hipLaunchKernelGGL(k_sap_ranges, dim3(grid(n, 256)), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_sap_cand, dim3(grid(n, 256)), dim3(256), 0, st, s);
    sap_scan(s, s.cand, n + 1, n + 1, nullptr, &s.cnt->ncand, st);
    hipLaunchKernelGGL(k_sap_nchunks, dim3(1), dim3(1), 0, st, s);
    hipLaunchKernelGGL(k_sap_chunks<false>, dim3(2048), dim3(256), 0, st, s);
    sap_scan(s, s.chunk, s.chunk_cap, 0, &s.cnt->nchunks, &s.cnt->nchunk_pairs, st);
    hipLaunchKernelGGL(k_sap_chunks<true>, dim3(2048), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_sap_finish, dim3(1), dim3(1), 0, st, s);
    const int rb = sap_red_blocks(n);
    hipLaunchKernelGGL(k_sap_axis_partial, dim3(rb), dim3(SAP_RED_THREADS), 0, st, s);
    hipLaunchKernelGGL(k_sap_axis_final, dim3(1), dim3(64), 0, st, s, rb);
}

}  // namespace pe
