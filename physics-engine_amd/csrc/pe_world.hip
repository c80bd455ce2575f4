// Single-world pipeline (the plain drop-in for one World::step) and its C ABI.
//
// Per step, on one HIP stream, no host synchronisation:
//   k_integrate        World::updateObjectStatus + applyExternalForce    (thread / body)
//   k_bp_keys + radix  BroadPhaseSweepAndPrune sort on aabb_min[axis]    (stable LSD, 4 passes)
//   k_bp_ranges/scan   per sorted body, the candidate run j > i with min_j <= max_i
//   k_bp_chunk_*       load-balanced sweep over all candidates, two passes (count, emit),
//                      so pairs come out in the reference's (i, j) sweep order
//   k_bp_axis          next axis = argmax variance of AABB centres
//   k_narrow           thread / pair: shape dispatch + 8-slot contact cache + depth sort
//   k_setup            thread / pair: solver rows (FrictionContactConstraint::initSequentialImpulse)
//   k_colour           one 1024-thread workgroup: constraint-graph colouring (claims in LDS)
//   k_gs_grid          cooperative launch over the whole GPU: Gauss-Seidel sweeps colour by
//                      colour with a grid barrier between colours, then velocity sync
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_prims.hpp"
#include "pe_pipeline.hpp"
#include "pe_host.hpp"
#include "../../include/pe_amd.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

static thread_local std::string g_err;
void pe_set_error(const std::string& msg) { g_err = msg; }

using namespace pe;

namespace {

constexpr int MAX_COLORS = 1024;
constexpr int CHUNK = 32;            // candidates per thread in the sweep
static_assert(CHUNK <= 32, "a chunk's test outcomes are one 32-bit mask");
constexpr int SOLVE_THREADS = 1024;
constexpr int CLAIM_LDS = 12288;     // bodies whose colouring claims fit in LDS (48 KiB)
// Capacity errors, OR-ed into DevCounts::overflow by the step that hits them and kept until the
// host reads and reports them (pe_world_sync, pe_world_broadphase): a result that lost pairs or
// contact points is never returned as a success.
constexpr int OVF_PAIRS = 1;         // the broadphase found more pairs than the pair capacity
constexpr int OVF_CHUNKS = 4;        // the sweep had more candidates than its chunk capacity
constexpr int OVF_CLIP = 8;          // a clipped polygon outgrew its buffer (contact points lost)

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
    int32_t pad[2];
};

struct Dev {
    BodyView b;
    PairStore ps;           // pairs [pcap], contact slots and rows [pcap*8]
    int32_t* shape_list;    // [pcap] pairs handed from k_narrow to k_narrow_shapes (from the front)
                            // and to k_narrow_sat (from the back)
    int32_t* order;         // [pcap]
    int32_t* act;           // [3*pcap] colouring work lists
    int32_t* color_start;   // [pcap+1]
    int32_t* claim;         // [n] (global claims when n > CLAIM_LDS)
    uint32_t* keys;         // [n] x2 radix buffers
    uint32_t* vals;
    uint32_t* keys2;
    uint32_t* vals2;
    float* skey;            // [n] sorted keys as floats
    uint32_t* cand;         // [n+1] candidate counts -> offsets
    uint32_t* chunk;        // [chunk capacity] pairs per chunk -> offsets
    uint32_t* chunk_hit;    // [chunk capacity] per chunk, bit k: its candidate k is a pair
    uint32_t* scan_part;    // [scan_tiles] tile totals of the grid-wide scans
    uint32_t* hist;         // radix histograms
    double* red;            // [6 * nred] axis reduction partials
    DevCounts* cnt;
    int pcap;
    int chunk_cap;
    float dt;
    float g[3];
    int iterations;
};

// ------------------------------------------------------------------ stage 0: integrate
// World::updateObjectStatus (world.cpp:35-119) with zero sleep thresholds, then
// World::applyExternalForce (world.cpp:121-127), then SequentialImpulseSolver::setupSolver's
// clearTempVelocity (sequential_impulse_solver.cpp:17-20).
__global__ void k_integrate(Dev d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) integrate_body(d.b, i, d.dt, d.g);
}

// Recomputes AABBs and scales only (after pe_world_set_state).
__global__ void k_aabb_only(Dev d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) refresh_body(d.b, i);
}

// ------------------------------------------------------------------ stage 1: broadphase
__global__ void k_bp_keys(Dev d) {
    const int ax = d.cnt->axis;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) {
        d.keys[i] = float_key(d.b.at(F_AMIN + ax, i) + 0.0f);
        d.vals[i] = (uint32_t)i;
    }
}

// After the sort: sorted float keys and, per sorted position s, the number of candidates
// j > s with key_j <= max_i (the reference's sweep breaks at the first key_j > max_i).
__global__ void k_bp_ranges(Dev d) {
    const int n = d.b.n;
    const int ax = d.cnt->axis;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int i = (int)d.vals[s];
        d.skey[s] = d.b.at(F_AMIN + ax, i) + 0.0f;
    }
}
__global__ void k_bp_cand(Dev d) {
    const int n = d.b.n;
    const int ax = d.cnt->axis;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int i = (int)d.vals[s];
        const float mx = d.b.at(F_AMAX + ax, i);
        int lo = s + 1, hi = n;  // first position in [s+1, n) with skey > mx
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (d.skey[mid] > mx) hi = mid; else lo = mid + 1;
        }
        d.cand[s] = (uint32_t)(lo - s - 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) d.cand[n] = 0;
}

// Locates candidate t: sorted position s with cand[s] <= t < cand[s+1] (cand holds offsets).
__device__ __forceinline__ int bp_find(const uint32_t* off, int n, uint32_t t) {
    int lo = 0, hi = n;  // last s with off[s] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool EMIT>
__global__ void k_bp_chunks(Dev d) {
    const int n = d.b.n;
    const uint32_t S = d.cnt->ncand;
    const uint32_t nchunks = min((S + CHUNK - 1) / CHUNK, (uint32_t)d.chunk_cap);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        uint32_t t = c * CHUNK;
        const uint32_t tend = min(S, t + CHUNK);
        int s = bp_find(d.cand, n, t);
        while (d.cand[s + 1] <= t) s++;
        if (EMIT) {
            // the counting pass left the outcome of each of the chunk's tests in chunk_hit[c]
            uint32_t out = d.chunk[c];
            uint32_t hit = d.chunk_hit[c];
            const uint32_t t0 = t;
            int cur_s = s;
            while (hit) {
                const uint32_t tt = t0 + (uint32_t)(__ffs(hit) - 1);
                hit &= hit - 1;
                while (d.cand[cur_s + 1] <= tt) cur_s++;
                if ((int)out < d.pcap) {
                    d.ps.pa[out] = (int)d.vals[cur_s];
                    d.ps.pb[out] = (int)d.vals[cur_s + 1 + (int)(tt - d.cand[cur_s])];
                }
                out++;
            }
        } else {
            uint32_t cntv = 0, hit = 0;
            const uint32_t t0 = t;
            int cur_s = s, ci = (int)d.vals[s];
            float A[8], B[8];
            ld_rec<2>(d.b.abr(ci), A);   // the run body's record stays in registers
            for (; t < tend; t++) {
                while (d.cand[cur_s + 1] <= t) {
                    cur_s++;
                    ci = (int)d.vals[cur_s];
                    ld_rec<2>(d.b.abr(ci), A);
                }
                const int j = (int)d.vals[cur_s + 1 + (int)(t - d.cand[cur_s])];
                ld_rec<2>(d.b.abr(j), B);
                if (bp_valid_rec(A, B)) {
                    cntv++;
                    hit |= 1u << (t - t0);
                }
            }
            d.chunk[c] = cntv;
            d.chunk_hit[c] = hit;
        }
    }
}

// Number of live sweep chunks, for the chunk scan.
__global__ void k_bp_nchunks(Dev d, uint32_t* nchunks) {
    const uint32_t want = (d.cnt->ncand + CHUNK - 1) / CHUNK;
    *nchunks = min(want, (uint32_t)d.chunk_cap);
    if (want > (uint32_t)d.chunk_cap) d.cnt->overflow |= OVF_CHUNKS;
}

__global__ void k_bp_finish(Dev d) {
    const uint32_t total = d.cnt->nchunk_pairs;
    if (total > (uint32_t)d.pcap) d.cnt->overflow |= OVF_PAIRS;
    d.cnt->npairs = total > (uint32_t)d.pcap ? d.pcap : (int32_t)total;
}

// Next SAP axis (broad_phase_sweep_and_prune.cpp:37-40): variance of AABB centres, reduced in
// FP64 over fixed blocks (deterministic), argmax with the reference's tie rule.
constexpr int RED_THREADS = 256;
__global__ void __launch_bounds__(RED_THREADS) k_bp_axis_partial(Dev d) {
    __shared__ double sh[6][RED_THREADS];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float c = (d.b.at(F_AMIN + k, i) + d.b.at(F_AMAX + k, i)) * 0.5f;
            acc[k] += (double)c;
            acc[3 + k] += (double)(c * c);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < 6; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 6; k++) d.red[k * gridDim.x + blockIdx.x] = sh[k][0];
}
__global__ void k_bp_axis_final(Dev d, int nblocks) {
    if (threadIdx.x != 0) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblocks; b++)
        for (int k = 0; k < 6; k++) s[k] += d.red[k * nblocks + b];
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = s[3 + k] - s[k] * s[k] / (double)d.b.n;
    int axis = 0;
    if (v[1] > v[0]) axis = 1;
    if (v[2] > v[axis]) axis = 2;
    d.cnt->axis = axis;
}

// ------------------------------------------------------------------ stage 2: narrowphase
// Box and sphere pairs here; pairs with a cylinder or a mesh are listed for k_narrow_shapes.
__global__ void __launch_bounds__(128) k_narrow(Dev d) {
    const int np = d.cnt->npairs;
    const int prev = d.cnt->prev_npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const bool simple = pair_is_simple(d.b, d.ps, p);
        const bool sat = !simple && pair_is_sat(d.b, d.ps, p);
        const int slot = wave_append(!simple && !sat, &d.cnt->nshape);
        const int sslot = wave_append(sat, &d.cnt->nsat);
        if (simple) narrow_pair<false>(d.b, d.ps, p, prev);
        else if (sat) d.shape_list[d.pcap - 1 - sslot] = p;
        else d.shape_list[slot] = p;
    }
}
__global__ void __launch_bounds__(64) k_narrow_shapes(Dev d) {
    const int ns = d.cnt->nshape;
    const int prev = d.cnt->prev_npairs;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ns; q += gridDim.x * blockDim.x)
        narrow_pair<true>(d.b, d.ps, d.shape_list[q], prev);
}
// One wave per listed separating-axis pair.
__global__ void __launch_bounds__(256) k_narrow_sat(Dev d) {
    const int ns = d.cnt->nsat;
    const int prev = d.cnt->prev_npairs;
    const int nw = gridDim.x * (blockDim.x / WAVE);
    for (int q = blockIdx.x * (blockDim.x / WAVE) + (int)(threadIdx.x / WAVE); q < ns; q += nw)
        narrow_pair<true, WAVE>(d.b, d.ps, d.shape_list[d.pcap - 1 - q], prev);
}
__global__ void k_narrow_finish(Dev d) {
    d.cnt->prev_npairs = d.cnt->npairs;
    d.cnt->nshape = 0;
    d.cnt->nsat = 0;
    d.cnt->npoints = 0;   // summed by k_count_points next
}

// ------------------------------------------------------------------ stage 3: solver
// SequentialImpulseSolver::setupSolver (sequential_impulse_solver.cpp:11-58), thread / pair.
__global__ void k_setup(Dev d) {
    const int np = d.cnt->npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) setup_pair(d.b, d.ps, p, d.dt);
}

// Constraint-graph colouring of the pairs with contact points, one workgroup (colour_pairs in
// pe_pipeline.hpp): writes the solve order, colour starts and colour count.
template <bool LDS_CLAIM>
__global__ void __launch_bounds__(SOLVE_THREADS) k_colour(Dev d) {
    extern __shared__ int s_claim[];
    __shared__ int wsum[SOLVE_THREADS / WAVE + 1];
    const BodyView& b = d.b;
    const int np = d.cnt->npairs;
    int* claim = LDS_CLAIM ? s_claim : d.claim;
    for (int i = threadIdx.x; i < b.n; i += blockDim.x) claim[i] = colour::FREE;
    int* act = d.act;
    int nact = 0;
    for (int base = 0; base < np; base += blockDim.x) {
        const int p = base + threadIdx.x;
        int flag = 0;
        if (p < np && d.ps.npts[p] > 0) flag = !(b.iat(I_KIN, d.ps.pa[p]) && b.iat(I_KIN, d.ps.pb[p]));
        int tot;
        const int ex = block_exclusive_scan(flag, wsum, &tot);
        if (flag) act[nact + ex] = p;
        nact += tot;
    }
    __syncthreads();
    int* bufs[3] = {act, act + d.pcap, act + 2 * (size_t)d.pcap};
    const int ncol = colour_pairs<LDS_CLAIM>(b, d.ps, bufs, nact, d.order, d.color_start, claim, 0, 0, wsum);
    if (threadIdx.x == 0) {
        d.cnt->ncolors = ncol;
        d.cnt->nordered = d.color_start[ncol];
    }
}

// Gauss-Seidel sweeps (SequentialImpulseSolver::solve, sequential_impulse_solver.cpp:60-81) in
// colour order across the whole GPU: the pairs of one colour share no dynamic body, so they run
// in parallel; a grid barrier separates colours.  Then syncTempVelocity.  Launched cooperatively
// (all workgroups co-resident).
constexpr int GS_THREADS = 256;

// Sense-reversing barrier over all workgroups of a cooperative launch (every workgroup is
// resident, so every wait ends).  The last arrival resets the count and bumps the generation.
__device__ __forceinline__ void grid_barrier(uint32_t* count, uint32_t* gen, uint32_t nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = __hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == nblocks - 1) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(GS_THREADS) k_gs_grid(Dev d) {
    const int ncol = d.cnt->ncolors;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    for (int it = 0; it < d.iterations; it++) {
        for (int c = 0; c < ncol; c++) {
            const int c0 = d.color_start[c], c1 = d.color_start[c + 1];
            for (int q = c0 + gtid; q < c1; q += gsz) iterate_pair(d.b, d.ps, d.order[q]);
            __threadfence();   // release this colour's velocity updates
            grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);
            __threadfence();   // acquire: later colours read velocities other CUs wrote
        }
    }
    for (int i = gtid; i < d.b.n; i += gsz) sync_body(d.b, i);
}

// Contact points of this narrowphase: per-workgroup sums added to the counter k_narrow_finish
// cleared; a pair whose clipper lost points (bit 1 of its swap word) raises OVF_CLIP.
__global__ void __launch_bounds__(256) k_count_points(Dev d) {
    __shared__ int wsum[256 / WAVE + 1];
    const int np = d.cnt->npairs;
    int acc = 0, lost = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        acc += d.ps.npts[p];
        lost |= d.ps.swap[p] & 2;
    }
    int tot;
    block_exclusive_scan(acc, wsum, &tot);
    if (threadIdx.x == 0 && tot) atomicAdd(&d.cnt->npoints, tot);
    if (lost) atomicOr(&d.cnt->overflow, OVF_CLIP);
}

}  // namespace

// ====================================================================== host side
// Workgroups of the cooperative Gauss-Seidel sweep: enough that a colour's pairs run about one
// per thread, few enough that the grid barrier between colours stays cheap.  Measured at rest
// (ms per step, 4 / 8 / 16 / 32 / 64 workgroups of 256): 10k box rain 3.01 / 2.52 / 2.47 / 2.71 /
// 3.62; 50k shape mix 9.27 / 8.31 / 5.99 / 6.49 / 7.36.
static int default_gs_blocks(int n, int max_blocks) { return std::min(max_blocks, std::max(4, std::min(16, n / 600))); }

struct pe_world {
    Dev d{};
    int n = 0, stride = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    int rs_blocks = 0;
    int red_blocks = 0;
    int gs_blocks = 0;      // workgroups of the cooperative Gauss-Seidel kernel
    int gs_max_blocks = 0;  // co-resident limit
    bool timing = false;
    hipEvent_t ev[5] = {};
    double stage_ms[4] = {0, 0, 0, 0};
    int n_dynamic = 0;
    bool has_shapes = false;   // any cylinder / convex / concave body
    DevMeshes meshes;

    ~pe_world() {
        if (stream) hipStreamSynchronize(stream);
        for (void* p : allocs) hipFree(p);
        meshes.release();
        for (auto& e : ev) if (e) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
    template <class T> int alloc(T** p, size_t count) {
        void* q = nullptr;
        PE_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        PE_HIP(hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), stream));
        allocs.push_back(q);
        *p = (T*)q;
        return PE_OK;
    }
    int grid(int work, int block) const { return std::max(1, std::min((work + block - 1) / block, 4096)); }

    int launch_integrate() {
        hipLaunchKernelGGL(k_integrate, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        return PE_OK;
    }
    // Grid-wide exclusive scan of n elements (of *count when count is given; at most cap).
    void scan(uint32_t* data, int cap, int n, const uint32_t* count, uint32_t* total) {
        const int nb = std::max(1, (cap + SCAN_TILE - 1) / SCAN_TILE);
        hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(1024), 0, stream, data, n, count, d.scan_part);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, stream, d.scan_part, nb, total);
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, stream, data, n, count, d.scan_part);
    }
    int launch_broadphase() {
        hipLaunchKernelGGL(k_bp_keys, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        uint32_t *ki = d.keys, *vi = d.vals, *ko = d.keys2, *vo = d.vals2;
        for (int pass = 0; pass < 4; pass++) {
            hipLaunchKernelGGL(k_radix_hist, dim3(rs_blocks), dim3(RS_THREADS), 0, stream, ki, n, pass * 8, d.hist, rs_blocks);
            hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, stream, d.hist, 256 * rs_blocks, (uint32_t*)nullptr);
            hipLaunchKernelGGL(k_radix_scatter, dim3(rs_blocks), dim3(RS_THREADS), 0, stream, ki, vi, ko, vo, n, pass * 8,
                               d.hist, rs_blocks);
            std::swap(ki, ko);
            std::swap(vi, vo);
        }
        // after 4 passes the sorted data is back in keys/vals
        hipLaunchKernelGGL(k_bp_ranges, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        hipLaunchKernelGGL(k_bp_cand, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        scan(d.cand, n + 1, n + 1, nullptr, &d.cnt->ncand);
        hipLaunchKernelGGL(k_bp_nchunks, dim3(1), dim3(1), 0, stream, d, &d.cnt->nchunks);
        hipLaunchKernelGGL(k_bp_chunks<false>, dim3(2048), dim3(256), 0, stream, d);
        scan(d.chunk, d.chunk_cap, 0, &d.cnt->nchunks, &d.cnt->nchunk_pairs);
        hipLaunchKernelGGL(k_bp_chunks<true>, dim3(2048), dim3(256), 0, stream, d);
        hipLaunchKernelGGL(k_bp_finish, dim3(1), dim3(1), 0, stream, d);
        hipLaunchKernelGGL(k_bp_axis_partial, dim3(red_blocks), dim3(RED_THREADS), 0, stream, d);
        hipLaunchKernelGGL(k_bp_axis_final, dim3(1), dim3(64), 0, stream, d, red_blocks);
        return PE_OK;
    }
    int launch_narrowphase() {
        hipLaunchKernelGGL(k_narrow, dim3(grid(d.pcap, 128)), dim3(128), 0, stream, d);
        if (has_shapes) {
            hipLaunchKernelGGL(k_narrow_shapes, dim3(grid(d.pcap, 64)), dim3(64), 0, stream, d);
            hipLaunchKernelGGL(k_narrow_sat, dim3(std::min(grid(d.pcap, 4), 4096)), dim3(256), 0, stream, d);
        }
        hipLaunchKernelGGL(k_narrow_finish, dim3(1), dim3(1), 0, stream, d);
        hipLaunchKernelGGL(k_count_points, dim3(grid(d.pcap, 1024)), dim3(256), 0, stream, d);
        return PE_OK;
    }
    int launch_solve() {
        hipLaunchKernelGGL(k_setup, dim3(grid(d.pcap, 128)), dim3(128), 0, stream, d);
        if (n <= CLAIM_LDS)
            hipLaunchKernelGGL(k_colour<true>, dim3(1), dim3(SOLVE_THREADS), n * sizeof(int), stream, d);
        else
            hipLaunchKernelGGL(k_colour<false>, dim3(1), dim3(SOLVE_THREADS), 0, stream, d);
        void* args[] = {(void*)&d};
        PE_HIP(hipLaunchCooperativeKernel((const void*)k_gs_grid, dim3(gs_blocks), dim3(GS_THREADS), args, 0, stream));
        return PE_OK;
    }
    int step() {
        if (timing) hipEventRecord(ev[0], stream);
        launch_integrate();
        if (timing) hipEventRecord(ev[1], stream);
        launch_broadphase();
        if (timing) hipEventRecord(ev[2], stream);
        launch_narrowphase();
        if (timing) hipEventRecord(ev[3], stream);
        launch_solve();
        if (timing) hipEventRecord(ev[4], stream);
        PE_HIP(hipGetLastError());
        if (timing) {
            PE_HIP(hipEventSynchronize(ev[4]));
            for (int k = 0; k < 4; k++) {
                float ms = 0;
                hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
                stage_ms[k] += ms;
            }
        }
        return PE_OK;
    }
};

namespace {
int check_world(pe_world* w) {
    if (!w) { pe_set_error("null world"); return PE_ERR_INVALID; }
    return PE_OK;
}
// Reports the capacity errors the steps since the last report raised (PE_ERR_CAPACITY, with the
// causes in the message) and clears them; c holds counters read after the stream drained.
int report_overflow(pe_world* w, const DevCounts& c) {
    if (c.overflow == 0) return PE_OK;
    std::string m = "capacity exceeded:";
    if (c.overflow & OVF_PAIRS) m += " more broadphase pairs than max_pairs (" + std::to_string(w->d.pcap) + ")";
    if (c.overflow & OVF_CHUNKS) m += " more sweep candidates than the sweep's chunk capacity";
    if (c.overflow & OVF_CLIP) m += " a clipped contact polygon outgrew its buffer";
    PE_HIP(hipMemsetAsync(&w->d.cnt->overflow, 0, sizeof(int32_t), w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    pe_set_error(m);
    return PE_ERR_CAPACITY;
}
}  // namespace

extern "C" {

const char* pe_last_error(void) { return g_err.c_str(); }

int pe_device_count(int32_t* count) {
    int c = 0;
    PE_HIP(hipGetDeviceCount(&c));
    if (count) *count = c;
    return PE_OK;
}

int pe_world_create(const pe_scene_desc* sc, const pe_world_params* prm, int32_t max_pairs, pe_world** out) {
    if (!sc || !prm || !out || sc->n <= 0) { pe_set_error("pe_world_create: invalid argument"); return PE_ERR_INVALID; }
    if (int rc = check_scene(sc)) return rc;
    auto* w = new pe_world();
    const int n = sc->n;
    w->n = n;
    w->stride = (n + 63) & ~63;
    const int pcap = max_pairs > 0 ? max_pairs : std::max(1024, 16 * n);
    auto fail = [&](int rc) { delete w; return rc; };
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) { pe_set_error("stream create failed"); return fail(PE_ERR_DEVICE); }
    Dev& d = w->d;
    d.pcap = pcap;
    d.ps.cs = (size_t)pcap * 8;
    d.dt = (float)prm->dt;
    for (int k = 0; k < 3; k++) d.g[k] = (float)prm->gravity[k];
    d.iterations = prm->iterations;
    d.b.n = n;
    d.b.stride = w->stride;
    const size_t S = w->stride;
    int rc = PE_OK;
    rc |= w->alloc(&d.b.f, NF * S);
    rc |= w->alloc(&d.b.i, NI * S);
    rc |= w->alloc(&d.ps.pa, pcap);
    rc |= w->alloc(&d.ps.pb, pcap);
    rc |= w->alloc(&d.ps.cp, (size_t)CW * pcap * 8);
    rc |= w->alloc(&d.ps.cd, (size_t)pcap * 8);
    rc |= w->alloc(&d.b.cb, (size_t)CBW * S);
    rc |= w->alloc(&d.b.sb, (size_t)SBW * S);
    rc |= w->alloc(&d.b.ab, (size_t)8 * S);
    rc |= w->alloc(&d.ps.npts, pcap);
    rc |= w->alloc(&d.ps.swap, pcap);
    rc |= w->alloc(&d.ps.rows, (size_t)RW * pcap * 8);
    rc |= w->alloc(&d.order, pcap);
    rc |= w->alloc(&d.shape_list, pcap);
    rc |= w->alloc(&d.act, 3 * (size_t)pcap);
    rc |= w->alloc(&d.color_start, pcap + 1);
    rc |= w->alloc(&d.claim, n);
    rc |= w->alloc(&d.keys, n);
    rc |= w->alloc(&d.vals, n);
    rc |= w->alloc(&d.keys2, n);
    rc |= w->alloc(&d.vals2, n);
    rc |= w->alloc(&d.skey, n);
    rc |= w->alloc(&d.cand, n + 1);
    w->rs_blocks = (n + RS_TILE - 1) / RS_TILE;
    rc |= w->alloc(&d.hist, 256 * (size_t)w->rs_blocks);
    w->red_blocks = std::max(1, std::min(256, (n + RED_THREADS - 1) / RED_THREADS));
    rc |= w->alloc(&d.red, 6 * (size_t)w->red_blocks);
    rc |= w->alloc(&d.cnt, 1);
    // sweep chunks: candidates are bounded by n * (n - 1) / 2; size for pcap * 64 candidates
    d.chunk_cap = (int)std::min<int64_t>((int64_t)pcap * 64 / CHUNK + 1, (int64_t)1 << 26);
    rc |= w->alloc(&d.chunk, d.chunk_cap);
    rc |= w->alloc(&d.chunk_hit, d.chunk_cap);
    rc |= w->alloc(&d.scan_part, (std::max(d.chunk_cap, n + 1) + SCAN_TILE - 1) / SCAN_TILE + 1);
    if (rc != PE_OK) return fail(PE_ERR_DEVICE);

    // host staging of the body store (FP32) and of the scene's meshes
    std::vector<float> hf(NF * S, 0.0f);
    std::vector<int32_t> hi(NI * S, 0);
    HostMeshes hm;
    if (int rc2 = stage_bodies(sc, hf, hi, S, 0, hm)) return fail(rc2);
    if (int rc2 = w->meshes.upload(hm, w->stream)) return fail(rc2);
    d.b.ms = w->meshes.view();
    int ndyn = 0;
    for (int i = 0; i < n; i++) {
        ndyn += sc->kinematic[i] ? 0 : 1;
        w->has_shapes |= sc->shape_type[i] > PE_SHAPE_SPHERE;
    }
    w->n_dynamic = ndyn;
    if (hipMemcpyAsync(d.b.f, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice, w->stream) != hipSuccess ||
        hipMemcpyAsync(d.b.i, hi.data(), hi.size() * sizeof(int32_t), hipMemcpyHostToDevice, w->stream) != hipSuccess) {
        pe_set_error("upload failed");
        return fail(PE_ERR_DEVICE);
    }
    hipLaunchKernelGGL(k_aabb_only, dim3(w->grid(n, 256)), dim3(256), 0, w->stream, d);
    for (auto& e : w->ev) hipEventCreate(&e);
    {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gs_grid, GS_THREADS, 0) != hipSuccess || per_cu < 1) {
            pe_set_error("cannot size the cooperative solver grid");
            return fail(PE_ERR_DEVICE);
        }
        w->gs_max_blocks = cus * per_cu;
        w->gs_blocks = default_gs_blocks(n, w->gs_max_blocks);
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) { pe_set_error("init sync failed"); return fail(PE_ERR_DEVICE); }
    *out = w;
    return PE_OK;
}

int pe_world_destroy(pe_world* w) {
    delete w;
    return PE_OK;
}

int pe_world_step(pe_world* w, int32_t nsteps) {
    if (int rc = check_world(w)) return rc;
    for (int s = 0; s < nsteps; s++)
        if (int rc = w->step()) return rc;
    return PE_OK;
}
int pe_world_integrate(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    w->launch_integrate();
    PE_HIP(hipGetLastError());
    return PE_OK;
}
int pe_world_broadphase(pe_world* w, int32_t* npairs) {
    if (int rc = check_world(w)) return rc;
    w->launch_broadphase();
    PE_HIP(hipGetLastError());
    if (npairs) {
        DevCounts c;
        PE_HIP(hipMemcpyAsync(&c, w->d.cnt, sizeof(c), hipMemcpyDeviceToHost, w->stream));
        PE_HIP(hipStreamSynchronize(w->stream));
        *npairs = c.npairs;
        if (int rc = report_overflow(w, c)) return rc;
    }
    return PE_OK;
}
int pe_world_narrowphase(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    w->launch_narrowphase();
    PE_HIP(hipGetLastError());
    return PE_OK;
}
int pe_world_solve(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    w->launch_solve();
    PE_HIP(hipGetLastError());
    return PE_OK;
}
int pe_world_sync(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    PE_HIP(hipStreamSynchronize(w->stream));
    DevCounts c;
    PE_HIP(hipMemcpy(&c, w->d.cnt, sizeof(c), hipMemcpyDeviceToHost));
    return report_overflow(w, c);
}

static int get_field(pe_world* w, int f0, int nf, float* out) {
    const size_t S = w->stride;
    std::vector<float> tmp((size_t)nf * S);
    PE_HIP(hipMemcpyAsync(tmp.data(), w->d.b.f + (size_t)f0 * S, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    for (int i = 0; i < w->n; i++)
        for (int k = 0; k < nf; k++) out[(size_t)i * nf + k] = tmp[(size_t)k * S + i];
    return PE_OK;
}
static int set_field(pe_world* w, int f0, int nf, const float* in) {
    const size_t S = w->stride;
    std::vector<float> tmp((size_t)nf * S);
    PE_HIP(hipMemcpyAsync(tmp.data(), w->d.b.f + (size_t)f0 * S, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    for (int i = 0; i < w->n; i++)
        for (int k = 0; k < nf; k++) tmp[(size_t)k * S + i] = in[(size_t)i * nf + k];
    PE_HIP(hipMemcpyAsync(w->d.b.f + (size_t)f0 * S, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    return PE_OK;
}

int pe_world_get_state(pe_world* w, float* pos, float* rot, float* v, float* av) {
    if (int rc = check_world(w)) return rc;
    int rc = PE_OK;
    if (pos && (rc = get_field(w, F_POS, 3, pos))) return rc;
    if (rot && (rc = get_field(w, F_ROT, 9, rot))) return rc;
    if (v && (rc = get_field(w, F_V, 3, v))) return rc;
    if (av && (rc = get_field(w, F_W, 3, av))) return rc;
    return PE_OK;
}
int pe_world_set_state(pe_world* w, const float* pos, const float* rot, const float* v, const float* av) {
    if (int rc = check_world(w)) return rc;
    int rc = PE_OK;
    if (pos && (rc = set_field(w, F_POS, 3, pos))) return rc;
    if (rot && (rc = set_field(w, F_ROT, 9, rot))) return rc;
    if (v && (rc = set_field(w, F_V, 3, v))) return rc;
    if (av && (rc = set_field(w, F_W, 3, av))) return rc;
    hipLaunchKernelGGL(k_aabb_only, dim3(w->grid(w->n, 256)), dim3(256), 0, w->stream, w->d);
    PE_HIP(hipStreamSynchronize(w->stream));
    return PE_OK;
}
int pe_world_get_aabbs(pe_world* w, float* mn, float* mx) {
    if (int rc = check_world(w)) return rc;
    int rc = PE_OK;
    if (mn && (rc = get_field(w, F_AMIN, 3, mn))) return rc;
    if (mx && (rc = get_field(w, F_AMAX, 3, mx))) return rc;
    return PE_OK;
}
static int read_counts(pe_world* w, DevCounts& c) {
    PE_HIP(hipMemcpyAsync(&c, w->d.cnt, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    return PE_OK;
}
int pe_world_get_pairs(pe_world* w, int32_t* pairs, int32_t cap, int32_t* npairs) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    if (npairs) *npairs = c.npairs;
    const int m = std::min(cap, c.npairs);
    if (pairs && m > 0) {
        std::vector<int32_t> a(m), b(m);
        PE_HIP(hipMemcpy(a.data(), w->d.ps.pa, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        PE_HIP(hipMemcpy(b.data(), w->d.ps.pb, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int i = 0; i < m; i++) { pairs[2 * i] = a[i]; pairs[2 * i + 1] = b[i]; }
    }
    return PE_OK;
}
int pe_world_get_contacts(pe_world* w, int32_t* npoints, float* points, int32_t cap) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    const int m = std::min(cap, c.npairs);
    if (m <= 0) return PE_OK;
    std::vector<int32_t> np(m);
    PE_HIP(hipMemcpy(np.data(), w->d.ps.npts, m * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (npoints) std::copy(np.begin(), np.end(), npoints);
    if (points) {
        // slots at or beyond the point count are reported invalid; every slot's distance is reported
        std::vector<float> rec((size_t)m * 8 * CW), dist((size_t)m * 8);
        PE_HIP(hipMemcpy(rec.data(), w->d.ps.cp, rec.size() * sizeof(float), hipMemcpyDeviceToHost));
        PE_HIP(hipMemcpy(dist.data(), w->d.ps.cd, dist.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t s = 0; s < (size_t)m * 8; s++) {
            float* o = points + s * CF;
            const bool valid = (int)(s % 8) < np[s / 8];
            for (int f = 0; f < CW; f++) o[f] = valid ? rec[s * CW + f] : (f < 3 ? REAL_MAX : 0.0f);
            o[12] = dist[s];
        }
    }
    return PE_OK;
}
int pe_world_get_impulses(pe_world* w, float* impulses, int32_t cap) {
    if (int rc = check_world(w)) return rc;
    if (!impulses) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    const int m = std::min(cap, c.npairs);
    if (m <= 0) return PE_OK;
    std::vector<int32_t> np(m);
    std::vector<float> rows((size_t)m * 8 * RW);
    PE_HIP(hipMemcpy(np.data(), w->d.ps.npts, m * sizeof(int32_t), hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(rows.data(), w->d.ps.rows, rows.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t s = 0; s < (size_t)m * 8; s++) {
        const bool valid = (int)(s % 8) < np[s / 8];
        for (int k = 0; k < 3; k++) impulses[s * 3 + k] = valid ? rows[s * RW + R_LN + k] : 0.0f;
    }
    return PE_OK;
}
int pe_world_get_solve_order(pe_world* w, int32_t* order, int32_t cap, int32_t* nordered, int32_t* color_start,
                             int32_t ccap, int32_t* ncolors) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    if (nordered) *nordered = c.nordered;
    if (ncolors) *ncolors = c.ncolors;
    if (order && cap > 0 && c.nordered > 0)
        PE_HIP(hipMemcpy(order, w->d.order, std::min(cap, c.nordered) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (color_start && ccap > 0)
        PE_HIP(hipMemcpy(color_start, w->d.color_start, std::min(ccap, c.ncolors + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PE_OK;
}
int pe_world_get_axis(pe_world* w, int32_t* axis) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    if (axis) *axis = c.axis;
    return PE_OK;
}
int pe_world_set_axis(pe_world* w, int32_t axis) {
    if (int rc = check_world(w)) return rc;
    if (axis < 0 || axis > 2) { pe_set_error("axis out of range"); return PE_ERR_INVALID; }
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    c.axis = axis;
    PE_HIP(hipMemcpy(w->d.cnt, &c, sizeof(c), hipMemcpyHostToDevice));
    return PE_OK;
}
int pe_world_set_timing(pe_world* w, int32_t enable) {
    if (int rc = check_world(w)) return rc;
    w->timing = enable != 0;
    return PE_OK;
}
int pe_world_get_stage_times(pe_world* w, double* ms) {
    if (int rc = check_world(w)) return rc;
    for (int k = 0; k < 4; k++) ms[k] = w->stage_ms[k];
    return PE_OK;
}
int pe_world_set_solver_blocks(pe_world* w, int32_t blocks) {
    if (int rc = check_world(w)) return rc;
    if (blocks < 0) { pe_set_error("blocks must be >= 0"); return PE_ERR_INVALID; }
    w->gs_blocks = blocks == 0 ? default_gs_blocks(w->n, w->gs_max_blocks) : std::min(blocks, w->gs_max_blocks);
    return PE_OK;
}

int pe_world_get_counts(pe_world* w, int64_t* counts) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    counts[0] = w->n;
    counts[1] = w->n_dynamic;
    counts[2] = c.npairs;
    counts[3] = c.npoints;
    counts[4] = c.ncolors;
    return PE_OK;
}

}  // extern "C"
