// Batched and served stepping: many independent worlds per launch.
//
// One step of every world is three launches:
//   K1 k_world_broad<BLOCK>  one workgroup per world: integrate (thread / body), SAP keys,
//                            bitonic sort of (key, index) in LDS, candidate runs, load-balanced
//                            sweep in (i, j) order, next axis; the world's pairs are appended to
//                            a narrowphase work list
//   K2 k_narrow_list         one thread per pair of every world: narrowphase + contact cache
//   K3 k_group_solve<BLOCK>  one workgroup per group of worlds of a size class: solver rows and
//                            colouring (LDS lists), then colour-ordered Gauss-Seidel sweeps on LDS
//                            copies of the group's bodies and pairs, velocity sync
// Every routine is the one the single-world pipeline uses (pe_pipeline.hpp), and the sort,
// sweep order and colouring rule are the same, so a world stepped here is bit-identical to
// the same world stepped alone -- batching never changes a result.
//
// Worlds are grouped into size classes; K1 and K3 of each class launch with the block size an
// autotuner picked for it and LDS sized for the largest world in the class.  The launches of a
// step are captured once into a hipGraph and replayed.
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_prims.hpp"
#include "pe_pipeline.hpp"
#include "pe_host.hpp"
#include "../../include/pe_amd.h"
#include "../../include/pe_amd_batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <numeric>
#include <vector>

using namespace pe;

namespace {

struct WorldDesc {      // static per world
    int32_t body_off;
    int32_t n;
    int32_t pair_off;
    int32_t pair_cap;
    int32_t cs_off;     // colour starts of the global-memory solve path: pair_cap + 1 slots of d.cstart
    int32_t pad[3];
};
struct WorldState {     // per world, updated by every step
    int32_t npairs;
    int32_t prev_npairs;
    int32_t axis;
    int32_t ncolors;
    int32_t npoints;
    int32_t overflow;   // sticky capacity-error bits (WOVF_*) since the world was created / admitted
    int32_t pad[2];
};
// Capacity errors of a world: its pairs beyond its pair capacity were dropped, or a clipped
// contact polygon outgrew its buffer.  Each is also OR-ed into BatchDev::ovf_any, which the host
// reads at every synchronisation (one word) and reports as PE_ERR_CAPACITY.
constexpr int WOVF_PAIRS = 1;
constexpr int WOVF_CLIP = 8;

struct BatchDev {
    BodyView b;
    PairStore ps;
    const WorldDesc* desc;
    WorldState* st;
    int32_t* act;       // [3 * total pair cap]
    int32_t* order;     // [total pair cap]
    int32_t* cstart;    // [total pair cap + nworlds]
    float dt;
    float g[3];
    int iterations;
    uint64_t* phase_ts;  // optional [nworlds][PHASES] timestamps (null: off)
    // Narrowphase work lists, one segment per size class (NarrowSeg): the global pair slots of the
    // class's worlds, appended by its K1 and worked by its K2.
    int32_t* nl_pair;    // [total pair cap] listed pair slots
    int32_t* nl_limit;   // [total pair cap] per listed pair: first slot of its world that starts fresh
    int32_t* nl_shape;   // [total pair cap] per segment, list entries whose pair involves a cylinder or a
                         // mesh: from the front those of k_narrow_shape_list, from the back those of
                         // k_narrow_sat_list (pair_on_wave)
    int32_t* nl_count;   // [3 * MAX_CLASSES] per class: listed pairs, shape entries, wave entries
    int32_t* ovf_any;    // [1] OR of the capacity-error bits raised since the host last read it
};
constexpr int MAX_CLASSES = 32;
// A class's narrowphase list segment: entries [off, off + cap) of the list arrays, counters
// nl_count[3 * cls .. 3 * cls + 2].
struct NarrowSeg {
    int off, cap, cls;
};
// Every class's segment, for the shape kernels, which run once over all classes.
struct NarrowSegs {
    int n;
    int off[MAX_CLASSES], cap[MAX_CLASSES];
};

constexpr uint64_t KEY_PAD = ~0ull;
constexpr int PHASES = 8;   // start, integrate, sort, sweep, narrow, setup, colour, solve

#define PE_PHASE(k)                                                                   \
    do {                                                                              \
        if (d.phase_ts) {                                                             \
            __syncthreads();                                                          \
            if (tid == 0) d.phase_ts[(size_t)w * PHASES + (k)] = wall_clock64();      \
        }                                                                             \
    } while (0)

// LDS layout for a class whose largest world has nmax bodies.
__host__ __device__ inline int np2(int n) { int p = 1; while (p < n) p <<= 1; return p; }
// K1's LDS: sort keys, sorted keys as floats, candidate counts, axis maxima, then (for worlds
// small enough: up to 64 KiB in all) the bodies' AABB records.
__host__ __device__ inline size_t broad_lds_head(int nmax) {
    const int P = np2(nmax > 2 ? nmax : 2);
    const size_t head = (size_t)P * 8 + (size_t)nmax * 4 /*skey*/ + (size_t)(nmax + 1) * 4 /*cand*/ + (size_t)nmax * 4 /*smax*/;
    return (head + 15) & ~(size_t)15;
}
__host__ __device__ inline bool broad_rec_in_lds(int nmax) { return broad_lds_head(nmax) + (size_t)nmax * 32 <= 64 * 1024; }
__host__ inline size_t broad_lds_bytes(int nmax) {
    return broad_lds_head(nmax) + (broad_rec_in_lds(nmax) ? (size_t)nmax * 32 : 0);
}
// LDS of K3 for a group of GB bodies and GP pairs.  The preparation uses per body its claim and
// inverse mass (8 B) and per pair its local bodies, global slot, claim priority, order slot,
// three colouring-list slots and first row (36 B); the sweeps then reuse the same memory for the body state
// (64 B per body) and per pair its local bodies, global slot and first row (12 B) plus the solve
// order and colour starts.  end covers both; a group is solved on the global path when it does
// not fit the launch's LDS (both halves evaluate this the same way).
struct PrepLayout {
    size_t bs, pab, gi, pr, ord, buf, cs, rs, end;
    __host__ __device__ PrepLayout(int GB, int GP) {
        auto al16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
        bs = al16((size_t)GB * 4);
        pab = al16(bs + (size_t)GB * 4);
        gi = al16(pab + (size_t)GP * 4);
        pr = al16(gi + (size_t)GP * 4);
        ord = al16(pr + (size_t)GP * 4);
        buf = al16(ord + (size_t)GP * 4);
        cs = al16(buf + (size_t)GP * 12);
        rs = al16(cs + (size_t)(GP + 2) * 4);
        const size_t prep_end = rs + (size_t)(GP + 1) * 4;
        const size_t sweep_end =
            al16((size_t)GB * 64 + (size_t)(3 * GP + 1) * 4) + (size_t)GP * 4 + (size_t)(GP + 1) * 4;
        end = prep_end > sweep_end ? prep_end : sweep_end;
    }
};
// K3's launch LDS for a size class: the layout of a group of G = max(group bodies, largest world)
// bodies and 2 pairs per body (a group with more pairs takes the global path), capped so a
// workgroup fits a CU.
constexpr size_t SOLVE_LDS_CAP = 144 * 1024;
__host__ inline size_t solve_lds_bytes(int nmax, int group_bodies) {
    const int g = std::max(std::max(nmax, group_bodies), 1);
    return std::min(SOLVE_LDS_CAP, PrepLayout(g, 2 * g).end);
}
// The AABB record of local body i: from LDS when the world's records are there.
__device__ __forceinline__ void lds_rec(const float4* lrec, const BodyView& b, int off, int i, float (&r)[8]) {
    if (!lrec) {
        ld_rec<2>(b.abr(off + i), r);
        return;
    }
    const float4 a = lrec[2 * i], c = lrec[2 * i + 1];
    r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
    r[4] = c.x; r[5] = c.y; r[6] = c.z; r[7] = c.w;
}
// K1: integrate + sweep-and-prune for one world per workgroup; appends the world's pairs to
// the narrowphase work list.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_world_broad(BatchDev d, const int32_t* __restrict__ wlist, int nmax, NarrowSeg g) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int wsum[BLOCK / WAVE + 1];
    __shared__ double red[6][BLOCK / WAVE];
    __shared__ int s_base;
    const int w = wlist[blockIdx.x];
    const WorldDesc wd = d.desc[w];
    const int n = wd.n, off = wd.body_off, poff = wd.pair_off;
    const BodyView& b = d.b;
    const int tid = threadIdx.x;
    const int P2 = np2(max(nmax, 2));
    uint64_t* key = (uint64_t*)lds;
    float* skey = (float*)(key + P2);
    uint32_t* cand = (uint32_t*)(skey + nmax);
    float* smax = (float*)(cand + nmax + 1);   // per body (unsorted): aabb_max[axis]
    // per body (unsorted): its AABB record, as integrate writes it to the body's record (null: the
    // sweep reads the records from the body store)
    float4* lrec = broad_rec_in_lds(nmax) ? (float4*)(lds + broad_lds_head(nmax)) : nullptr;
    const int axis = d.st[w].axis;
    const int prev_np = d.st[w].prev_npairs;

    PE_PHASE(0);
    // ---- stage 0: integrate (world.cpp:35-127)
    // with each body's sort key (aabb_min[axis] + 0, index) into LDS, from the AABB it just made
    for (int i = tid; i < n; i += BLOCK) {
        const Box3 a = integrate_body<false, true>(b, off + i, d.dt, d.g);
        const float m = (axis == 0 ? a.mn.x : (axis == 1 ? a.mn.y : a.mn.z)) + 0.0f;
        key[i] = ((uint64_t)float_key(m) << 32) | (uint32_t)i;
        smax[i] = axis == 0 ? a.mx.x : (axis == 1 ? a.mx.y : a.mx.z);
        if (lrec) {
            lrec[2 * i] = make_float4(a.mn.x, a.mn.y, a.mn.z, a.kin);
            lrec[2 * i + 1] = make_float4(a.mx.x, a.mx.y, a.mx.z, a.coll);
        }
    }
    for (int i = n + tid; i < P2; i += BLOCK) key[i] = KEY_PAD;
    __syncthreads();
    PE_PHASE(1);

    // ---- stage 1: sweep and prune (broad_phase_sweep_and_prune.cpp:6-41)
    for (int k = 2; k <= P2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P2; i += BLOCK) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = key[i], c = key[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) { key[i] = c; key[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int s = tid; s < n; s += BLOCK) skey[s] = key_float((uint32_t)(key[s] >> 32));
    __syncthreads();
    PE_PHASE(2);
    // candidate run of sorted position s: j in (s, first skey > max_i)
    int carry = 0;
    for (int base = 0; base < n; base += BLOCK) {
        const int s = base + tid;
        int c = 0;
        if (s < n) {
            const float mx = smax[(int)(uint32_t)key[s]];
            int lo = s + 1, hi = n;
            while (lo < hi) { const int mid = (lo + hi) >> 1; if (skey[mid] > mx) hi = mid; else lo = mid + 1; }
            c = lo - s - 1;
        }
        int tot;
        const int ex = block_exclusive_scan(c, wsum, &tot);
        if (s < n) cand[s] = (uint32_t)(carry + ex);
        carry += tot;
    }
    if (tid == 0) cand[n] = (uint32_t)carry;
    __syncthreads();
    const uint32_t S = cand[n];
    const uint32_t C = (S + BLOCK - 1) / BLOCK;
    const uint32_t t0 = min(S, (uint32_t)tid * C), t1 = min(S, t0 + C);
    int npairs;
    {
        uint32_t cnt = 0;
        int s0 = 0;
        if (t0 < t1) {
            int lo = 0, hi = n;
            while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (cand[mid] <= t0) lo = mid; else hi = mid; }
            s0 = lo;
        }
        // a thread's candidates come in runs of one body i: its AABB record stays in registers.
        // With at most 64 candidates a thread keeps the outcome of each test in a bit mask, and
        // the emitting pass reads only the sorted order, not the records again.
        const bool masked = t1 - t0 <= 64;
        uint64_t hit = 0;
        int s = s0, si = -1, i = 0;
        float A[8], Bj[8];
        for (uint32_t t = t0; t < t1; t++) {
            while (cand[s + 1] <= t) s++;
            if (s != si) {
                si = s;
                i = (int)(uint32_t)key[s];
                lds_rec(lrec, b, off, i, A);
            }
            const int j = (int)(uint32_t)key[s + 1 + (int)(t - cand[s])];
            lds_rec(lrec, b, off, j, Bj);
            if (bp_valid_rec(A, Bj)) {
                cnt++;
                if (masked) hit |= 1ull << (t - t0);
            }
        }
        int tot;
        int out = block_exclusive_scan((int)cnt, wsum, &tot);
        npairs = min(tot, wd.pair_cap);
        if (tid == 0 && tot > wd.pair_cap) {
            d.st[w].overflow |= WOVF_PAIRS;
            atomicOr(d.ovf_any, WOVF_PAIRS);
        }
        s = s0;
        if (masked) {
            while (hit) {
                const uint32_t t = t0 + (uint32_t)(__ffsll((unsigned long long)hit) - 1);
                hit &= hit - 1;
                while (cand[s + 1] <= t) s++;
                if (out < wd.pair_cap) {
                    d.ps.pa[poff + out] = off + (int)(uint32_t)key[s];
                    d.ps.pb[poff + out] = off + (int)(uint32_t)key[s + 1 + (int)(t - cand[s])];
                }
                out++;
            }
        } else {
            si = -1;
            for (uint32_t t = t0; t < t1; t++) {
                while (cand[s + 1] <= t) s++;
                if (s != si) {
                    si = s;
                    i = (int)(uint32_t)key[s];
                    lds_rec(lrec, b, off, i, A);
                }
                const int j = (int)(uint32_t)key[s + 1 + (int)(t - cand[s])];
                lds_rec(lrec, b, off, j, Bj);
                if (bp_valid_rec(A, Bj)) {
                    if (out < wd.pair_cap) { d.ps.pa[poff + out] = off + i; d.ps.pb[poff + out] = off + j; }
                    out++;
                }
            }
        }
    }
    // next axis: variance of AABB centres (broad_phase_sweep_and_prune.cpp:37-40)
    {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int i = tid; i < n; i += BLOCK) {
            float r[8];
            lds_rec(lrec, b, off, i, r);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float c = (r[k] + r[4 + k]) * 0.5f;
                acc[k] += (double)c;
                acc[3 + k] += (double)(c * c);
            }
        }
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double v = acc[k];
            for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
            if ((tid & (WAVE - 1)) == 0) red[k][tid / WAVE] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double s6[6];
            for (int k = 0; k < 6; k++) { s6[k] = 0; for (int q = 0; q < BLOCK / WAVE; q++) s6[k] += red[k][q]; }
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = s6[3 + k] - s6[k] * s6[k] / (double)n;
            int ax = 0;
            if (v[1] > v[0]) ax = 1;
            if (v[2] > v[ax]) ax = 2;
            d.st[w].axis = ax;
        }
    }
    __syncthreads();

    PE_PHASE(3);
    // hand the pairs to the class's narrowphase list (the order of worlds in it does not matter)
    if (tid == 0) {
        s_base = g.off + atomicAdd(&d.nl_count[3 * g.cls], npairs);
        d.st[w].npairs = npairs;
    }
    __syncthreads();
    for (int p = tid; p < npairs; p += BLOCK) {
        d.nl_pair[s_base + p] = poff + p;
        d.nl_limit[s_base + p] = poff + prev_np;
    }
}

// K2: SimpleNarrowPhase::calcContactResults (simple_narrow_phase.cpp:7-41) over the pairs of
// all worlds at once, one thread per pair.
// Box and sphere pairs here; pairs with a cylinder or a mesh are listed for K2b.
__global__ void __launch_bounds__(128) k_narrow_list(BatchDev d, NarrowSeg g) {
    const int cnt = d.nl_count[3 * g.cls];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int q = g.off + k;
        const int p = d.nl_pair[q];
        const bool simple = pair_is_simple(d.b, d.ps, p);
        const bool sat = !simple && pair_on_wave(d.b, d.ps, p);
        const int slot = wave_append(!simple && !sat, &d.nl_count[3 * g.cls + 1]);
        const int sslot = wave_append(sat, &d.nl_count[3 * g.cls + 2]);
        if (simple) narrow_pair<false>(d.b, d.ps, p, d.nl_limit[q]);
        else if (sat) d.nl_shape[g.off + g.cap - 1 - sslot] = q;
        else d.nl_shape[g.off + slot] = q;
    }
}
// Entry k of the concatenation, over all classes, of the shape entries (which = 1, from the front
// of each segment) or of the separating-axis entries (which = 2, from the back): its list index.
// pre[c] = entries of the classes before c (pre[n] = all), in LDS.
__device__ __forceinline__ int seg_entry(const BatchDev& d, const NarrowSegs& g, const int* pre, int which, int k) {
    int c = 0;
    while (c + 1 < g.n && pre[c + 1] <= k) c++;
    const int j = k - pre[c];
    return which == 1 ? d.nl_shape[g.off[c] + j] : d.nl_shape[g.off[c] + g.cap[c] - 1 - j];
}
__device__ __forceinline__ void seg_prefix(const BatchDev& d, const NarrowSegs& g, int which, int* pre) {
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int c = 0; c < g.n; c++) { pre[c] = acc; acc += d.nl_count[3 * c + which]; }
        pre[g.n] = acc;
    }
    __syncthreads();
}
// K2b: the listed pairs of every class that involve a cylinder, a convex mesh or a concave mesh
// and are not wave pairs (a box or sphere with a cylinder).
// One launch for all classes: these kernels carry large scratch frames, and launching them per
// class multiplied their launch cost (measured: K2 of 96 shape worlds 2.6 -> 4.7 ms).
__global__ void __launch_bounds__(64) k_narrow_shape_list(BatchDev d, NarrowSegs g) {
    __shared__ int pre[MAX_CLASSES + 1];
    seg_prefix(d, g, 1, pre);
    const int cnt = pre[g.n];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int q = seg_entry(d, g, pre, 1, k);
        narrow_pair<true>(d.b, d.ps, d.nl_pair[q], d.nl_limit[q]);
    }
}

// K2c: the listed wave pairs (pair_on_wave: separating-axis pairs, a sphere or cylinder with a
// mesh) of every class, one wave each.
__global__ void __launch_bounds__(256) k_narrow_sat_list(BatchDev d, NarrowSegs g) {
    __shared__ int pre[MAX_CLASSES + 1];
    seg_prefix(d, g, 2, pre);
    const int cnt = pre[g.n];
    const int nw = gridDim.x * (blockDim.x / WAVE);
    for (int k = blockIdx.x * (blockDim.x / WAVE) + (int)(threadIdx.x / WAVE); k < cnt; k += nw) {
        const int q = seg_entry(d, g, pre, 2, k);
        narrow_pair<true, WAVE>(d.b, d.ps, d.nl_pair[q], d.nl_limit[q]);
    }
}

__global__ void k_list_reset(BatchDev d) {
    for (int k = threadIdx.x; k < 3 * MAX_CLASSES; k += blockDim.x) d.nl_count[k] = 0;
}

// K3 for one world on the global row store (claims in LDS): solver setup, colouring,
// colour-ordered Gauss-Seidel sweeps, velocity sync (sequential_impulse_solver.cpp:11-81).
template <int BLOCK>
__device__ void solve_world_global(const BatchDev& d, int w, int* claim, int* wsum) {
    const WorldDesc wd = d.desc[w];
    const int n = wd.n, off = wd.body_off, poff = wd.pair_off;
    const BodyView& b = d.b;
    const int tid = threadIdx.x;
    const int npairs = d.st[w].npairs;
    for (int p = tid; p < npairs; p += BLOCK) setup_pair(b, d.ps, poff + p, d.dt);
    for (int i = tid; i < n; i += BLOCK) claim[i] = 0x7fffffff;
    __syncthreads();
    int* act = d.act + 3 * (size_t)poff;
    int nact = 0, npts = 0;
    for (int base = 0; base < npairs; base += BLOCK) {
        const int p = base + tid;
        int flag = 0, np_ = 0;
        if (p < npairs) {
            np_ = d.ps.npts[poff + p];
            if (np_ > 0) flag = !(b.iat(I_KIN, d.ps.pa[poff + p]) && b.iat(I_KIN, d.ps.pb[poff + p]));
            if (d.ps.swap[poff + p] & 2) {   // the narrowphase's clipper lost points of this pair
                atomicOr(&d.st[w].overflow, WOVF_CLIP);
                atomicOr(d.ovf_any, WOVF_CLIP);
            }
        }
        int tot;
        const int ex = block_exclusive_scan(flag, wsum, &tot);
        if (flag) act[nact + ex] = poff + p;
        nact += tot;
        int totp;
        block_exclusive_scan(np_, wsum, &totp);
        npts += totp;
    }
    __syncthreads();
    int* order = d.order + poff;
    int* cs = d.cstart + wd.cs_off;
    int* bufs[3] = {act, act + wd.pair_cap, act + 2 * (size_t)wd.pair_cap};
    const int ncol = colour_pairs<true>(b, d.ps, bufs, nact, order, cs, claim, off, poff, wsum);
    if (tid == 0) {
        d.st[w].ncolors = ncol;
        d.st[w].npairs = npairs;
        d.st[w].prev_npairs = npairs;
        d.st[w].npoints = npts;
    }
    __syncthreads();
    for (int it = 0; it < d.iterations; it++) {
        for (int c = 0; c < ncol; c++) {
            const int c0 = cs[c], c1 = cs[c + 1];
            for (int q = c0 + tid; q < c1; q += BLOCK) iterate_pair(b, d.ps, order[q]);
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += BLOCK) sync_body_rec(b, off + i);
    __syncthreads();
}

constexpr int MAX_GROUP_WORLDS = 64;

// A group of consecutive worlds of a size class (worlds are laid out in class order, so the
// group's bodies and pair slots are contiguous) solved as one system: its pairs never share a
// body across worlds, so the colouring yields each world's own colour classes and the sweeps
// apply to every world the same updates in the same order as when it is solved alone.
struct Group {
    const int32_t* wl;
    int nw, B0, GB, GP;
};
__device__ __forceinline__ bool group_on_lds(const PrepLayout& L, int GB, int lds_bytes) {
    return L.end <= (size_t)lds_bytes && GB <= 65535;
}
// Reads the group's geometry; s_wb[k] = first local pair of world k (s_wb[nw] = GP).
__device__ __forceinline__ Group load_group(const BatchDev& d, const int32_t* wlist, int2 g, int* s_wb) {
    Group G;
    G.wl = wlist + g.x;
    G.nw = g.y;
    G.B0 = d.desc[G.wl[0]].body_off;
    G.GB = d.desc[G.wl[G.nw - 1]].body_off + d.desc[G.wl[G.nw - 1]].n - G.B0;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < G.nw; k++) { s_wb[k] = acc; acc += d.st[G.wl[k]].npairs; }
        s_wb[G.nw] = acc;
    }
    __syncthreads();
    G.GP = s_wb[G.nw];
    return G;
}
// World k of the group holding local pair lp.
__device__ __forceinline__ int world_of(const int* s_wb, int nw, int lp) {
    int lo = 0, hi = nw - 1;
    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_wb[mid] <= lp) lo = mid; else hi = mid - 1; }
    return lo;
}

// First half of K3: solver rows (FrictionContactConstraint::initSequentialImpulse, to the global
// row store) and the colouring of a group; the solve order (local pair indices) and colour starts
// go to the first world's order and colour-start regions for the sweeps.  A group that does not fit LDS is
// solved here completely, world by world, on the global path.
template <int BLOCK>
__device__ __forceinline__ void group_prep(const BatchDev& d, const int32_t* __restrict__ wlist,
                                           const int2* __restrict__ groups, int lds_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int wsum[BLOCK / WAVE + 1];
    __shared__ int s_wb[MAX_GROUP_WORLDS + 1];
    __shared__ int s_wpts[MAX_GROUP_WORLDS];
    const int tid = threadIdx.x;
    const BodyView& b = d.b;
    const Group G = load_group(d, wlist, groups[blockIdx.x], s_wb);
    const int w = G.wl[0];   // for the phase timestamps
    PE_PHASE(4);
    const PrepLayout Lo(G.GB, G.GP);
    if (!group_on_lds(Lo, G.GB, lds_bytes)) {
        for (int k = 0; k < G.nw; k++) solve_world_global<BLOCK>(d, G.wl[k], (int*)lds, wsum);
        return;
    }
    const int GB = G.GB, GP = G.GP, nw = G.nw, B0 = G.B0;
    if (tid < nw) s_wpts[tid] = 0;
    int* claim = (int*)lds;
    float* imk = (float*)(lds + Lo.bs);   // inverse mass, negative for kinematic bodies
    int* pab = (int*)(lds + Lo.pab);
    int* gidx = (int*)(lds + Lo.gi);
    int* pri = (int*)(lds + Lo.pr);
    int* order = (int*)(lds + Lo.ord);
    int* bufs[3] = {(int*)(lds + Lo.buf), (int*)(lds + Lo.buf) + GP, (int*)(lds + Lo.buf) + 2 * GP};
    int* cs = (int*)(lds + Lo.cs);
    int* rs = (int*)(lds + Lo.rs);
    for (int i = tid; i < GB; i += BLOCK) {
        claim[i] = 0x7fffffff;
        imk[i] = b.iat(I_KIN, B0 + i) ? -1.0f : 0.0f;
    }
    __syncthreads();
    int nact = 0, nrows = 0;
    for (int base = 0; base < GP; base += BLOCK) {
        const int lp = base + tid;
        int flag = 0, np_ = 0;
        if (lp < GP) {
            const int k = world_of(s_wb, nw, lp);
            const int gs = d.desc[G.wl[k]].pair_off + (lp - s_wb[k]);
            const int pa = d.ps.pa[gs], pb = d.ps.pb[gs];
            np_ = d.ps.npts[gs];
            if (d.ps.swap[gs] & 2) {   // the narrowphase's clipper lost points of this pair
                atomicOr(&d.st[G.wl[k]].overflow, WOVF_CLIP);
                atomicOr(d.ovf_any, WOVF_CLIP);
            }
            gidx[lp] = gs;
            pri[lp] = colour::prio(lp - s_wb[k], np_);
            pab[lp] = (pa - B0) | ((pb - B0) << 16);
            if (np_ > 0) {
                flag = !(imk[pa - B0] < 0.0f && imk[pb - B0] < 0.0f);
                atomicAdd(&s_wpts[k], np_);
            }
        }
        int tot;
        const int ex = block_exclusive_scan(flag, wsum, &tot);
        if (flag) bufs[0][nact + ex] = lp;
        nact += tot;
        const int exr = block_exclusive_scan(np_, wsum, &tot);
        if (lp < GP) rs[lp] = nrows + exr;
        nrows += tot;
    }
    if (tid == 0) rs[GP] = nrows;
    __syncthreads();
    // ---- stage 3: solver (sequential_impulse_solver.cpp:11-81)
    // Solver rows, one thread per contact point of the group (a pair's rows are independent):
    // row r belongs to the last local pair whose first row is at most r.
    for (int r = tid; r < nrows; r += BLOCK) {
        int lo = 0, hi = GP - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rs[mid] <= r) lo = mid;
            else hi = mid - 1;
        }
        setup_row(b, d.ps, gidx[lo], B0 + (pab[lo] & 0xffff), B0 + (pab[lo] >> 16), r - rs[lo], d.dt);
    }
    PE_PHASE(5);
    const int ncol = colour_group(pab, imk, pri, bufs, nact, order, cs, claim);
    PE_PHASE(6);
    const WorldDesc w0 = d.desc[G.wl[0]];
    for (int q = tid; q < nact; q += BLOCK) d.order[w0.pair_off + q] = order[q];
    for (int c = tid; c <= ncol; c += BLOCK) d.cstart[w0.cs_off + c] = cs[c];
    for (int k = tid; k < nw; k += BLOCK) {
        const int wk = G.wl[k];
        const int np = s_wb[k + 1] - s_wb[k];
        d.st[wk].ncolors = ncol;
        d.st[wk].npairs = np;
        d.st[wk].prev_npairs = np;
        d.st[wk].npoints = s_wpts[k];
    }
}

// The colour-ordered passes of a group: thread tid solves the pairs at order slots
// cs[c] + tid + k * BLOCK of every colour c in every iteration, so it alone reads and writes their
// rows.  pair(lp, true, cur, nxt) loads the first two rows of local pair lp, pair(lp, false, cur,
// nxt) solves the pair from them; the first rows of a thread's first pair of the next colour are
// loaded right after its pairs of this colour are done, ahead of the barrier.
template <int BLOCK, class PairFn>
__device__ __forceinline__ void sweep_passes(int iterations, int ncol, const int* cs, const int* order, int tid,
                                             PairFn pair) {
    float cur[RW], nxt[RW];
    int pend = -1;   // local pair whose first rows cur and nxt hold
    auto fetch = [&](int c) {
        const int q = cs[c] + tid;
        pend = q < cs[c + 1] ? order[q] : -1;
        if (pend >= 0) pair(pend, true, cur, nxt);
    };
    if (ncol > 0) fetch(0);
    for (int it = 0; it < iterations; it++) {
        for (int c = 0; c < ncol; c++) {
            if (pend >= 0) {
                pair(pend, false, cur, nxt);
                for (int q = cs[c] + tid + BLOCK; q < cs[c + 1]; q += BLOCK) {
                    const int lp = order[q];
                    pair(lp, true, cur, nxt);
                    pair(lp, false, cur, nxt);
                }
            }
            const bool last = it + 1 == iterations && c + 1 == ncol;
            if (!last) fetch(c + 1 == ncol ? 0 : c + 1);
            __syncthreads();
        }
    }
}

// Second half of K3: the colour-ordered Gauss-Seidel sweeps of a prepared group on LDS copies of its bodies
// and pairs (rows in the global row store: staging them in LDS measured slower, as it lowers the
// workgroups a CU holds), then syncTempVelocity.  Its passes are chains of dependent arithmetic
// separated by barriers.
template <int BLOCK>
__device__ __forceinline__ void group_sweep(const BatchDev& d, const int32_t* __restrict__ wlist,
                                            const int2* __restrict__ groups, int prep_lds_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int wsum[BLOCK / WAVE + 1];
    __shared__ int s_wb[MAX_GROUP_WORLDS + 1];
    const int tid = threadIdx.x;
    const BodyView& b = d.b;
    const Group G = load_group(d, wlist, groups[blockIdx.x], s_wb);
    if (!group_on_lds(PrepLayout(G.GB, G.GP), G.GB, prep_lds_bytes)) return;   // solved by group_prep
    const int GB = G.GB, GP = G.GP, nw = G.nw, B0 = G.B0;
    const int w = G.wl[0];
    auto al16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    LdsSolve L;
    L.bs = (float4*)lds;
    L.pab = (int*)(lds + (size_t)GB * 64);
    int* gidx = L.pab + GP;
    L.rs = gidx + GP;
    int* order = (int*)(lds + al16((size_t)GB * 64 + (size_t)(3 * GP + 1) * 4));
    L.n = GB;
    const WorldDesc w0 = d.desc[w];
    const int ncol = d.st[w].ncolors;
    for (int i = tid; i < GB; i += BLOCK) lds_load_body(L, b, B0 + i, i);
    int npts = 0;
    for (int base = 0; base < GP; base += BLOCK) {
        const int lp = base + tid;
        int np_ = 0;
        if (lp < GP) {
            const int k = world_of(s_wb, nw, lp);
            const int gs = d.desc[G.wl[k]].pair_off + (lp - s_wb[k]);
            gidx[lp] = gs;
            L.pab[lp] = (d.ps.pa[gs] - B0) | ((d.ps.pb[gs] - B0) << 16);
            np_ = d.ps.npts[gs];
        }
        int tot;
        const int ex = block_exclusive_scan(np_, wsum, &tot);
        if (lp < GP) L.rs[lp] = npts + ex;
        npts += tot;
    }
    if (tid == 0) L.rs[GP] = npts;
    const int nord = d.cstart[w0.cs_off + ncol];
    for (int q = tid; q < nord; q += BLOCK) order[q] = d.order[w0.pair_off + q];
    int* cs = order + nord;   // colour starts (ncol + 1 <= GP + 1 entries fit the layout's bound)
    for (int q = tid; q <= ncol; q += BLOCK) cs[q] = d.cstart[w0.cs_off + q];
    __syncthreads();
    sweep_passes<BLOCK>(d.iterations, ncol, cs, order, tid, [&](int lp, bool first, float (&cur)[RW], float (&nxt)[RW]) {
        float* rows = d.ps.rows + (size_t)gidx[lp] * 8 * RW;
        if (first) prefetch_pair_rows(L, lp, rows, cur, nxt);
        else iterate_pair_lds(L, lp, rows, cur, nxt);
    });
    PE_PHASE(7);
    // the solved velocities into the solver records' v, w (floats 17..22; float 16 is the last
    // inverse-inertia entry, kept as it is)
    for (int i = tid; i < GB; i += BLOCK) {
        const float4 q0 = L.bs[4 * i], q1 = L.bs[4 * i + 1], q3 = L.bs[4 * i + 3];
        if (q1.z < 0.0f) continue;
        float4* sr = reinterpret_cast<float4*>(b.sbr(B0 + i));
        sr[4] = make_float4(q3.w, q0.x, q0.y, q0.z);
        sr[5] = make_float4(q0.w, q1.x, q1.y, 0.0f);
    }
}

// K3: a group's preparation (solver rows, colouring) and its Gauss-Seidel sweeps in one launch, so
// the sweeps of a group start as soon as its own preparation is done rather than after every group's.
// WAVES is the minimum number of waves per SIMD the register allocation must allow: 4 (at most
// 128 VGPRs) for classes whose groups are solved in LDS -- the sweeps are chains of dependent
// arithmetic, so the more workgroups a CU holds the better; it spills only on the global path,
// which such classes take only for a group with more than 2 pairs per body.  Classes of worlds too
// large for LDS run the global path and use the unconstrained build (WAVES = 1).
template <int BLOCK, int WAVES>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_group_solve(BatchDev d, const int32_t* __restrict__ wlist, const int2* __restrict__ groups, int lds_bytes) {
    group_prep<BLOCK>(d, wlist, groups, lds_bytes);
    __syncthreads();
    group_sweep<BLOCK>(d, wlist, groups, lds_bytes);
}

__global__ void k_refresh_range(BodyView b, int first, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) refresh_body(b, first + i);
}

// ------------------------------------------------------------------ host-side helpers
// Pair slots per world: 16 per body (a resting cubic packing of equal boxes has 13 AABB pairs per
// body: 26 neighbours across faces, edges and corners).  A world that finds more has its extra
// pairs dropped and raises WOVF_PAIRS, reported as PE_ERR_CAPACITY.
int body_pair_cap(int n) { return 16 * n + 64; }

// Allows every instantiation the full LDS of a CU (worlds up to 4096 bodies need ~64 KiB).
template <int B> void set_lds_one() {
    const int bytes = 150 * 1024;
    hipFuncSetAttribute((const void*)k_world_broad<B>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    hipFuncSetAttribute((const void*)k_group_solve<B, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    hipFuncSetAttribute((const void*)k_group_solve<B, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}
void set_lds_limits() {
    static bool done = false;
    if (done) return;
    done = true;
    set_lds_one<64>(); set_lds_one<128>(); set_lds_one<256>(); set_lds_one<512>(); set_lds_one<1024>();
}

// Launches K1 (phase 0) or K3 (phase 1) of one class: dispatch over the block sizes.
// K1 of a class: one workgroup per world of wl[0..count).  K3: one workgroup per group of
// groups[0..ngroups) (first index into the world list, world count).
template <int B>
void launch_one(int phase, const BatchDev& d, const int32_t* wl, int count, int nmax, const int2* groups, int ngroups,
                int group_bodies, NarrowSeg g, hipStream_t s) {
    if (phase == 0) {
        hipLaunchKernelGGL(k_world_broad<B>, dim3(count), dim3(B), broad_lds_bytes(nmax), s, d, wl, nmax, g);
    } else {
        const size_t lb = solve_lds_bytes(nmax, group_bodies);
        if (lb < SOLVE_LDS_CAP)
            hipLaunchKernelGGL((k_group_solve<B, 4>), dim3(ngroups), dim3(B), lb, s, d, wl, groups, (int)lb);
        else
            hipLaunchKernelGGL((k_group_solve<B, 1>), dim3(ngroups), dim3(B), lb, s, d, wl, groups, (int)lb);
    }
}
hipError_t launch_class(int phase, int block, const BatchDev& d, const int32_t* wl, int count, int nmax, const int2* groups,
                        int ngroups, int group_bodies, NarrowSeg g, hipStream_t s) {
    switch (block) {
        case 64: launch_one<64>(phase, d, wl, count, nmax, groups, ngroups, group_bodies, g, s); break;
        case 128: launch_one<128>(phase, d, wl, count, nmax, groups, ngroups, group_bodies, g, s); break;
        case 256: launch_one<256>(phase, d, wl, count, nmax, groups, ngroups, group_bodies, g, s); break;
        case 512: launch_one<512>(phase, d, wl, count, nmax, groups, ngroups, group_bodies, g, s); break;
        default: launch_one<1024>(phase, d, wl, count, nmax, groups, ngroups, group_bodies, g, s); break;
    }
    return hipGetLastError();
}

constexpr int DEFAULT_GROUP_BODIES = 256;   // bodies solved together per K3 workgroup


struct Class {
    int lo, hi;         // body-count range [lo, hi]
    int block;
    int group_bodies = DEFAULT_GROUP_BODIES;   // K3 groups: consecutive worlds up to this many bodies
    int count = 0;      // worlds in the class
    int nmax = 0;
    int list_off = 0;   // offset into the device world list
    int group_off = 0;  // offset into the device group list
    int ngroups = 0;
    double cost = -1.0; // measured K1 + K2 + K3 device time per step in ms (< 0: not measured)
    int stream = 0;     // 0: the main stream, k > 0: side stream k - 1
    int nl_off = 0;     // narrowphase list segment: the summed pair capacity of the classes before it
    int nl_cap = 0;     // and of its own worlds
};

// Groups of consecutive class members (first index into the world list, count) holding at most
// group_bodies bodies and MAX_GROUP_WORLDS worlds each (at least one world); group_bodies = 0
// gives one world per group.
void build_groups(Class& c, const std::vector<int32_t>& members, const std::vector<int>& nbodies, bool contiguous,
                  std::vector<int2>& out) {
    c.group_off = (int)out.size();
    int k = 0;
    const int m = (int)members.size();
    while (k < m) {
        int cnt = 1, tot = nbodies[members[k]];
        if (contiguous && c.group_bodies > 0)
            while (k + cnt < m && cnt < MAX_GROUP_WORLDS && tot + nbodies[members[k + cnt]] <= c.group_bodies) tot += nbodies[members[k + cnt++]];
        out.push_back(make_int2(c.list_off + k, cnt));
        k += cnt;
    }
    c.ngroups = (int)out.size() - c.group_off;
}

// Side streams the size classes of K1 and K3 run on concurrently (a class's launch alone rarely
// fills the GPU: its runtime is set by its largest worlds' dependent passes, not by its work).
// Classes are dealt round-robin onto the main stream and up to MAX_SIDE side streams, forked
// from and joined back to the main stream by events, so the step stays one ordered unit on the
// main stream (and a captured graph holds the same fork / join dependencies).
constexpr int MAX_SIDE = 3;
struct ClassStreams {
    hipStream_t side[MAX_SIDE] = {};
    hipEvent_t fork = nullptr, join[MAX_SIDE] = {};
    int n = 0;
    int create() {
        if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return PE_ERR_DEVICE;
        for (; n < MAX_SIDE; n++) {
            if (hipStreamCreateWithFlags(&side[n], hipStreamNonBlocking) != hipSuccess) return PE_ERR_DEVICE;
            if (hipEventCreateWithFlags(&join[n], hipEventDisableTiming) != hipSuccess) return PE_ERR_DEVICE;
        }
        return PE_OK;
    }
    void release() {
        for (int k = 0; k < n; k++) {
            if (side[k]) { hipStreamSynchronize(side[k]); hipStreamDestroy(side[k]); }
            if (join[k]) hipEventDestroy(join[k]);
        }
        if (fork) hipEventDestroy(fork);
        n = 0;
        fork = nullptr;
    }
};

// Per-stage timing: an event pair on the main stream around each stage (0 = K1 integrate +
// broadphase of all classes, 1 = K2 narrowphase, 2 = K3 solve of all classes), so a stage's
// device time is its span from before its first launch to after its last, without the gaps
// between stages.
struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> stage;
    int used = 0;
    ~LaunchTimer() { for (auto e : ev) hipEventDestroy(e); }
    hipError_t begin(int st, hipStream_t s) {
        if (used + 2 > (int)ev.size()) {
            for (int k = 0; k < 2; k++) {
                hipEvent_t e;
                hipError_t r = hipEventCreate(&e);
                if (r != hipSuccess) return r;
                ev.push_back(e);
            }
            stage.resize(ev.size() / 2);
        }
        stage[used / 2] = st;
        return hipEventRecord(ev[used], s);
    }
    hipError_t end(hipStream_t s) { hipError_t r = hipEventRecord(ev[used + 1], s); used += 2; return r; }
};

// Deals the classes onto nstreams streams, longest first onto the least-loaded stream (the
// heaviest class gets the main stream).  A class's load is its measured cost, or before any
// measurement its largest world's body count (the length of its longest chain of passes).
void schedule_classes(std::vector<Class>& classes, int nstreams) {
    std::vector<int> ord;
    for (int c = 0; c < (int)classes.size(); c++)
        if (classes[c].count > 0) ord.push_back(c);
    bool measured = true;
    for (int c : ord) measured = measured && classes[c].cost >= 0.0;
    auto load = [&](int c) { return measured ? classes[c].cost : (double)classes[c].nmax; };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return load(a) > load(b); });
    std::vector<double> acc(std::max(1, nstreams), 0.0);
    for (int c : ord) {
        const int k = (int)(std::min_element(acc.begin(), acc.end()) - acc.begin());
        classes[c].stream = k;
        acc[k] += load(c);
    }
}

// K2 of one class over its list segment (box and sphere pairs; the shape pairs it lists run in
// launch_shapes); the grid scales with the segment's pair capacity (16 slots per body; about
// twelve slots per listed pair: about a pair and a half per thread) and the kernel strides over
// what was listed.
void launch_narrow(const BatchDev& d, const Class& c, int ci, hipStream_t t) {
    const NarrowSeg g{c.nl_off, c.nl_cap, ci};
    hipLaunchKernelGGL(k_narrow_list, dim3((unsigned)std::max(1, std::min(4096, c.nl_cap / 2048))), dim3(128), 0, t, d, g);
}
// K2b and K2c over the shape pairs every class listed.
void launch_shapes(const BatchDev& d, const std::vector<Class>& classes, hipStream_t t) {
    NarrowSegs g{};
    g.n = (int)classes.size();
    for (int c = 0; c < g.n; c++) { g.off[c] = classes[c].nl_off; g.cap[c] = classes[c].nl_cap; }
    hipLaunchKernelGGL(k_narrow_shape_list, dim3(4096), dim3(64), 0, t, d, g);
    hipLaunchKernelGGL(k_narrow_sat_list, dim3(2048), dim3(256), 0, t, d, g);
}

// One step of every listed world: list reset, then per class K1, K2 over its own pairs and K3.
// Pipelined (worlds of boxes and spheres): classes share nothing within a step, so each class's
// three launch groups run in sequence on the stream schedule_classes dealt it (the main stream s
// or a side stream, heaviest class first on each), and the classes of different streams overlap
// freely: no stage waits for every class.  Staged (worlds with other shapes, and every timed
// step): K1 of all classes, then K2 of all classes followed by one launch of the shape kernels
// over every class's shape pairs (launch_shapes), then K3 of all classes, the classes of each
// stage concurrently on their streams.  With a timer each stage is timed as stage 0, 1 and 2;
// with class_tm every class runs on s alone, in turn, its K1 + K2 and its K3 each timed by their
// own event pair (stage 3 + 2 * class + 0 or 1; the shape kernels run untimed between the two).
int launch_all(const BatchDev& d, const std::vector<Class>& classes, const int32_t* dlist, const int2* dgroups, hipStream_t s,
               const ClassStreams& cs, bool shapes, LaunchTimer* tm = nullptr, bool class_tm = false) {
    pe_clear_stale_error();
    hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, s, d);
    std::vector<int> ord;
    int nused = 0;
    for (int c = 0; c < (int)classes.size(); c++)
        if (classes[c].count > 0) { ord.push_back(c); nused = std::max(nused, classes[c].stream + 1); }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return classes[a].nmax > classes[b].nmax; });
    const bool fork = !class_tm && nused > 1 && nused <= cs.n + 1;
    auto stream_of = [&](const Class& c) { return fork && c.stream > 0 ? cs.side[c.stream - 1] : s; };
    auto k1 = [&](int ci, hipStream_t t) {
        const Class& c = classes[ci];
        return launch_class(0, c.block, d, dlist + c.list_off, c.count, c.nmax, nullptr, 0, 0,
                            NarrowSeg{c.nl_off, c.nl_cap, ci}, t);
    };
    auto k3 = [&](int ci, hipStream_t t) {
        const Class& c = classes[ci];
        return launch_class(1, c.block, d, dlist, c.count, c.nmax, dgroups + c.group_off, c.ngroups, c.group_bodies,
                            NarrowSeg{c.nl_off, c.nl_cap, ci}, t);
    };
    auto fork_out = [&]() {
        if (!fork) return;
        hipEventRecord(cs.fork, s);
        for (int k = 1; k < nused; k++) hipStreamWaitEvent(cs.side[k - 1], cs.fork, 0);
    };
    auto join_in = [&]() {
        if (!fork) return;
        for (int k = 1; k < nused; k++) {
            hipEventRecord(cs.join[k - 1], cs.side[k - 1]);
            hipStreamWaitEvent(s, cs.join[k - 1], 0);
        }
    };
    hipError_t e = hipSuccess;
    if (!tm && !shapes) {
        fork_out();
        for (int ci : ord) {
            const hipStream_t t = stream_of(classes[ci]);
            if ((e = k1(ci, t)) != hipSuccess) break;
            launch_narrow(d, classes[ci], ci, t);
            if ((e = k3(ci, t)) != hipSuccess) break;
        }
        join_in();
    } else if (class_tm) {
        for (int ci : ord) {
            tm->begin(3 + 2 * ci, s);
            if ((e = k1(ci, s)) != hipSuccess) break;
            launch_narrow(d, classes[ci], ci, s);
            tm->end(s);
        }
        if (shapes) launch_shapes(d, classes, s);
        for (int ci : ord) {
            tm->begin(3 + 2 * ci + 1, s);
            if ((e = k3(ci, s)) != hipSuccess) break;
            tm->end(s);
        }
    } else {
        if (tm) tm->begin(0, s);
        fork_out();
        for (int ci : ord) if ((e = k1(ci, stream_of(classes[ci]))) != hipSuccess) break;
        join_in();
        if (tm) tm->end(s);
        if (tm) tm->begin(1, s);
        fork_out();
        for (int ci : ord) launch_narrow(d, classes[ci], ci, stream_of(classes[ci]));
        join_in();
        if (shapes) launch_shapes(d, classes, s);
        if (tm) tm->end(s);
        if (tm) tm->begin(2, s);
        fork_out();
        for (int ci : ord) if (e == hipSuccess && (e = k3(ci, stream_of(classes[ci]))) != hipSuccess) break;
        join_in();
        if (tm) tm->end(s);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) { pe_set_error(std::string("launch: ") + hipGetErrorString(e)); return PE_ERR_DEVICE; }
    return PE_OK;
}

// Default class table: body-count ranges and their block sizes.
std::vector<Class> default_classes() {
    return {{1, 24, 64}, {25, 64, 64}, {65, 160, 128}, {161, 384, 256}, {385, 1024, 256}, {1025, 1 << 30, 512}};
}

}  // namespace

// ====================================================================== batch object
struct pe_batch {
    int nworlds = 0;
    int total_bodies = 0, stride = 0;
    int64_t total_pair_cap = 0;
    std::vector<WorldDesc> hdesc;
    BatchDev d{};
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;
    std::vector<void*> snap;       // copies of the state allocations (pe_batch_snapshot)
    std::vector<Class> classes;
    int32_t* dlist = nullptr;      // world ids grouped by class
    int2* dgroups = nullptr;       // K3 groups per class
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int64_t dyn_bodies = 0;

    DevMeshes meshes;
    bool has_shapes = false;   // any cylinder / convex / concave body
    ClassStreams side;
    ~pe_batch() {
        if (stream) hipStreamSynchronize(stream);
        side.release();
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        for (void* p : allocs) hipFree(p);
        for (void* p : snap) if (p) hipFree(p);
        meshes.release();
        if (stream) hipStreamDestroy(stream);
    }
    template <class T> int alloc(T** p, size_t count) {
        void* q = nullptr;
        PE_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        PE_HIP(hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), stream));
        allocs.push_back(q);
        alloc_bytes.push_back(std::max<size_t>(count, 1) * sizeof(T));
        *p = (T*)q;
        return PE_OK;
    }
    // Copies every state allocation to (save) or from (restore) its snapshot copy; the class
    // lists are configuration, not state, and are left alone.
    int copy_state(bool save) {
        if (save && snap.empty()) {
            snap.assign(allocs.size(), nullptr);
            for (size_t k = 0; k < allocs.size(); k++) {
                if (allocs[k] == (void*)dlist || allocs[k] == (void*)dgroups) continue;
                PE_HIP(hipMalloc(&snap[k], alloc_bytes[k]));
            }
        }
        if (snap.size() != allocs.size()) { pe_set_error("no snapshot to restore"); return PE_ERR_INVALID; }
        for (size_t k = 0; k < allocs.size(); k++) {
            if (!snap[k]) continue;
            PE_HIP(hipMemcpyAsync(save ? snap[k] : allocs[k], save ? allocs[k] : snap[k], alloc_bytes[k],
                                  hipMemcpyDeviceToDevice, stream));
        }
        PE_HIP(hipStreamSynchronize(stream));
        return PE_OK;
    }
    // Groups worlds into classes (largest worlds first inside a class) and uploads the list.
    int assign_classes() {
        std::vector<std::vector<int>> members(classes.size());
        for (int w = 0; w < nworlds; w++) {
            const int n = hdesc[w].n;
            for (size_t c = 0; c < classes.size(); c++)
                if (n >= classes[c].lo && n <= classes[c].hi) { members[c].push_back(w); break; }
        }
        std::vector<int32_t> list;
        std::vector<int2> grp;
        std::vector<int> nb(nworlds);
        for (int w = 0; w < nworlds; w++) nb[w] = hdesc[w].n;
        int nl = 0;
        for (size_t c = 0; c < classes.size(); c++) {
            auto& m = members[c];
            std::stable_sort(m.begin(), m.end(), [&](int a, int b) { return hdesc[a].n > hdesc[b].n; });
            classes[c].count = (int)m.size();
            classes[c].nmax = m.empty() ? 0 : hdesc[m[0]].n;
            classes[c].nl_off = nl;
            classes[c].nl_cap = 0;
            for (int w : m) classes[c].nl_cap += hdesc[w].pair_cap;
            nl += classes[c].nl_cap;
            classes[c].list_off = (int)list.size();
            list.insert(list.end(), m.begin(), m.end());
            // the layout follows this order (largest worlds first, ties by index), so consecutive
            // members are contiguous in the body and pair stores
            build_groups(classes[c], m, nb, true, grp);
        }
        if (!dlist && alloc(&dlist, std::max(1, nworlds)) != PE_OK) return PE_ERR_DEVICE;
        if (!dgroups && alloc(&dgroups, std::max(1, nworlds)) != PE_OK) return PE_ERR_DEVICE;
        PE_HIP(hipMemcpyAsync(dlist, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        PE_HIP(hipMemcpyAsync(dgroups, grp.data(), grp.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
        PE_HIP(hipStreamSynchronize(stream));
        schedule_classes(classes, side.n + 1);
        return PE_OK;
    }
    int launch_step() { return launch_all(d, classes, dlist, dgroups, stream, side, has_shapes); }
    void drop_graph() {
        if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { hipGraphDestroy(graph); graph = nullptr; }
    }
};

extern "C" {

int pe_batch_world_footprint(const pe_scene_desc* scene, int64_t* bytes) {
    if (!scene || !bytes) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    const int64_t n = scene->n, pc = body_pair_cap(scene->n);
    *bytes = n * (int64_t)(NF + NI + CBW + SBW + 8) * 4 + pc * (8 * (int64_t)(CW + 1 + RW) * 4 + 9 * 4 + 4 * 4) + (int64_t)sizeof(WorldDesc) +
             (int64_t)sizeof(WorldState);
    return PE_OK;
}

namespace {
// Batched and served worlds step with the reference's default (zero) sleep thresholds, bodies that
// live forever and no compound shapes: those need the single-world path.
int check_batched_scene(const pe_scene_desc* sc) {
    if (int rc = check_scene(sc)) return rc;
    for (int i = 0; i < sc->n; i++)
        if (sc->shape_type[i] == PE_SHAPE_COMPOUND) {
            pe_set_error("body " + std::to_string(i) + " is a compound: compounds need a pe_world");
            return PE_ERR_UNSUPPORTED;
        }
    if (sc->lifetime)
        for (int i = 0; i < sc->n; i++)
            if (!(sc->lifetime[i] >= 1e30)) {
                pe_set_error("body " + std::to_string(i) + " has a finite lifetime: lifetimes need a pe_world");
                return PE_ERR_INVALID;
            }
    return PE_OK;
}
}  // namespace

int pe_batch_create(int32_t nworlds, const pe_scene_desc* scenes, const pe_world_params* prm, pe_batch** out) {
    if (nworlds <= 0 || !scenes || !prm || !out) { pe_set_error("pe_batch_create: invalid argument"); return PE_ERR_INVALID; }
    for (int w = 0; w < nworlds; w++) {
        if (int rc = check_batched_scene(&scenes[w])) return rc;
        if (scenes[w].n > PE_BATCH_MAX_WORLD_BODIES) { pe_set_error("world too large for the batched kernel"); return PE_ERR_INVALID; }
    }
    set_lds_limits();
    auto* B = new pe_batch();
    auto fail = [&](int rc) { delete B; return rc; };
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess || B->side.create() != PE_OK) {
        pe_set_error("stream");
        return fail(PE_ERR_DEVICE);
    }
    B->nworlds = nworlds;
    B->hdesc.resize(nworlds);
    int boff = 0;
    int64_t poff = 0;
    // lay worlds out largest first (ties by index), the order of every size class's world list
    std::vector<int> lay(nworlds);
    std::iota(lay.begin(), lay.end(), 0);
    std::stable_sort(lay.begin(), lay.end(), [&](int a, int b) { return scenes[a].n > scenes[b].n; });
    int rank = 0;
    for (int w : lay) {
        B->hdesc[w].body_off = boff;
        B->hdesc[w].n = scenes[w].n;
        B->hdesc[w].pair_off = (int32_t)poff;
        B->hdesc[w].pair_cap = body_pair_cap(scenes[w].n);
        B->hdesc[w].cs_off = (int32_t)poff + rank++;
        boff += scenes[w].n;
        poff += B->hdesc[w].pair_cap;
    }
    if (poff > (int64_t)1 << 30) { pe_set_error("batch too large"); return fail(PE_ERR_INVALID); }
    B->total_bodies = boff;
    B->stride = (boff + 63) & ~63;
    B->total_pair_cap = poff;
    BatchDev& d = B->d;
    d.dt = (float)prm->dt;
    for (int k = 0; k < 3; k++) d.g[k] = (float)prm->gravity[k];
    d.iterations = prm->iterations;
    d.b.n = boff;
    d.b.stride = B->stride;
    d.ps.cs = (size_t)poff * 8;
    const size_t S = B->stride;
    int rc = PE_OK;
    rc |= B->alloc(&d.b.f, NF * S);
    rc |= B->alloc(&d.b.i, NI * S);
    rc |= B->alloc(&d.b.cb, (size_t)CBW * S);
    rc |= B->alloc(&d.b.sb, (size_t)SBW * S);
    rc |= B->alloc(&d.b.ab, (size_t)8 * S);
    rc |= B->alloc(&d.ps.pa, poff);
    rc |= B->alloc(&d.ps.pb, poff);
    rc |= B->alloc(&d.ps.cp, (size_t)CW * poff * 8);
    rc |= B->alloc(&d.ps.cd, (size_t)poff * 8);
    rc |= B->alloc(&d.ps.npts, poff);
    rc |= B->alloc(&d.ps.swap, poff);
    rc |= B->alloc(&d.ps.rows, (size_t)RW * poff * 8);
    rc |= B->alloc(&d.act, 3 * (size_t)poff);
    rc |= B->alloc(&d.order, poff);
    rc |= B->alloc(&d.cstart, poff + nworlds);
    rc |= B->alloc(&d.nl_pair, poff);
    rc |= B->alloc(&d.nl_limit, poff);
    rc |= B->alloc(&d.nl_shape, poff);
    rc |= B->alloc(&d.nl_count, 3 * MAX_CLASSES);
    rc |= B->alloc(&d.ovf_any, 1);
    WorldDesc* ddesc = nullptr;
    WorldState* dst = nullptr;
    rc |= B->alloc(&ddesc, nworlds);
    rc |= B->alloc(&dst, nworlds);
    if (rc != PE_OK) return fail(PE_ERR_DEVICE);
    d.desc = ddesc;
    d.st = dst;
    std::vector<float> hf(NF * S, 0.0f);
    std::vector<int32_t> hi(NI * S, 0);
    int64_t ndyn = 0;
    HostMeshes hm;
    for (int w = 0; w < nworlds; w++) {
        if (int rc2 = stage_bodies(&scenes[w], hf, hi, S, B->hdesc[w].body_off, hm)) return fail(rc2);
        for (int i = 0; i < scenes[w].n; i++) {
            ndyn += scenes[w].kinematic[i] ? 0 : 1;
            B->has_shapes |= scenes[w].shape_type[i] > PE_SHAPE_SPHERE;
        }
    }
    if (int rc2 = B->meshes.upload(hm, B->stream)) return fail(rc2);
    d.b.ms = B->meshes.view();
    B->dyn_bodies = ndyn;
    if (hipMemcpyAsync(d.b.f, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, B->stream) != hipSuccess ||
        hipMemcpyAsync(d.b.i, hi.data(), hi.size() * 4, hipMemcpyHostToDevice, B->stream) != hipSuccess ||
        hipMemcpyAsync(ddesc, B->hdesc.data(), nworlds * sizeof(WorldDesc), hipMemcpyHostToDevice, B->stream) != hipSuccess) {
        pe_set_error("upload failed");
        return fail(PE_ERR_DEVICE);
    }
    hipLaunchKernelGGL(k_refresh_range, dim3(std::min(4096, (boff + 255) / 256)), dim3(256), 0, B->stream, d.b, 0, boff);
    B->classes = default_classes();
    if (B->assign_classes() != PE_OK) return fail(PE_ERR_DEVICE);
    if (hipStreamSynchronize(B->stream) != hipSuccess) { pe_set_error("init sync"); return fail(PE_ERR_DEVICE); }
    *out = B;
    return PE_OK;
}

int pe_batch_destroy(pe_batch* B) {
    delete B;
    return PE_OK;
}

int pe_batch_set_config(pe_batch* B, int32_t nclasses, const int32_t* upper, const int32_t* block) {
    if (!B || nclasses <= 0 || !upper || !block) { pe_set_error("invalid config"); return PE_ERR_INVALID; }
    if (nclasses > MAX_CLASSES) { pe_set_error("at most 32 size classes"); return PE_ERR_INVALID; }
    std::vector<Class> cl;
    int lo = 1;
    for (int c = 0; c < nclasses; c++) {
        const int bs = block[c];
        if (bs != 64 && bs != 128 && bs != 256 && bs != 512 && bs != 1024) { pe_set_error("block size must be 64..1024, power of 2"); return PE_ERR_INVALID; }
        const int hi = c == nclasses - 1 ? (1 << 30) : upper[c];
        if (hi < lo) { pe_set_error("class bounds must increase"); return PE_ERR_INVALID; }
        cl.push_back(Class{lo, hi, bs, c < (int)B->classes.size() ? B->classes[c].group_bodies : DEFAULT_GROUP_BODIES});
        lo = hi + 1;
    }
    B->classes = cl;
    B->drop_graph();
    return B->assign_classes();
}

int pe_batch_set_solver_groups(pe_batch* B, int32_t nclasses, const int32_t* group_bodies) {
    if (!B || !group_bodies || nclasses != (int32_t)B->classes.size()) { pe_set_error("one value per class expected"); return PE_ERR_INVALID; }
    for (int c = 0; c < nclasses; c++) {
        if (group_bodies[c] < 0 || group_bodies[c] > 2048) { pe_set_error("group bodies must be 0..2048"); return PE_ERR_INVALID; }
        B->classes[c].group_bodies = group_bodies[c];
        B->classes[c].cost = -1.0;
    }
    B->drop_graph();
    return B->assign_classes();
}

int pe_batch_get_solver_groups(pe_batch* B, int32_t* nclasses, int32_t* group_bodies, int32_t cap) {
    if (!B || !nclasses) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    *nclasses = (int32_t)B->classes.size();
    for (int c = 0; c < (int)B->classes.size() && c < cap; c++)
        if (group_bodies) group_bodies[c] = B->classes[c].group_bodies;
    return PE_OK;
}

int pe_batch_get_config(pe_batch* B, int32_t* nclasses, int32_t* upper, int32_t* block, int32_t* count, int32_t cap) {
    if (!B || !nclasses) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    *nclasses = (int32_t)B->classes.size();
    for (int c = 0; c < (int)B->classes.size() && c < cap; c++) {
        if (upper) upper[c] = std::min(B->classes[c].hi, 1 << 30);
        if (block) block[c] = B->classes[c].block;
        if (count) count[c] = B->classes[c].count;
    }
    return PE_OK;
}

int pe_batch_capture_graph(pe_batch* B) {
    if (!B) { pe_set_error("null batch"); return PE_ERR_INVALID; }
    B->drop_graph();
    PE_HIP(hipStreamSynchronize(B->stream));
    PE_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    int rc = B->launch_step();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(B->stream, &g);
    if (rc != PE_OK) return rc;
    if (e != hipSuccess) { pe_set_error(std::string("capture: ") + hipGetErrorString(e)); return PE_ERR_DEVICE; }
    B->graph = g;
    PE_HIP(hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0));
    return PE_OK;
}

int pe_batch_step(pe_batch* B, int32_t nsteps) {
    if (!B) { pe_set_error("null batch"); return PE_ERR_INVALID; }
    for (int s = 0; s < nsteps; s++) {
        if (B->gexec) PE_HIP(hipGraphLaunch(B->gexec, B->stream));
        else if (int rc = B->launch_step()) return rc;
    }
    return PE_OK;
}

// Drains the batch stream, then reports the capacity errors raised since the last report
// (PE_ERR_CAPACITY naming the worlds; their flags stay set for pe_batch_get_counts).
int pe_batch_sync(pe_batch* B) {
    if (!B) { pe_set_error("null batch"); return PE_ERR_INVALID; }
    PE_HIP(hipStreamSynchronize(B->stream));
    int32_t any = 0;
    PE_HIP(hipMemcpy(&any, B->d.ovf_any, sizeof(any), hipMemcpyDeviceToHost));
    if (any == 0) return PE_OK;
    std::vector<WorldState> st(B->nworlds);
    PE_HIP(hipMemcpy(st.data(), B->d.st, st.size() * sizeof(WorldState), hipMemcpyDeviceToHost));
    PE_HIP(hipMemsetAsync(B->d.ovf_any, 0, sizeof(int32_t), B->stream));
    PE_HIP(hipStreamSynchronize(B->stream));
    std::string m = "capacity exceeded in worlds";
    int listed = 0, total = 0;
    for (int w = 0; w < B->nworlds; w++) {
        if (!st[w].overflow) continue;
        total++;
        if (listed++ < 8)
            m += " " + std::to_string(w) + ((st[w].overflow & WOVF_PAIRS) ? " (pairs beyond " + std::to_string(B->hdesc[w].pair_cap) + ")" : "") +
                 ((st[w].overflow & WOVF_CLIP) ? " (clip buffer)" : "");
    }
    if (total > listed) m += " and " + std::to_string(total - listed) + " more";
    pe_set_error(m);
    return PE_ERR_CAPACITY;
}

int pe_batch_get_world_state(pe_batch* B, int32_t w, float* pos, float* rot, float* v, float* av) {
    if (!B || w < 0 || w >= B->nworlds) { pe_set_error("world index out of range"); return PE_ERR_INVALID; }
    PE_HIP(hipStreamSynchronize(B->stream));
    const WorldDesc& wd = B->hdesc[w];
    // a batched world's pose and velocities live in its collision and solver records
    std::vector<float> cb((size_t)wd.n * CBW), sb((size_t)wd.n * SBW);
    PE_HIP(hipMemcpy(cb.data(), B->d.b.cb + (size_t)wd.body_off * CBW, cb.size() * 4, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(sb.data(), B->d.b.sb + (size_t)wd.body_off * SBW, sb.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < wd.n; i++) {
        const float* c = cb.data() + (size_t)i * CBW;
        const float* s = sb.data() + (size_t)i * SBW;
        for (int k = 0; k < 3; k++) {
            if (pos) pos[(size_t)i * 3 + k] = c[CB_P + k];
            if (v) v[(size_t)i * 3 + k] = s[SB_V + k];
            if (av) av[(size_t)i * 3 + k] = s[SB_W + k];
        }
        if (rot) for (int k = 0; k < 9; k++) rot[(size_t)i * 9 + k] = c[CB_R + k];
    }
    return PE_OK;
}

int pe_batch_get_counts(pe_batch* B, int64_t* counts) {
    if (!B || !counts) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    std::vector<WorldState> st(B->nworlds);
    PE_HIP(hipMemcpyAsync(st.data(), B->d.st, st.size() * sizeof(WorldState), hipMemcpyDeviceToHost, B->stream));
    PE_HIP(hipStreamSynchronize(B->stream));
    int64_t pairs = 0, points = 0, maxcol = 0, ovf = 0;
    for (auto& s : st) { pairs += s.npairs; points += s.npoints; maxcol = std::max<int64_t>(maxcol, s.ncolors); ovf += s.overflow != 0; }
    counts[0] = B->total_bodies;
    counts[1] = B->dyn_bodies;
    counts[2] = pairs;
    counts[3] = points;
    counts[4] = maxcol;
    counts[5] = ovf;
    counts[6] = B->nworlds;
    return PE_OK;
}

}  // extern "C"

// ====================================================================== phase profile
// Runs one step with per-world phase timestamps (wall_clock64, 100 MHz on gfx9) into ts
// [nworlds][PHASES], zero where a phase was not recorded: the solver phases of a group of worlds
// are stamped on its first world only.
static int phase_step(pe_batch* B, uint64_t* h) {
    uint64_t* ts = nullptr;
    const size_t bytes = (size_t)B->nworlds * PHASES * 8;
    PE_HIP(hipMalloc(&ts, bytes));
    int rc = hipMemsetAsync(ts, 0, bytes, B->stream) == hipSuccess ? PE_OK : PE_ERR_DEVICE;
    B->d.phase_ts = ts;
    B->drop_graph();
    if (rc == PE_OK) rc = B->launch_step();
    B->d.phase_ts = nullptr;
    if (rc == PE_OK && (hipMemcpyAsync(h, ts, bytes, hipMemcpyDeviceToHost, B->stream) != hipSuccess ||
                        hipStreamSynchronize(B->stream) != hipSuccess)) {
        pe_set_error("phase profile: copy failed");
        rc = PE_ERR_DEVICE;
    }
    hipFree(ts);
    return rc;
}

extern "C" int pe_batch_phase_trace(pe_batch* B, uint64_t* ts) {
    if (!B || !ts) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    return phase_step(B, ts);
}

// The sum over worlds of each phase's duration in microseconds (phases whose two stamps exist).
extern "C" int pe_batch_phase_profile(pe_batch* B, double* us /*8*/) {
    if (!B || !us) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    std::vector<uint64_t> h((size_t)B->nworlds * PHASES);
    const int rc = phase_step(B, h.data());
    for (int k = 0; k < PHASES; k++) us[k] = 0;
    if (rc != PE_OK) return rc;
    for (int w = 0; w < B->nworlds; w++)
        for (int k = 1; k < PHASES; k++) {
            const uint64_t a = h[(size_t)w * PHASES + k - 1], b = h[(size_t)w * PHASES + k];
            if (a && b) us[k - 1] += (double)(b - a) / 100.0;
        }
    return rc;
}

// ====================================================================== per-stage times
// Runs nsteps eagerly with hipEvents between the three launch groups of each step and returns
// the summed device milliseconds of K1 (integrate + broadphase), K2 (narrowphase) and K3
// (solver setup + colouring + Gauss-Seidel).
extern "C" int pe_batch_stage_times(pe_batch* B, int32_t nsteps, double* ms /*3*/) {
    if (!B || !ms || nsteps <= 0) { pe_set_error("invalid argument"); return PE_ERR_INVALID; }
    ms[0] = ms[1] = ms[2] = 0;
    LaunchTimer tm;
    for (int s = 0; s < nsteps; s++) {
        tm.used = 0;
        if (int rc = launch_all(B->d, B->classes, B->dlist, B->dgroups, B->stream, B->side, B->has_shapes, &tm)) return rc;
        PE_HIP(hipEventSynchronize(tm.ev[tm.used - 1]));
        for (int k = 0; k < tm.used / 2; k++) {
            float t = 0;
            PE_HIP(hipEventElapsedTime(&t, tm.ev[2 * k], tm.ev[2 * k + 1]));
            ms[tm.stage[k]] += t;
        }
    }
    return PE_OK;
}

// Runs nsteps eagerly with every size class launched alone, in turn, each launch timed by its own
// event pair; ms[2c] / ms[2c + 1] are class c's summed K1 / K3 device milliseconds.  The
// per-step sums become the classes' costs, which place them on streams for concurrent steps.
extern "C" int pe_batch_class_times(pe_batch* B, int32_t nsteps, double* ms, int32_t cap) {
    const int nc = B ? (int)B->classes.size() : 0;
    if (!B || !ms || nsteps <= 0 || cap < 2 * nc) { pe_set_error("invalid argument"); return PE_ERR_INVALID; }
    for (int k = 0; k < 2 * nc; k++) ms[k] = 0;
    LaunchTimer tm;
    for (int s = 0; s < nsteps; s++) {
        tm.used = 0;
        if (int rc = launch_all(B->d, B->classes, B->dlist, B->dgroups, B->stream, B->side, B->has_shapes, &tm, true)) return rc;
        if (tm.used) PE_HIP(hipEventSynchronize(tm.ev[tm.used - 1]));
        for (int k = 0; k < tm.used / 2; k++) {
            float t = 0;
            PE_HIP(hipEventElapsedTime(&t, tm.ev[2 * k], tm.ev[2 * k + 1]));
            ms[tm.stage[k] - 3] += t;
        }
    }
    for (int c = 0; c < nc; c++) B->classes[c].cost = B->classes[c].count ? (ms[2 * c] + ms[2 * c + 1]) / nsteps : -1.0;
    schedule_classes(B->classes, B->side.n + 1);
    B->drop_graph();
    return PE_OK;
}

// ====================================================================== snapshot
extern "C" int pe_batch_snapshot(pe_batch* B) {
    if (!B) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    return B->copy_state(true);
}
extern "C" int pe_batch_restore(pe_batch* B) {
    if (!B) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    return B->copy_state(false);
}

// ====================================================================== timed batch steps
extern "C" int pe_batch_step_timed(pe_batch* B, int32_t nsteps, double* ms) {
    if (!B || !ms) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    hipEvent_t e0, e1;
    PE_HIP(hipEventCreate(&e0));
    PE_HIP(hipEventCreate(&e1));
    PE_HIP(hipEventRecord(e0, B->stream));
    int rc = pe_batch_step(B, nsteps);
    PE_HIP(hipEventRecord(e1, B->stream));
    PE_HIP(hipEventSynchronize(e1));
    float t = 0;
    PE_HIP(hipEventElapsedTime(&t, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t;
    return rc;
}

// ====================================================================== server
namespace {

constexpr int NSTAGE = NF + NI;          // staged columns per body
constexpr int NRESULT = 18;              // pos 3, rot 9, v 3, w 3

// Writes a staged [NSTAGE][n] block into the arena columns of bodies [off, off+n), then the
// derived quantities (AABB, scale, world inverse inertia).
__global__ void k_admit(BodyView b, const float* __restrict__ stage, int off, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int f = 0; f < NF; f++) b.at(f, off + i) = stage[(size_t)f * n + i];
        for (int f = 0; f < NI; f++) b.iat(f, off + i) = __float_as_int(stage[(size_t)(NF + f) * n + i]);
        refresh_body(b, off + i);
    }
}
// Packs pos/rot/v/w of bodies [off, off+n) (from their records) into out[NRESULT][n], then the
// world's capacity-error bits (int) into out[NRESULT * n].
__global__ void k_gather(BodyView b, const WorldState* __restrict__ st, int slot, float* __restrict__ out, int off, int n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[(size_t)NRESULT * n] = __int_as_float(st[slot].overflow);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* c = b.cbr(off + i);   // the records hold a served world's pose and velocities
        const float* s = b.sbr(off + i);
        for (int k = 0; k < 3; k++) out[(size_t)k * n + i] = c[CB_P + k];
        for (int k = 0; k < 9; k++) out[(size_t)(3 + k) * n + i] = c[CB_R + k];
        for (int k = 0; k < 3; k++) out[(size_t)(12 + k) * n + i] = s[SB_V + k];
        for (int k = 0; k < 3; k++) out[(size_t)(15 + k) * n + i] = s[SB_W + k];
    }
}
__global__ void k_reset_state(WorldState* st, int w) {
    WorldState z{};
    st[w] = z;
}

// First-fit allocator over [0, cap) with coalescing frees.
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;    // offset -> length
    void init(int64_t cap) { free_.clear(); if (cap > 0) free_[0] = cap; }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t off = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[off + len] = rest;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t len) {
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

struct Req {
    int64_t id;
    int n;
    int horizon;
    int done = 0;
    int slot = -1;           // world slot while live
    int64_t boff = -1, poff = -1;
    int pcap = 0;
    int64_t stage_off = 0;       // offset of its upload in the staging buffer
    std::vector<float> staged;   // [NSTAGE][n]
    HostMeshes meshes;           // its convex / concave meshes, prepared at submission
    int64_t mrec_off = -1, marena_off = -1;   // their records / arena range while live
    int64_t footprint = 0;
    bool cancelled = false;
    bool prepared = false;       // its upload is in flight (admitted at the next tick)
    std::chrono::steady_clock::time_point t_submit, t_admit;
};

int64_t req_footprint(int n) {
    const int64_t pc = body_pair_cap(n);
    return (int64_t)n * (NSTAGE + CBW + SBW + 8) * 4 + pc * (8 * (int64_t)(CW + 1 + RW) * 4 + 10 * 4) + 4 * (int64_t)(pc + 1) + 64;
}
// Device bytes of a request's prepared meshes (records and arena).
int64_t mesh_footprint(const HostMeshes& m) {
    return (int64_t)m.rec.size() * 4 + (int64_t)m.arena.size() * (int64_t)sizeof(float4);
}

}  // namespace

struct pe_server {
    pe_server_config cfg{};
    hipStream_t cs = nullptr;   // compute
    hipStream_t xs = nullptr;   // copies
    std::vector<void*> allocs;
    BatchDev d{};
    int64_t max_bodies = 0, max_pairs = 0;
    RangeAlloc bodies, pairs;
    std::vector<int> free_slots;
    std::vector<WorldDesc> hdesc;
    std::deque<Req*> queue;
    std::vector<Req*> live;           // by admission order
    std::vector<Req*> prepared;       // uploads in flight, admitted at the next tick
    std::map<int64_t, Req*> all;
    std::map<int64_t, std::vector<float>> results;   // id -> [NRESULT][n], then the capacity-error bits
    std::vector<double> wait_ms;      // per admitted request: submission -> admission
    std::vector<double> latency_ms;   // per finished request: submission -> result available
    double queue_sum = 0;             // queued requests summed over ticks (mean queue depth)
    int64_t queue_ticks = 0, queue_max = 0, cancelled_n = 0, capacity_errors = 0;
    std::vector<int64_t> finished_ids;
    int64_t next_id = 1;
    float* dstage[2] = {nullptr, nullptr};
    int64_t stage_cap = 0;
    int stage_buf = 0;
    hipEvent_t ev_upload[2] = {nullptr, nullptr};
    float* dresult = nullptr;
    int64_t result_cap = 0;
    int32_t* dlist = nullptr;
    int2* dgroups = nullptr;
    std::vector<Class> classes;
    bool dirty = true;
    int64_t live_footprint = 0;
    std::vector<double> tick_ms;
    double body_steps = 0;
    int64_t admitted = 0;
    hipEvent_t ev_done = nullptr;

    DevMeshes meshes;   // the unit cylinder mesh
    int32_t* mrec = nullptr;   // [max_meshes * MRW] mesh records of the live requests
    float4* marena = nullptr;  // [mesh_arena_cap] their vertices, faces, edges, BVH
    int64_t max_meshes = 0, mesh_arena_cap = 0;
    RangeAlloc mesh_recs, mesh_mem;
    bool has_shapes = false;   // a cylinder or a mesh was ever submitted
    ClassStreams side;
    ~pe_server() {
        if (cs) hipStreamSynchronize(cs);
        if (xs) hipStreamSynchronize(xs);
        side.release();
        for (void* p : allocs) hipFree(p);
        meshes.release();
        for (auto& e : ev_upload) if (e) hipEventDestroy(e);
        if (ev_done) hipEventDestroy(ev_done);
        if (cs) hipStreamDestroy(cs);
        if (xs) hipStreamDestroy(xs);
        for (auto& kv : all) delete kv.second;
    }
    template <class T> int alloc(T** p, size_t count) {
        void* q = nullptr;
        PE_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        PE_HIP(hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), cs));
        allocs.push_back(q);
        *p = (T*)q;
        return PE_OK;
    }

    int rebuild_lists() {
        std::vector<std::vector<int>> members(classes.size());
        for (Req* r : live) {
            for (size_t c = 0; c < classes.size(); c++)
                if (r->n >= classes[c].lo && r->n <= classes[c].hi) { members[c].push_back(r->slot); break; }
        }
        std::vector<int32_t> list;
        std::vector<int2> grp;
        std::vector<int> nb(hdesc.size());
        for (size_t w = 0; w < hdesc.size(); w++) nb[w] = hdesc[w].n;
        int nl = 0;   // the live worlds' pair capacities sum to at most the pair arena's size
        for (size_t c = 0; c < classes.size(); c++) {
            auto& m = members[c];
            std::stable_sort(m.begin(), m.end(), [&](int a, int b) { return hdesc[a].n > hdesc[b].n; });
            classes[c].count = (int)m.size();
            classes[c].nmax = m.empty() ? 0 : hdesc[m[0]].n;
            classes[c].nl_off = nl;
            classes[c].nl_cap = 0;
            for (int w : m) classes[c].nl_cap += hdesc[w].pair_cap;
            nl += classes[c].nl_cap;
            classes[c].list_off = (int)list.size();
            list.insert(list.end(), m.begin(), m.end());
            build_groups(classes[c], m, nb, false, grp);   // arena ranges are not contiguous: one world per group
        }
        if (!list.empty()) PE_HIP(hipMemcpyAsync(dlist, list.data(), list.size() * 4, hipMemcpyHostToDevice, cs));
        if (!grp.empty()) PE_HIP(hipMemcpyAsync(dgroups, grp.data(), grp.size() * sizeof(int2), hipMemcpyHostToDevice, cs));
        PE_HIP(hipStreamSynchronize(cs));  // the host list buffer is reused
        schedule_classes(classes, side.n + 1);
        dirty = false;
        return PE_OK;
    }

    // Picks queued requests that fit and starts their uploads on the copy stream.
    int prepare_admissions() {
        const int buf = stage_buf;
        int64_t used = 0;
        while (!queue.empty()) {
            Req* r = queue.front();
            const int64_t fp = r->footprint;
            const int64_t need = (int64_t)r->n * NSTAGE;
            if (free_slots.empty() || live_footprint + fp > cfg.hbm_budget_bytes || used + need > stage_cap) break;
            const int64_t bo = bodies.alloc(r->n);
            if (bo < 0) break;
            const int pc = body_pair_cap(r->n);
            const int64_t po = pairs.alloc(pc + 1);
            if (po < 0) { bodies.release(bo, r->n); break; }
            const int nm = r->meshes.count();
            const int64_t nar = (int64_t)r->meshes.arena.size();
            int64_t mo = -1, ao = -1;
            if (nm > 0) {
                mo = mesh_recs.alloc(nm);
                ao = mo < 0 ? -1 : mesh_mem.alloc(std::max<int64_t>(nar, 1));
                if (mo < 0 || ao < 0) {
                    if (mo >= 0) mesh_recs.release(mo, nm);
                    bodies.release(bo, r->n);
                    pairs.release(po, pc + 1);
                    break;
                }
                // the meshes move to their arena range, and the bodies name them by server-wide id
                r->meshes.rebase((int32_t)ao);
                float* sp = r->staged.data() + (size_t)F_SP * r->n;
                const int32_t* ty = reinterpret_cast<const int32_t*>(r->staged.data()) + (size_t)(NF + I_TYPE) * r->n;
                for (int i = 0; i < r->n; i++) {
                    if (ty[i] != SHAPE_CONVEX && ty[i] != SHAPE_CONCAVE) continue;
                    int32_t id;
                    std::memcpy(&id, sp + i, 4);
                    id += (int32_t)mo;
                    std::memcpy(sp + i, &id, 4);
                }
                PE_HIP(hipMemcpyAsync(mrec + mo * MRW, r->meshes.rec.data(), r->meshes.rec.size() * 4, hipMemcpyHostToDevice, xs));
                if (nar > 0)
                    PE_HIP(hipMemcpyAsync(marena + ao, r->meshes.arena.data(), nar * sizeof(float4), hipMemcpyHostToDevice, xs));
            }
            r->mrec_off = mo;
            r->marena_off = ao;
            queue.pop_front();
            r->boff = bo;
            r->poff = po;
            r->pcap = pc;
            r->slot = free_slots.back();
            free_slots.pop_back();
            live_footprint += fp;
            PE_HIP(hipMemcpyAsync(dstage[buf] + used, r->staged.data(), need * 4, hipMemcpyHostToDevice, xs));
            r->staged.clear();
            r->staged.shrink_to_fit();
            r->stage_off = used;
            r->prepared = true;
            prepared.push_back(r);
            used += need;
        }
        PE_HIP(hipEventRecord(ev_upload[buf], xs));
        return PE_OK;
    }

    int admit_prepared() {
        if (prepared.empty()) return PE_OK;
        const int buf = stage_buf;
        PE_HIP(hipStreamWaitEvent(cs, ev_upload[buf], 0));
        const auto now = std::chrono::steady_clock::now();
        for (Req* r : prepared) {
            r->prepared = false;
            if (r->cancelled) {   // cancelled while its upload was in flight: never stepped
                release(r);
                forget(r);
                continue;
            }
            r->t_admit = now;
            wait_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
            const int64_t soff = r->stage_off;
            // the pair range holds pcap + 1 slots: its start doubles as the colour-start region
            WorldDesc wd{(int32_t)r->boff, r->n, (int32_t)r->poff, r->pcap, (int32_t)r->poff, {0, 0, 0}};
            hdesc[r->slot] = wd;
            PE_HIP(hipMemcpyAsync((void*)(d.desc + r->slot), &hdesc[r->slot], sizeof(WorldDesc), hipMemcpyHostToDevice, cs));
            hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, cs, d.st, r->slot);
            hipLaunchKernelGGL(k_admit, dim3((r->n + 255) / 256), dim3(256), 0, cs, d.b, dstage[buf] + soff, (int)r->boff, r->n);
            live.push_back(r);
            admitted++;
        }
        PE_HIP(hipStreamSynchronize(cs));  // descriptor staging on the host is reused
        prepared.clear();
        dirty = true;
        return PE_OK;
    }

    // Drops a request that produced no result (cancelled) from the id table.
    void forget(Req* r) {
        all.erase(r->id);
        delete r;
        cancelled_n++;
    }
    void release(Req* r) {
        bodies.release(r->boff, r->n);
        pairs.release(r->poff, r->pcap + 1);
        if (r->mrec_off >= 0) {
            mesh_recs.release(r->mrec_off, r->meshes.count());
            mesh_mem.release(r->marena_off, std::max<int64_t>((int64_t)r->meshes.arena.size(), 1));
            r->mrec_off = r->marena_off = -1;
        }
        r->meshes = HostMeshes();
        free_slots.push_back(r->slot);
        live_footprint -= r->footprint;
        r->slot = -1;
    }

    int tick(int32_t* lw, int64_t* lb) {
        const auto t0 = std::chrono::steady_clock::now();
        // drop cancelled live worlds
        for (size_t i = 0; i < live.size();) {
            if (live[i]->cancelled) {
                Req* r = live[i];
                release(r);
                live.erase(live.begin() + i);
                forget(r);
                dirty = true;
            } else {
                i++;
            }
        }
        if (int rc = admit_prepared()) return rc;
        if (dirty) if (int rc = rebuild_lists()) return rc;
        int64_t nb = 0;
        for (Req* r : live) nb += r->n;
        if (!live.empty())
            if (int rc = launch_all(d, classes, dlist, dgroups, cs, side, has_shapes)) return rc;
        PE_HIP(hipEventRecord(ev_done, cs));
        // overlap: next admissions upload while this step runs
        stage_buf ^= 1;
        if (int rc = prepare_admissions()) return rc;
        PE_HIP(hipEventSynchronize(ev_done));
        body_steps += (double)nb;
        // retire every world that reached its horizon: gathers into the result buffer and one copy
        // back per round, as many rounds as the buffer needs (a world is never stepped past its
        // horizon)
        std::vector<Req*> fin;
        for (Req* r : live) if (++r->done >= r->horizon) fin.push_back(r);
        const auto t_fin = std::chrono::steady_clock::now();
        for (size_t k0 = 0; k0 < fin.size();) {
            int64_t off = 0;
            std::vector<int64_t> offs;
            size_t k1 = k0;
            for (; k1 < fin.size(); k1++) {
                Req* r = fin[k1];
                const int64_t need = (int64_t)r->n * NRESULT + 1;
                if (off + need > result_cap) break;   // result_cap holds at least one largest world
                hipLaunchKernelGGL(k_gather, dim3((r->n + 255) / 256), dim3(256), 0, cs, d.b, d.st, r->slot, dresult + off,
                                   (int)r->boff, r->n);
                offs.push_back(off);
                off += need;
            }
            std::vector<float> host(off);
            PE_HIP(hipMemcpyAsync(host.data(), dresult, off * 4, hipMemcpyDeviceToHost, cs));
            PE_HIP(hipStreamSynchronize(cs));
            for (size_t k = k0; k < k1; k++) {
                Req* r = fin[k];
                const int64_t o = offs[k - k0], len = (int64_t)r->n * NRESULT + 1;
                results[r->id] = std::vector<float>(host.begin() + o, host.begin() + o + len);
                finished_ids.push_back(r->id);
                latency_ms.push_back(std::chrono::duration<double, std::milli>(t_fin - r->t_submit).count());
                release(r);
                live.erase(std::find(live.begin(), live.end(), r));
                dirty = true;
            }
            k0 = k1;
        }
        queue_sum += (double)queue.size();
        queue_ticks++;
        queue_max = std::max<int64_t>(queue_max, (int64_t)queue.size());
        const auto t1 = std::chrono::steady_clock::now();
        if (nb > 0) tick_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
        if (lw) *lw = (int32_t)live.size();
        if (lb) {
            int64_t s = 0;
            for (Req* r : live) s += r->n;
            *lb = s;
        }
        return PE_OK;
    }
};

extern "C" {

int pe_server_footprint(const pe_scene_desc* scene, int64_t* bytes) {
    if (!scene || !bytes) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    if (int rc = check_scene(scene)) return rc;
    std::vector<float> hf((size_t)NF * scene->n, 0.0f);
    std::vector<int32_t> hi((size_t)NI * scene->n, 0);
    HostMeshes hm;
    if (int rc = stage_bodies(scene, hf, hi, (size_t)scene->n, 0, hm)) return rc;
    *bytes = req_footprint(scene->n) + mesh_footprint(hm);
    return PE_OK;
}

int pe_server_create(const pe_server_config* cfg, pe_server** out) {
    if (!cfg || !out || cfg->max_worlds <= 0 || cfg->max_world_bodies <= 0 ||
        cfg->max_world_bodies > PE_BATCH_MAX_WORLD_BODIES || cfg->hbm_budget_bytes <= 0) {
        pe_set_error("pe_server_create: invalid config");
        return PE_ERR_INVALID;
    }
    set_lds_limits();
    auto* S = new pe_server();
    auto fail = [&](int rc) { delete S; return rc; };
    S->cfg = *cfg;
    if (hipStreamCreateWithFlags(&S->cs, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S->xs, hipStreamNonBlocking) != hipSuccess || S->side.create() != PE_OK) {
        pe_set_error("streams");
        return fail(PE_ERR_DEVICE);
    }
    // arena sizes from the budget: per body its store and 8 pair slots of contacts and rows
    const int64_t per_body = req_footprint(1024) / 1024 + 1;
    S->stage_cap = (int64_t)cfg->max_world_bodies * NSTAGE * 8;            // floats per staging buffer
    S->result_cap = ((int64_t)cfg->max_world_bodies * NRESULT + 1) * 64;  // floats: 64 largest worlds
    // a sixteenth of the budget (at most 8 GiB) holds the requests' meshes: records for one mesh per KiB
    S->mesh_arena_cap = std::min<int64_t>(cfg->hbm_budget_bytes / 16, (int64_t)8 << 30) / (int64_t)sizeof(float4);
    S->mesh_arena_cap = std::min<int64_t>(std::max<int64_t>(S->mesh_arena_cap, 4096), ((int64_t)1 << 31) - 1);
    S->max_meshes = std::min<int64_t>(std::max<int64_t>(S->mesh_arena_cap / 64, 64), (int64_t)1 << 24);
    const int64_t fixed = 2 * S->stage_cap * 4 + S->result_cap * 4 + (int64_t)cfg->max_worlds * 64 +
                          S->mesh_arena_cap * (int64_t)sizeof(float4) + S->max_meshes * MRW * 4;
    S->max_bodies = std::max<int64_t>((cfg->hbm_budget_bytes - fixed) / per_body, cfg->max_world_bodies);
    S->max_bodies = std::min<int64_t>(S->max_bodies, (int64_t)1 << 28);
    S->max_pairs = 8 * S->max_bodies + 65 * (int64_t)cfg->max_worlds;
    if (S->max_pairs > ((int64_t)1 << 31) / 8 - 1) { pe_set_error("budget too large for 32-bit pair indices"); return fail(PE_ERR_INVALID); }
    BatchDev& d = S->d;
    d.dt = (float)cfg->params.dt;
    for (int k = 0; k < 3; k++) d.g[k] = (float)cfg->params.gravity[k];
    d.iterations = cfg->params.iterations;
    d.b.n = (int)S->max_bodies;
    d.b.stride = (int)((S->max_bodies + 63) & ~63);
    d.ps.cs = (size_t)S->max_pairs * 8;
    const size_t SB = d.b.stride;
    const int64_t P = S->max_pairs;
    int rc = PE_OK;
    rc |= S->alloc(&d.b.f, NF * SB);
    rc |= S->alloc(&d.b.i, NI * SB);
    rc |= S->alloc(&d.b.cb, (size_t)CBW * SB);
    rc |= S->alloc(&d.b.sb, (size_t)SBW * SB);
    rc |= S->alloc(&d.b.ab, (size_t)8 * SB);
    rc |= S->alloc(&d.ps.pa, P);
    rc |= S->alloc(&d.ps.pb, P);
    rc |= S->alloc(&d.ps.cp, (size_t)CW * P * 8);
    rc |= S->alloc(&d.ps.cd, (size_t)P * 8);
    rc |= S->alloc(&d.ps.npts, P);
    rc |= S->alloc(&d.ps.swap, P);
    rc |= S->alloc(&d.ps.rows, (size_t)RW * P * 8);
    rc |= S->alloc(&d.act, 3 * (size_t)P);
    rc |= S->alloc(&d.order, P);
    rc |= S->alloc(&d.cstart, P + cfg->max_worlds);
    rc |= S->alloc(&d.nl_pair, P);
    rc |= S->alloc(&d.nl_limit, P);
    rc |= S->alloc(&d.nl_shape, P);
    rc |= S->alloc(&d.nl_count, 3 * MAX_CLASSES);
    rc |= S->alloc(&d.ovf_any, 1);
    WorldDesc* ddesc = nullptr;
    WorldState* dst = nullptr;
    rc |= S->alloc(&ddesc, cfg->max_worlds);
    rc |= S->alloc(&dst, cfg->max_worlds);
    rc |= S->alloc(&S->dstage[0], S->stage_cap);
    rc |= S->alloc(&S->dstage[1], S->stage_cap);
    rc |= S->alloc(&S->dresult, S->result_cap);
    rc |= S->alloc(&S->dlist, cfg->max_worlds);
    rc |= S->alloc(&S->dgroups, cfg->max_worlds);
    rc |= S->alloc(&S->mrec, (size_t)S->max_meshes * MRW);
    rc |= S->alloc(&S->marena, (size_t)S->mesh_arena_cap);
    if (rc != PE_OK) return fail(PE_ERR_DEVICE);
    {
        HostMeshes none;
        if (int rc2 = S->meshes.upload(none, S->cs)) return fail(rc2);
        d.b.ms = MeshSet{S->mrec, S->marena, S->meshes.cyl};
    }
    S->mesh_recs.init(S->max_meshes);
    S->mesh_mem.init(S->mesh_arena_cap);
    d.desc = ddesc;
    d.st = dst;
    S->bodies.init(S->max_bodies);
    S->pairs.init(S->max_pairs);
    S->hdesc.assign(cfg->max_worlds, WorldDesc{0, 0, 0, 0, 0, {0, 0, 0}});
    for (int w = cfg->max_worlds - 1; w >= 0; w--) S->free_slots.push_back(w);
    S->classes = default_classes();
    for (auto& e : S->ev_upload) if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(PE_ERR_DEVICE);
    if (hipEventCreateWithFlags(&S->ev_done, hipEventDisableTiming) != hipSuccess) return fail(PE_ERR_DEVICE);
    if (hipStreamSynchronize(S->cs) != hipSuccess) { pe_set_error("init sync"); return fail(PE_ERR_DEVICE); }
    *out = S;
    return PE_OK;
}

int pe_server_destroy(pe_server* S) {
    delete S;
    return PE_OK;
}

int pe_server_submit(pe_server* S, const pe_scene_desc* scene, int32_t horizon, int64_t* id) {
    if (!S || !scene || horizon <= 0) { pe_set_error("invalid argument"); return PE_ERR_INVALID; }
    if (int rc = check_batched_scene(scene)) return rc;
    if (scene->n > S->cfg.max_world_bodies) {
        pe_set_error("request larger than the server admits");
        return PE_ERR_BUSY;
    }
    Req* r = new Req();
    r->n = scene->n;
    r->horizon = horizon;
    r->staged.assign((size_t)NSTAGE * scene->n, 0.0f);
    std::vector<float> hf((size_t)NF * scene->n, 0.0f);
    std::vector<int32_t> hi((size_t)NI * scene->n, 0);
    if (int rc = stage_bodies(scene, hf, hi, (size_t)scene->n, 0, r->meshes)) { delete r; return rc; }
    r->footprint = req_footprint(scene->n) + mesh_footprint(r->meshes);
    if (r->footprint > S->cfg.hbm_budget_bytes || r->meshes.count() > S->max_meshes ||
        (int64_t)r->meshes.arena.size() > S->mesh_arena_cap) {
        delete r;
        pe_set_error("request larger than the server admits");
        return PE_ERR_BUSY;
    }
    r->id = S->next_id++;
    r->t_submit = std::chrono::steady_clock::now();
    for (int i = 0; i < scene->n; i++) S->has_shapes |= scene->shape_type[i] > PE_SHAPE_SPHERE;
    std::copy(hf.begin(), hf.end(), r->staged.begin());
    std::memcpy(r->staged.data() + hf.size(), hi.data(), hi.size() * 4);
    S->all[r->id] = r;
    S->queue.push_back(r);
    if (id) *id = r->id;
    return PE_OK;
}

int pe_server_tick(pe_server* S, int32_t* live_worlds, int64_t* live_bodies) {
    if (!S) { pe_set_error("null server"); return PE_ERR_INVALID; }
    if (S->live.empty() && S->prepared.empty()) {
        // nothing on the device yet: start uploads now so they admit at this tick
        if (int rc = S->prepare_admissions()) return rc;
    }
    return S->tick(live_worlds, live_bodies);
}

int pe_server_poll(pe_server* S, int64_t* ids, int32_t cap, int32_t* count) {
    if (!S || !count) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    const int m = std::min<int>(cap, (int)S->finished_ids.size());
    for (int i = 0; i < m; i++) ids[i] = S->finished_ids[i];
    S->finished_ids.erase(S->finished_ids.begin(), S->finished_ids.begin() + m);
    *count = m;
    return PE_OK;
}

int pe_server_result(pe_server* S, int64_t id, float* pos, float* rot, float* v, float* w) {
    if (!S) { pe_set_error("null server"); return PE_ERR_INVALID; }
    auto it = S->results.find(id);
    if (it == S->results.end()) { pe_set_error("no result for this id"); return PE_ERR_NOT_FOUND; }
    const auto& r = it->second;
    const int n = (int)((r.size() - 1) / NRESULT);
    int32_t ovf;
    std::memcpy(&ovf, &r[(size_t)NRESULT * n], 4);
    if (ovf == 0) {
        for (int i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++) if (pos) pos[3 * i + k] = r[(size_t)k * n + i];
            for (int k = 0; k < 9; k++) if (rot) rot[9 * i + k] = r[(size_t)(3 + k) * n + i];
            for (int k = 0; k < 3; k++) if (v) v[3 * i + k] = r[(size_t)(12 + k) * n + i];
            for (int k = 0; k < 3; k++) if (w) w[3 * i + k] = r[(size_t)(15 + k) * n + i];
        }
    }
    S->results.erase(it);
    auto a = S->all.find(id);
    if (a != S->all.end()) { delete a->second; S->all.erase(a); }
    if (ovf != 0) {
        S->capacity_errors++;
        pe_set_error(std::string("request ") + std::to_string(id) + " exceeded a capacity while stepping:" +
                     ((ovf & WOVF_PAIRS) ? " pairs beyond its pair capacity (16 per body + 64)" : "") +
                     ((ovf & WOVF_CLIP) ? " a clipped contact polygon outgrew its buffer" : "") + "; no result");
        return PE_ERR_CAPACITY;
    }
    return PE_OK;
}

int pe_server_cancel(pe_server* S, int64_t id) {
    if (!S) { pe_set_error("null server"); return PE_ERR_INVALID; }
    auto it = S->all.find(id);
    if (it == S->all.end()) { pe_set_error("unknown request"); return PE_ERR_NOT_FOUND; }
    Req* r = it->second;
    if (S->results.count(id)) { pe_set_error("request already finished"); return PE_ERR_INVALID; }
    auto q = std::find(S->queue.begin(), S->queue.end(), r);
    if (q != S->queue.end()) {   // still queued: it holds nothing on the device
        S->queue.erase(q);
        S->forget(r);
        return PE_OK;
    }
    r->cancelled = true;   // live or uploading: released at the next tick
    return PE_OK;
}

int pe_server_stats(pe_server* S, double* out) {
    if (!S || !out) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    std::vector<double> t = S->tick_ms;
    std::sort(t.begin(), t.end());
    auto pct = [&](double q) { return t.empty() ? 0.0 : t[std::min(t.size() - 1, (size_t)std::floor(q * (t.size() - 1) + 0.5))]; };
    out[0] = (double)S->tick_ms.size();
    out[1] = pct(0.50);
    out[2] = pct(0.99);
    out[3] = t.empty() ? 0.0 : std::accumulate(t.begin(), t.end(), 0.0) / t.size();
    out[4] = S->body_steps;
    out[5] = (double)S->live.size();
    out[6] = (double)S->queue.size();
    out[7] = (double)S->admitted;
    return PE_OK;
}

int pe_server_request_stats(pe_server* S, double* out) {
    if (!S || !out) { pe_set_error("null argument"); return PE_ERR_INVALID; }
    auto pct = [](std::vector<double> t, double q) {
        if (t.empty()) return 0.0;
        std::sort(t.begin(), t.end());
        return t[std::min(t.size() - 1, (size_t)std::floor(q * (t.size() - 1) + 0.5))];
    };
    out[0] = pct(S->wait_ms, 0.50);
    out[1] = pct(S->wait_ms, 0.99);
    out[2] = pct(S->latency_ms, 0.50);
    out[3] = pct(S->latency_ms, 0.99);
    out[4] = S->queue_ticks ? S->queue_sum / (double)S->queue_ticks : 0.0;
    out[5] = (double)S->queue_max;
    out[6] = (double)S->latency_ms.size();
    out[7] = (double)S->cancelled_n;
    out[8] = (double)S->capacity_errors;
    return PE_OK;
}

}  // extern "C"
