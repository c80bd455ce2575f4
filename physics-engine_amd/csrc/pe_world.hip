// Single-world pipeline (the plain drop-in for one World::step) and its C ABI.
//
// Per step, on one HIP stream, no host synchronisation:
//   k_integrate        World::updateObjectStatus + applyExternalForce    (thread / body)
//   k_sap_*            sweep-and-prune broadphase over the body store (pe_sap.hpp): stable
//                      radix sort on aabb_min[axis], candidate runs, load-balanced sweep in the
//                      reference's (i, j) order, next axis
//   k_narrow           thread / pair: shape dispatch + 8-slot contact cache + depth sort
//   k_setup            thread / pair: solver rows (FrictionContactConstraint::initSequentialImpulse)
//   k_colour           one 1024-thread workgroup: constraint-graph colouring (claims in LDS)
//   k_gs_grid          cooperative launch over the whole GPU: Gauss-Seidel sweeps colour by
//                      colour with a grid barrier between colours, then velocity sync
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_prims.hpp"
#include "pe_pipeline.hpp"
#include "pe_sap.hpp"
#include "pe_joints.hpp"
#include "pe_raycast.hpp"
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
constexpr int SOLVE_THREADS = 1024;
constexpr int CLAIM_LDS = 12288;     // bodies whose colouring claims fit in LDS (48 KiB)

struct Dev {
    BodyView b;
    PairStore ps;           // pairs [pcap], contact slots and rows [pcap*8]
    SapView sap;            // the broadphase over the body store's AABBs (pairs to ps.pa / ps.pb)
    int32_t* shape_list;    // [pcap] pairs handed from k_narrow to k_narrow_shapes (from the front)
                            // and to k_narrow_sat (from the back)
    int32_t* comp_list;     // [pcap] pairs with a compound body (worlds with compounds; else null)
    // colouring by rounds: per-body degree offsets [n + 1] and fill cursors [n], adjacency [2 pcap],
    // colour per pair [pcap], pairs per colour [pcap + 1]
    uint32_t* jp_deg;
    int32_t* jp_fill;
    int2* jp_adj;        // per dynamic body, its active pairs: (pair, claim priority)
    int32_t* jp_col;
    int32_t* jp_ccount;
    int32_t* order;         // [pcap]
    int32_t* act;           // [3*pcap] colouring work lists
    int32_t* color_start;   // [pcap+1]
    int32_t* claim;         // [n] (global claims when n > CLAIM_LDS)
    DevCounts* cnt;
    int pcap;
    float dt;
    float g[3];
    int iterations;
    // joints (World::addConstraint): definitions, per-step rows, sweep order grouped by colour
    const JointDef* joints;
    JointRow* jrows;
    const int32_t* jorder;   // [njoints] joint indices, colour by colour
    const int32_t* jcs;      // [njcol + 1] colour starts in jorder
    int njoints, njcol;
    // sleeping, lifetimes, removals (World::updateObjectStatus) and collision reports
    StatusView st;
    int status_on;           // some body may sleep, end its lifetime or has been removed
    int report_on;           // some body reports its contacts
    float* events;           // [pcap][EVW] the last narrowphase's collision reports, by pair
    int32_t* ev_flag;        // [pcap] 1: pair p has a report
};

// Collision report of a pair (World::execCollisionCallbacks, world.cpp:129-192): bodies A and B (int
// bits), mean contact position, normalised sum of the normals, mean distance, mean relative velocity.
constexpr int EVW = 12;

// ------------------------------------------------------------------ stage 0: integrate
// World::updateObjectStatus (world.cpp:35-119) with zero sleep thresholds, then
// World::applyExternalForce (world.cpp:121-127), then SequentialImpulseSolver::setupSolver's
// clearTempVelocity (sequential_impulse_solver.cpp:17-20).
__global__ void k_integrate(Dev d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) integrate_body(d.b, i, d.dt, d.g);
}
// The same with sleeping and lifetimes (world.cpp:85-117).
__global__ void k_integrate_status(Dev d) {
    const StatusView st = d.st;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x)
        integrate_body<true>(d.b, i, d.dt, d.g, &st);
}

// DefaultRaycastSolver::performRaycast (default_raycast_solver.cpp:11-86), one 64-lane workgroup per
// ray over the world's bodies.  Skipped: removed bodies, ignore-collision bodies, the call's ignore
// set (sorted).  The reference walks the bodies in order and ends the query, reporting no hit, at the
// first mesh body whose AABB the unbounded ray misses, its result keeping the best hit of the bodies
// before it; so the first pass finds that body (the lowest such index), the second takes the nearest
// hit among the bodies before it, ties to the lowest index (the reference replaces a hit only by a
// strictly nearer one), as a minimum over (distance key, index).
constexpr int RAY_IN = 7;    // start (3), direction (3), length
constexpr int RAY_OUT = 12;  // hit, body (int bits), distance, point (3), normal (3), pad
__global__ void __launch_bounds__(64) k_raycast(Dev d, const float* rays, const int32_t* ignore, int nignore, float* out) {
    __shared__ int s_abort[64];
    __shared__ uint64_t s_best[64];
    const int lane = threadIdx.x;
    const float* r = rays + (size_t)blockIdx.x * RAY_IN;
    const f3 start{r[0], r[1], r[2]}, dir{r[3], r[4], r[5]};
    const float len = r[6];
    const int n = d.b.n;
    auto skipped = [&](int i) {
        if (d.status_on && (d.st.flags[i] & ST_REMOVED)) return true;
        if (d.b.at(F_SP + 3, i) == COLL_IGNORE) return true;
        int lo = 0, hi = nignore;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ignore[mid] < i) lo = mid + 1; else hi = mid;
        }
        return lo < nignore && ignore[lo] == i;
    };
    int abort = 0x7fffffff;
    for (int i = lane; i < n; i += 64) {
        if (d.b.iat(I_TYPE, i) < SHAPE_CONVEX || skipped(i)) continue;
        float t;
        f3 hp, hn;
        if (!ray_hit_box(start, dir, ld3(d.b, F_AMIN, i), ld3(d.b, F_AMAX, i), t, hp, hn)) { abort = i; break; }
    }
    s_abort[lane] = abort;
    __syncthreads();
    for (int k = 0; k < 64; k++) abort = s_abort[k] < abort ? s_abort[k] : abort;
    const int lim = n < abort ? n : abort;
    uint64_t best = ~0ull;
    float bd = REAL_MAX;
    f3 bp{0.0f, 0.0f, 0.0f}, bn{0.0f, 0.0f, 0.0f};
    for (int i = lane; i < lim; i += 64) {
        if (skipped(i)) continue;
        const int type = d.b.iat(I_TYPE, i);
        const f3 o = ld3(d.b, F_POS, i);
        const m33 R = ld9(d.b, F_ROT, i);
        const float sp[3] = {d.b.at(F_SP, i), d.b.at(F_SP + 1, i), d.b.at(F_SP + 2, i)};
        float t = 0.0f;
        f3 hp, hn;
        bool ret;
        if (type == SHAPE_BOX) ret = ray_box(start, dir, len, sp, R, o, t, hp, hn);
        else if (type == SHAPE_SPHERE) ret = ray_sphere(start, dir, len, sp[0], o, t, hp, hn);
        else if (type == SHAPE_CYLINDER) ret = ray_cylinder(start, dir, len, sp[0], sp[1], R, o, t, hp, hn);
        else if (type != SHAPE_COMPOUND) ret = ray_mesh(d.b.ms, __float_as_int(sp[0]), start, dir, len, R, o, t, hp, hn);
        else {
            // each child in turn writes the same outputs, its hit or-ed in: the last child's distance and
            // point decide (default_raycast_solver.cpp:50-78)
            ret = false;
            const int2 kr = d.b.ms.krange[__float_as_int(sp[0])];
            for (int k = 0; k < kr.y; k++) {
                float csp[3];
                m33 Rc;
                f3 pc;
                const int ct = child_pose(d.b.ms, kr.x + k, R, o, csp, Rc, pc);
                if (ct == SHAPE_BOX) ret |= ray_box(start, dir, len, csp, Rc, pc, t, hp, hn);
                else if (ct == SHAPE_SPHERE) ret |= ray_sphere(start, dir, len, csp[0], pc, t, hp, hn);
                else if (ct == SHAPE_CYLINDER) ret |= ray_cylinder(start, dir, len, csp[0], csp[1], Rc, pc, t, hp, hn);
                else ret |= ray_mesh(d.b.ms, __float_as_int(csp[0]), start, dir, len, Rc, pc, t, hp, hn);
            }
        }
        if (ret && t <= len && t < REAL_MAX) {
            const uint64_t key = ((uint64_t)float_key(t + 0.0f) << 32) | (uint32_t)i;
            if (key < best) { best = key; bd = t; bp = hp; bn = hn; }
        }
    }
    __syncthreads();
    s_best[lane] = best;
    __syncthreads();
    uint64_t win = ~0ull;
    for (int k = 0; k < 64; k++) win = s_best[k] < win ? s_best[k] : win;
    float* o = out + (size_t)blockIdx.x * RAY_OUT;
    if (win == ~0ull) {
        if (lane == 0) {
            const float e[RAY_OUT] = {0.0f, __int_as_float(-1), REAL_MAX, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            st_rec<RAY_OUT / 4>(o, e);
        }
    } else if (best == win) {
        const float e[RAY_OUT] = {abort == 0x7fffffff ? 1.0f : 0.0f, __int_as_float((int)(uint32_t)win), bd,
                                  bp.x, bp.y, bp.z, bn.x, bn.y, bn.z, 0.0f, 0.0f, 0.0f};
        st_rec<RAY_OUT / 4>(o, e);
    }
}

// Recomputes AABBs and scales only (after pe_world_set_state); a removed body stays out of the sweep.
__global__ void k_aabb_only(Dev d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.b.n; i += gridDim.x * blockDim.x) {
        if (d.status_on && (d.st.flags[i] & ST_REMOVED)) retire_body(d.b, i);
        else refresh_body(d.b, i);
    }
}

// World::removeRigidBody (world.cpp:223-231) of body i.
__global__ void k_remove_body(Dev d, int i) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    d.st.flags[i] |= ST_REMOVED;
    retire_body(d.b, i);
}

// SimpleNarrowPhase's neighbour counts (simple_narrow_phase.cpp:43-62): for a pair with contact
// points, each non-kinematic body counts a static neighbour when the other sleeps or is kinematic,
// a dynamic one otherwise (integer atomics: the counts do not depend on the order).
__global__ void k_neighbours(Dev d) {
    const int np = d.cnt->npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        if (d.ps.npts[p] == 0) continue;
        const int a = d.ps.pa[p], b = d.ps.pb[p];
        const bool ka = d.b.iat(I_KIN, a) != 0, kb = d.b.iat(I_KIN, b) != 0;
        const bool sa = (d.st.flags[a] & ST_SLEEP) != 0, sb = (d.st.flags[b] & ST_SLEEP) != 0;
        if (!kb) atomicAdd((sa || ka) ? &d.st.scnt[b] : &d.st.dcnt[b], 1);
        if (!ka) atomicAdd((sb || kb) ? &d.st.scnt[a] : &d.st.dcnt[a], 1);
    }
}

// World::execCollisionCallbacks' arithmetic (world.cpp:131-151) for every pair with contact points
// whose bodies report: sums over the points in depth order, then the means and the normalised normal
// (Eigen's normalize: divide by the square root of the squared norm when it is positive).  Velocities
// are RigidBody::getLinearVelocityAtLocalPoint of the points' local positions (rigidbody.cpp:138-140)
// after gravity, as the callbacks see them.
__global__ void k_events(Dev d) {
    const int np = d.cnt->npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const int n = d.ps.npts[p];
        const int a = d.ps.pa[p], b = d.ps.pb[p];
        int flag = 0;
        if (n > 0 && (d.st.report[a] || d.st.report[b])) {
            flag = 1;
            f3 pos{0.0f, 0.0f, 0.0f}, nor{0.0f, 0.0f, 0.0f}, vel{0.0f, 0.0f, 0.0f};
            float depth = 0.0f;
            const f3 va = ld3(d.b, F_V, a), wa = ld3(d.b, F_W, a), vb = ld3(d.b, F_V, b), wb = ld3(d.b, F_W, b);
            for (int k = 0; k < n; k++) {
                const float* c = d.ps.cp + ((size_t)p * 8 + k) * CW;
                pos = add(pos, f3{c[C_WPOS], c[C_WPOS + 1], c[C_WPOS + 2]});
                nor = add(nor, f3{c[C_N], c[C_N + 1], c[C_N + 2]});
                depth += d.ps.cd[(size_t)p * 8 + k];
                vel = add(vel, add(va, cross(wa, f3{c[C_LA], c[C_LA + 1], c[C_LA + 2]})));
                vel = sub(vel, add(vb, cross(wb, f3{c[C_LB], c[C_LB + 1], c[C_LB + 2]})));
            }
            const float fn = (float)n;
            pos = f3{pos.x / fn, pos.y / fn, pos.z / fn};
            const float z = sqnorm3(nor);
            if (z > 0.0f) {
                const float r = sqrtf(z);
                nor = f3{nor.x / r, nor.y / r, nor.z / r};
            }
            depth = depth / fn;
            vel = f3{vel.x / fn, vel.y / fn, vel.z / fn};
            const float e[EVW] = {__int_as_float(a), __int_as_float(b), pos.x, pos.y, pos.z, nor.x, nor.y, nor.z,
                                  depth, vel.x, vel.y, vel.z};
            st_rec<EVW / 4>(d.events + (size_t)p * EVW, e);
        }
        d.ev_flag[p] = flag;
    }
}

// ------------------------------------------------------------------ stage 1: broadphase (pe_sap.hpp)

// ------------------------------------------------------------------ stage 2: narrowphase
// Box and sphere pairs here; pairs with a cylinder or a mesh are listed for k_narrow_sat (a wave
// each, pair_on_wave) or k_narrow_shapes (a thread each: a box or sphere with a cylinder).
__global__ void __launch_bounds__(128) k_narrow(Dev d) {
    const int np = d.cnt->npairs;
    const int prev = d.cnt->prev_npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const bool simple = pair_is_simple(d.b, d.ps, p);
        const bool comp = !simple && (d.b.iat(I_TYPE, d.ps.pa[p]) == SHAPE_COMPOUND ||
                                      d.b.iat(I_TYPE, d.ps.pb[p]) == SHAPE_COMPOUND);
        const bool sat = !simple && !comp && pair_on_wave(d.b, d.ps, p);
        const int slot = wave_append(!simple && !sat && !comp, &d.cnt->nshape);
        const int sslot = wave_append(sat, &d.cnt->nsat);
        const int cslot = wave_append(comp && d.comp_list, &d.cnt->ncomp);
        if (simple) narrow_pair<false>(d.b, d.ps, p, prev);
        else if (sat) d.shape_list[d.pcap - 1 - sslot] = p;
        else if (comp) {
            if (d.comp_list) d.comp_list[cslot] = p;
            else { d.ps.npts[p] = 0; d.ps.swap[p] = 0; }   // no compound table: no algorithm
        } else d.shape_list[slot] = p;
    }
}
__global__ void __launch_bounds__(64) k_narrow_shapes(Dev d) {
    const int ns = d.cnt->nshape;
    const int prev = d.cnt->prev_npairs;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ns; q += gridDim.x * blockDim.x)
        narrow_pair<true>(d.b, d.ps, d.shape_list[q], prev);
}
// One wave per listed wave pair (pair_on_wave: separating-axis pairs, a sphere or cylinder with a
// mesh).  Built for three waves per SIMD (at most 168 VGPRs; a little spilling): the 50k shape mix
// took 4.16 ms per step at two waves (248 VGPRs), 4.12 at three, 4.26 at four.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_narrow_sat(Dev d) {
    const int ns = d.cnt->nsat;
    const int prev = d.cnt->prev_npairs;
    const int nw = gridDim.x * (blockDim.x / WAVE);
    for (int q = blockIdx.x * (blockDim.x / WAVE) + (int)(threadIdx.x / WAVE); q < ns; q += nw)
        narrow_pair<true, WAVE>(d.b, d.ps, d.shape_list[d.pcap - 1 - q], prev);
}
// The pairs with a compound body, one thread each (64-thread workgroups: the contact cache and the
// BVH stacks of the separating-axis and mesh routines live in LDS rows of one lane each).
__global__ void __launch_bounds__(64) k_narrow_compound(Dev d) {
    const int nc = d.cnt->ncomp;
    const int prev = d.cnt->prev_npairs;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += gridDim.x * blockDim.x)
        narrow_pair<true, 1, true>(d.b, d.ps, d.comp_list[q], prev);
}
__global__ void k_narrow_finish(Dev d) {
    d.cnt->prev_npairs = d.cnt->npairs;
    d.cnt->nshape = 0;
    d.cnt->nsat = 0;
    d.cnt->ncomp = 0;
    d.cnt->npoints = 0;   // summed by k_count_points next
}

// ------------------------------------------------------------------ stage 3: solver
// SequentialImpulseSolver::setupSolver (sequential_impulse_solver.cpp:11-58), thread / pair.
__global__ void k_setup(Dev d) {
    const int np = d.cnt->npairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) setup_pair(d.b, d.ps, p, d.dt);
}

// BallJointConstraint / HingeJointConstraint / SliderJointConstraint::initSequentialImpulse, thread / joint
// (after the contacts' rows, as setupSolver initialises them, sequential_impulse_solver.cpp:49-51).
__global__ void k_joint_setup(Dev d) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d.njoints; k += gridDim.x * blockDim.x)
        joint_setup(d.b, d.joints[k], d.jrows[k], d.dt);
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


// Sense-reversing barrier over all workgroups of a cooperative launch (every workgroup is
// resident, so every wait ends).  The last arrival resets the count and bumps the generation.
// It also orders memory across the grid: the workgroup's writes before the barrier are visible to
// every workgroup after it.  One thread per workgroup fences (a release fence before arriving, an
// acquire fence after the wait: the workgroup barriers around them carry the ordering to its other
// threads), and the arrival and the polling are relaxed atomics, so the caches are written back and
// invalidated once per workgroup and barrier, not once per wave or poll.
__device__ __forceinline__ void grid_barrier(uint32_t* count, uint32_t* gen, uint32_t nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // also keeps the load of g before the arrival
        if (__hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nblocks - 1) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);   // after the reset
        } else {
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Constraint-graph colouring by rounds (Jones-Plassmann), the same colour classes as colour_pairs:
// colour_pairs builds each colour as the independent set its priority claims settle on, which is the
// greedy independent set in priority order of the pairs left (the candidate of lowest priority always
// holds its claims, and a pair joins exactly when no lower-priority neighbour joined); repeating that
// colour after colour gives each pair the first colour no lower-priority neighbour (a pair sharing a
// dynamic body) holds.  Here a pair takes that colour as soon as all its lower-priority neighbours
// have theirs, in rounds over the whole GPU, and the colours read are final when seen, so the result
// does not depend on timing.
__global__ void k_jp_init(Dev d) {
    const int n = d.b.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        d.jp_deg[i] = 0;
        if (i < n) d.jp_fill[i] = 0;
    }
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= d.pcap; c += gridDim.x * blockDim.x) d.jp_ccount[c] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.cnt->nact = 0;
        d.cnt->ncolors = 0;
        d.cnt->jp_rem[0] = 0;
        d.cnt->jp_rem[1] = 0;
        d.cnt->jp_rem[2] = 0;
    }
}
// The pairs to colour (contact points, not both bodies kinematic) and their dynamic bodies' degrees.
__global__ void k_jp_active(Dev d) {
    const int np = d.cnt->npairs;
    const int n_iter = (np + (int)(gridDim.x * blockDim.x) - 1) / (int)(gridDim.x * blockDim.x);
    for (int it = 0; it < n_iter; it++) {
        const int p = (it * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
        bool act = false;
        int a = 0, c = 0;
        bool ka = true, kc = true;
        if (p < np && d.ps.npts[p] > 0) {
            a = d.ps.pa[p];
            c = d.ps.pb[p];
            ka = d.b.iat(I_KIN, a) != 0;
            kc = d.b.iat(I_KIN, c) != 0;
            act = !(ka && kc);
        }
        const int slot = wave_append(act, &d.cnt->nact);
        if (act) {
            d.act[slot] = p;
            d.jp_col[p] = -1;
            if (!ka) atomicAdd(&d.jp_deg[a], 1u);
            if (!kc) atomicAdd(&d.jp_deg[c], 1u);
        }
    }
}
__global__ void k_jp_fill(Dev d) {
    const int nact = d.cnt->nact;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nact; k += gridDim.x * blockDim.x) {
        const int p = d.act[k];
        const int a = d.ps.pa[p], c = d.ps.pb[p];
        const int2 e = make_int2(p, colour::prio(p, d.ps.npts[p]));
        if (!d.b.iat(I_KIN, a)) d.jp_adj[d.jp_deg[a] + atomicAdd(&d.jp_fill[a], 1)] = e;
        if (!d.b.iat(I_KIN, c)) d.jp_adj[d.jp_deg[c] + atomicAdd(&d.jp_fill[c], 1)] = e;
    }
}
constexpr int JP_THREADS = 256;
__device__ __forceinline__ int jp_col_load(const int32_t* col, int q) {
    return __hip_atomic_load(&col[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Pair p's first-fit colour if every lower-priority neighbour is coloured, else -1.  The adjacency
// entries and their colours are loaded eight at a time (independent loads in flight together: a
// neighbour scan is a chain of L2 round trips otherwise).
constexpr int JP_BATCH = 8;
__device__ int jp_try(const Dev& d, int p) {
    const int pr = colour::prio(p, d.ps.npts[p]);
    const int body[2] = {d.ps.pa[p], d.ps.pb[p]};
    uint64_t used = 0;
    bool big = false;
    for (int s = 0; s < 2; s++) {
        const int x = body[s];
        if (d.b.iat(I_KIN, x)) continue;
        const int k0 = (int)d.jp_deg[x], k1 = (int)d.jp_deg[x + 1];
        for (int k = k0; k < k1; k += JP_BATCH) {
            int2 q[JP_BATCH];
            int cq[JP_BATCH];
#pragma unroll
            for (int j = 0; j < JP_BATCH; j++) q[j] = k + j < k1 ? d.jp_adj[k + j] : make_int2(p, pr);
#pragma unroll
            for (int j = 0; j < JP_BATCH; j++) {
                const bool lower = q[j].x != p && q[j].y < pr;
                cq[j] = lower ? jp_col_load(d.jp_col, q[j].x) : -2;
            }
            int wait = 0;
#pragma unroll
            for (int j = 0; j < JP_BATCH; j++) {
                wait |= cq[j] == -1;
                if (cq[j] >= 64) big = true;
                else if (cq[j] >= 0) used |= 1ull << cq[j];
            }
            if (wait) return -1;
        }
    }
    if (~used != 0ull) return __ffsll((unsigned long long)~used) - 1;
    // 64 colours and more around this pair: the first colour no lower neighbour holds
    (void)big;
    for (int c = 64;; c++) {
        bool taken = false;
        for (int s = 0; s < 2 && !taken; s++) {
            const int x = body[s];
            if (d.b.iat(I_KIN, x)) continue;
            for (int k = (int)d.jp_deg[x]; k < (int)d.jp_deg[x + 1] && !taken; k++) {
                const int2 q = d.jp_adj[k];
                if (q.x != p && q.y < pr && jp_col_load(d.jp_col, q.x) == c) taken = true;
            }
        }
        if (!taken) return c;
    }
}
// Rounds until every pair has its colour (each round colours at least the uncoloured pair of lowest
// priority, so the loop ends), then the colour starts and the solve order, colour by colour.  First-fit
// colours are dense (a pair of colour c has lower neighbours of every colour below c), so the colour
// count is the first empty bin.  The colour counts go through LDS histograms, one global add per
// workgroup and bin.  Launched cooperatively (all workgroups resident).
constexpr int JP_BINS = 64;
__global__ void __launch_bounds__(JP_THREADS) k_jp_colour(Dev d) {
    __shared__ int s_any;
    __shared__ int s_hist[JP_BINS], s_base[JP_BINS];
    const int nact = d.cnt->nact;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    // Round 0 takes every active pair, each later round the pairs the round before left uncoloured
    // (lists alternating in act + pcap and act + 2 pcap; the count of round r's list in
    // jp_rem[r % 3], cleared two rounds ahead, after every thread has read it).  Once at most
    // JP_THREADS pairs are left, workgroup 0 finishes them alone, with workgroup barriers.
    const int32_t* cur = d.act;
    int ncur = nact;
    int r = 0;
    for (;; r++) {
        int32_t* nxt = d.act + (size_t)(1 + (r & 1)) * d.pcap;
        if (gtid == 0) d.cnt->jp_rem[(r + 2) % 3] = 0;
        const int n_iter = (ncur + gsz - 1) / gsz;
        for (int it = 0; it < n_iter; it++) {
            const int k = it * gsz + gtid;
            int p = -1, c = 0;
            if (k < ncur) {
                p = cur[k];
                c = jp_try(d, p);
                if (c >= 0) __hip_atomic_store(&d.jp_col[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const int slot = wave_append(p >= 0 && c < 0, (int*)&d.cnt->jp_rem[(r + 1) % 3]);
            if (p >= 0 && c < 0) nxt[slot] = p;
        }
        grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);
        ncur = (int)__hip_atomic_load(&d.cnt->jp_rem[(r + 1) % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        cur = nxt;
        if (ncur <= JP_THREADS) break;
    }
    r++;
    if (blockIdx.x == 0) {
        const int p = (int)threadIdx.x < ncur ? cur[threadIdx.x] : -1;
        bool done = p < 0;
        for (;; r++) {
            if (threadIdx.x == 0) s_any = 0;
            __syncthreads();
            if (!done) {
                const int c = jp_try(d, p);
                if (c >= 0) {
                    __hip_atomic_store(&d.jp_col[p], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    done = true;
                } else {
                    s_any = 1;
                }
            }
            __syncthreads();
            if (!s_any) break;
        }
        if (threadIdx.x == 0) d.cnt->jp_rounds = ncur > 0 ? r + 1 : r;
    }
    grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);   // every colour is set
    // pairs per colour: LDS bins per workgroup, one global add per workgroup and bin
    for (int c = threadIdx.x; c < JP_BINS; c += blockDim.x) s_hist[c] = 0;
    __syncthreads();
    for (int k = gtid; k < nact; k += gsz) {
        const int c = d.jp_col[d.act[k]];
        if (c < JP_BINS) atomicAdd(&s_hist[c], 1);
        else atomicAdd(&d.jp_ccount[c], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < JP_BINS; c += blockDim.x)
        if (s_hist[c]) atomicAdd(&d.jp_ccount[c], s_hist[c]);
    grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);
    if (gtid == 0) {
        int run = 0, ncol = 0;
        while (ncol <= d.pcap && d.jp_ccount[ncol] > 0) {
            d.color_start[ncol] = run;
            run += d.jp_ccount[ncol];
            d.jp_ccount[ncol] = 0;
            ncol++;
        }
        d.color_start[ncol] = run;
        d.cnt->ncolors = ncol;
        d.cnt->nordered = run;
    }
    grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);
    // the solve order: each workgroup reserves its range of every colour once, then places its pairs
    for (int c = threadIdx.x; c < JP_BINS; c += blockDim.x) {
        s_base[c] = s_hist[c] ? atomicAdd(&d.jp_ccount[c], s_hist[c]) : 0;
        s_hist[c] = 0;
    }
    __syncthreads();
    for (int k = gtid; k < nact; k += gsz) {
        const int p = d.act[k];
        const int c = d.jp_col[p];
        const int at = c < JP_BINS ? s_base[c] + atomicAdd(&s_hist[c], 1) : atomicAdd(&d.jp_ccount[c], 1);
        d.order[d.color_start[c] + at] = p;
    }
}

// Gauss-Seidel sweeps (SequentialImpulseSolver::solve, sequential_impulse_solver.cpp:60-81) in
// colour order across the whole GPU: the pairs of one colour share no dynamic body, so they run
// in parallel; a grid barrier separates colours.  Then syncTempVelocity.  Launched cooperatively
// (all workgroups co-resident).
constexpr int GS_THREADS = 256;

__global__ void __launch_bounds__(GS_THREADS) k_gs_grid(Dev d) {
    const int ncol = d.cnt->ncolors;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    // A thread's first pair of each colour is loaded before the barrier ahead of it (gs_prefetch):
    // only the bodies' temp velocities are read after the barrier.
    GsPair G;
    bool have = false;
    auto fetch = [&](int c) {
        const int q = d.color_start[c] + gtid;
        have = q < d.color_start[c + 1];
        if (have) gs_prefetch(d.b, d.ps, d.order[q], G);
    };
    if (ncol > 0) fetch(0);
    // The last colours of at most GS_THREADS pairs each: every pair of such a colour falls to a thread
    // of workgroup 0, so between two of them a workgroup barrier orders the updates (the grid barrier
    // after the last one publishes them to the next iteration).
    int ctail = ncol;
    while (ctail > 0 && d.color_start[ctail] - d.color_start[ctail - 1] <= GS_THREADS) ctail--;
    for (int it = 0; it < d.iterations; it++) {
        for (int c = 0; c < ncol; c++) {
            const int c0 = d.color_start[c], c1 = d.color_start[c + 1];
            if (have) {
                gs_solve(d.b, d.ps, G);
                for (int q = c0 + gtid + gsz; q < c1; q += gsz) {
                    gs_prefetch(d.b, d.ps, d.order[q], G);
                    gs_solve(d.b, d.ps, G);
                }
            }
            if (c + 1 < ncol) fetch(c + 1);
            else if (it + 1 < d.iterations) fetch(0);
            if (c >= ctail && c + 1 < ncol) __syncthreads();
            else grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);   // later colours read these velocities
        }
        // the joints after the contacts, colour by colour (sequential_impulse_solver.cpp:66-68)
        for (int c = 0; c < d.njcol; c++) {
            for (int q = d.jcs[c] + gtid; q < d.jcs[c + 1]; q += gsz) {
                const int k = d.jorder[q];
                joint_iterate(d.b, d.joints[k], d.jrows[k]);
            }
            grid_barrier(&d.cnt->bar_count, &d.cnt->bar_gen, gridDim.x);
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
// (ms per step, 4 / 8 / 16 / 32 / 64 workgroups of 256): box rain of 1 000 bodies 0.45 / 0.47 / 0.49 /
// 0.49, of 5 000 bodies 1.29 / 1.14 / 1.10 / 1.14, of 10 000 bodies (colouring rounds) - / 1.94 / 1.76 /
// 1.72 / 1.82; 50k shape mix at step 280: 16 / 32 / 64 workgroups 7.29 / 7.06 / 7.08.
static int default_gs_blocks(int n, int max_blocks) { return std::min(max_blocks, std::max(4, std::min(32, n / 300))); }
// Bodies beyond which the colouring runs in rounds over the GPU rather than in one workgroup (box
// rain at rest, ms per step with one workgroup / rounds: 1 000 bodies 0.42 / 0.46, 2 500 bodies
// 0.67 / 0.69, 4 000 bodies 0.89 / 0.87, 6 000 bodies 1.26 / 1.14, 8 000 bodies 1.61 / 1.33).
constexpr int COLOUR_ROUNDS_FROM = 3000;

struct pe_world {
    Dev d{};
    int n = 0, stride = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    int gs_blocks = 0;      // workgroups of the cooperative Gauss-Seidel kernel
    int gs_max_blocks = 0;
    int jp_blocks = 1;                       // workgroups of the cooperative colouring  // co-resident limit
    int colouring = 0;                       // 0: by body count, 1: one workgroup's claims, 2: rounds
    bool timing = false;
    hipEvent_t ev[5] = {};
    hipEvent_t order_ev[2] = {};   // orders work on a caller's stream with the world's own stream
    double stage_ms[4] = {0, 0, 0, 0};
    int n_dynamic = 0;
    bool has_shapes = false;   // any cylinder / convex / concave body
    DevMeshes meshes;
    std::vector<JointDef> joints;            // host copies of the joint definitions
    std::vector<uint64_t> ignore_keys;       // sorted ignore-pair keys
    std::vector<char> kinematic;             // per body (joint colouring)
    std::vector<char> finite_life, has_joint, removed;   // per body (joints and removals exclude each other)
    void* joint_mem[4] = {};                 // device: definitions, rows, order, colour starts
    void* ignore_mem = nullptr;
    void* ray_mem = nullptr;                 // raycast inputs, ignore set and results
    size_t ray_bytes = 0;

    ~pe_world() {
        if (stream) hipStreamSynchronize(stream);
        for (void* p : allocs) hipFree(p);
        for (void* p : joint_mem) if (p) hipFree(p);
        if (ignore_mem) hipFree(ignore_mem);
        if (ray_mem) hipFree(ray_mem);
        meshes.release();
        for (auto& e : ev) if (e) hipEventDestroy(e);
        for (auto& e : order_ev) if (e) hipEventDestroy(e);
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
        if (d.status_on) hipLaunchKernelGGL(k_integrate_status, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        else hipLaunchKernelGGL(k_integrate, dim3(grid(n, 256)), dim3(256), 0, stream, d);
        return PE_OK;
    }
    int launch_broadphase() {
        sap_launch(d.sap, stream);
        return PE_OK;
    }
    int launch_narrowphase() {
        hipLaunchKernelGGL(k_narrow, dim3(grid(d.pcap, 128)), dim3(128), 0, stream, d);
        if (has_shapes) {
            hipLaunchKernelGGL(k_narrow_shapes, dim3(grid(d.pcap, 64)), dim3(64), 0, stream, d);
            hipLaunchKernelGGL(k_narrow_sat, dim3(std::min(grid(d.pcap, 4), 4096)), dim3(256), 0, stream, d);
        }
        if (d.comp_list) hipLaunchKernelGGL(k_narrow_compound, dim3(grid(d.pcap, 64)), dim3(64), 0, stream, d);
        hipLaunchKernelGGL(k_narrow_finish, dim3(1), dim3(1), 0, stream, d);
        hipLaunchKernelGGL(k_count_points, dim3(grid(d.pcap, 1024)), dim3(256), 0, stream, d);
        if (d.status_on) hipLaunchKernelGGL(k_neighbours, dim3(grid(d.pcap, 256)), dim3(256), 0, stream, d);
        if (d.report_on) hipLaunchKernelGGL(k_events, dim3(grid(d.pcap, 256)), dim3(256), 0, stream, d);
        return PE_OK;
    }
    int launch_solve() {
        hipLaunchKernelGGL(k_setup, dim3(grid(d.pcap, 128)), dim3(128), 0, stream, d);
        if (d.njoints > 0) hipLaunchKernelGGL(k_joint_setup, dim3(grid(d.njoints, 64)), dim3(64), 0, stream, d);
        if (colouring == 1 || (colouring == 0 && n <= COLOUR_ROUNDS_FROM)) {
            // claims in one workgroup's LDS (at most CLAIM_LDS bodies)
            hipLaunchKernelGGL(k_colour<true>, dim3(1), dim3(SOLVE_THREADS), n * sizeof(int), stream, d);
        } else {
            // colouring by rounds: the pairs and their bodies' adjacency, then the rounds (cooperative)
            hipLaunchKernelGGL(k_jp_init, dim3(grid(std::max(n + 1, d.pcap + 1), 256)), dim3(256), 0, stream, d);
            hipLaunchKernelGGL(k_jp_active, dim3(grid(d.pcap, 256)), dim3(256), 0, stream, d);
            sap_scan(d.sap, d.jp_deg, n + 1, n + 1, nullptr, &d.cnt->adj_total, stream);
            hipLaunchKernelGGL(k_jp_fill, dim3(grid(d.pcap, 256)), dim3(256), 0, stream, d);
            void* jargs[] = {(void*)&d};
            PE_HIP(hipLaunchCooperativeKernel((const void*)k_jp_colour, dim3(jp_blocks), dim3(JP_THREADS), jargs, 0, stream));
        }
        void* args[] = {(void*)&d};
        PE_HIP(hipLaunchCooperativeKernel((const void*)k_gs_grid, dim3(gs_blocks), dim3(GS_THREADS), args, 0, stream));
        return PE_OK;
    }
    int step() {
        pe_clear_stale_error();
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
    const int64_t want = max_pairs > 0 ? (int64_t)max_pairs : std::max<int64_t>(1024, 16 * (int64_t)n);
    if (want > ((int64_t)1 << colour::PRIO_INDEX_BITS)) {
        delete w;
        pe_set_error("pe_world_create: at most 2^28 pairs per world");
        return PE_ERR_INVALID;
    }
    const int pcap = (int)want;
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
    rc |= w->alloc(&d.cnt, 1);
    rc |= w->alloc(&d.jp_deg, (size_t)n + 1);
    rc |= w->alloc(&d.jp_fill, n);
    rc |= w->alloc(&d.jp_adj, 2 * (size_t)pcap);
    rc |= w->alloc(&d.jp_col, pcap);
    rc |= w->alloc(&d.jp_ccount, (size_t)pcap + 1);
    rc |= w->alloc(&d.st.flags, n);
    rc |= w->alloc(&d.st.sleep_time, n);
    rc |= w->alloc(&d.st.life, n);
    rc |= w->alloc(&d.st.last, n);
    rc |= w->alloc(&d.st.scnt, n);
    rc |= w->alloc(&d.st.dcnt, n);
    rc |= w->alloc(&d.st.report, n);
    rc |= w->alloc(&d.events, (size_t)EVW * pcap);
    rc |= w->alloc(&d.ev_flag, pcap);
    SapView& sp = d.sap;
    sp.n = n;
    sp.amin = d.b.f + (size_t)F_AMIN * w->stride;
    sp.amax = d.b.f + (size_t)F_AMAX * w->stride;
    sp.fs = w->stride;
    sp.bs = 1;
    sp.rec = d.b.ab;
    sp.cnt = d.cnt;
    sp.pa = d.ps.pa;
    sp.pb = d.ps.pb;
    sp.pcap = pcap;
    rc |= w->alloc(&sp.keys, n);
    rc |= w->alloc(&sp.vals, n);
    rc |= w->alloc(&sp.keys2, n);
    rc |= w->alloc(&sp.vals2, n);
    rc |= w->alloc(&sp.skey, n);
    rc |= w->alloc(&sp.srec, (size_t)8 * n);
    rc |= w->alloc(&sp.cand, n + 1);
    rc |= w->alloc(&sp.hist, 256 * (size_t)sap_radix_blocks(n));
    rc |= w->alloc(&sp.red, SAP_RED_VALUES * (size_t)sap_red_blocks(n));
    // sweep chunks: the candidates (bodies overlapping on the sweep axis) are at most n (n - 1) / 2,
    // and a pile that lands while the axis is still the fall direction comes close to that (the
    // reference's sweep too: the axis is chosen after the sweep); 8 bytes per 32 candidates, so the
    // bound itself is allocated, up to 2^31 candidates (the counters are 32-bit)
    const int64_t max_cand = std::min<int64_t>((int64_t)n * (n - 1) / 2, (int64_t)1 << 31);
    sp.chunk_cap = (int)std::max<int64_t>(max_cand / SAP_CHUNK + 1, 1024);
    rc |= w->alloc(&sp.chunk, sp.chunk_cap);
    rc |= w->alloc(&sp.chunk_hit, sp.chunk_cap);
    rc |= w->alloc(&sp.scan_part, sap_scan_parts(n, sp.chunk_cap));
    if (rc != PE_OK) return fail(PE_ERR_DEVICE);

    // host staging of the body store (FP32) and of the scene's meshes
    std::vector<float> hf(NF * S, 0.0f);
    std::vector<int32_t> hi(NI * S, 0);
    HostMeshes hm;
    if (int rc2 = stage_bodies(sc, hf, hi, S, 0, hm)) return fail(rc2);
    if (int rc2 = w->meshes.upload(hm, w->stream)) return fail(rc2);
    d.b.ms = w->meshes.view();
    int ndyn = 0;
    bool has_compound = false;
    w->kinematic.resize(n);
    for (int i = 0; i < n; i++) {
        ndyn += sc->kinematic[i] ? 0 : 1;
        w->kinematic[i] = sc->kinematic[i] ? 1 : 0;
        w->has_shapes |= sc->shape_type[i] > PE_SHAPE_SPHERE;
        has_compound |= sc->shape_type[i] == PE_SHAPE_COMPOUND;
    }
    if (has_compound && w->alloc(&d.comp_list, pcap) != PE_OK) return fail(PE_ERR_DEVICE);
    w->n_dynamic = ndyn;
    // RigidBody::_life_time (rigidbody.cpp:97): FP32 PE_REAL_MAX unless the scene sets one
    std::vector<float> life(n, 1e30f);
    w->finite_life.assign(n, 0);
    for (int i = 0; sc->lifetime && i < n; i++) {
        life[i] = (float)sc->lifetime[i];
        if (!(sc->lifetime[i] >= 1e30)) { w->finite_life[i] = 1; d.status_on = 1; }
    }
    w->has_joint.assign(n, 0);
    w->removed.assign(n, 0);
    if (hipMemcpyAsync(d.b.f, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice, w->stream) != hipSuccess ||
        hipMemcpyAsync(d.b.i, hi.data(), hi.size() * sizeof(int32_t), hipMemcpyHostToDevice, w->stream) != hipSuccess ||
        hipMemcpyAsync(d.st.life, life.data(), n * sizeof(float), hipMemcpyHostToDevice, w->stream) != hipSuccess) {
        pe_set_error("upload failed");
        return fail(PE_ERR_DEVICE);
    }
    for (void* p : {(void*)d.st.flags, (void*)d.st.sleep_time, (void*)d.st.last, (void*)d.st.scnt, (void*)d.st.dcnt,
                    (void*)d.st.report})
        if (hipMemsetAsync(p, 0, n * sizeof(int32_t), w->stream) != hipSuccess) { pe_set_error("upload failed"); return fail(PE_ERR_DEVICE); }
    if (hipMemsetAsync(d.ev_flag, 0, pcap * sizeof(int32_t), w->stream) != hipSuccess) { pe_set_error("upload failed"); return fail(PE_ERR_DEVICE); }
    hipLaunchKernelGGL(k_aabb_only, dim3(w->grid(n, 256)), dim3(256), 0, w->stream, d);
    for (auto& e : w->ev) hipEventCreate(&e);
    for (auto& e : w->order_ev) hipEventCreateWithFlags(&e, hipEventDisableTiming);
    {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gs_grid, GS_THREADS, 0) != hipSuccess || per_cu < 1) {
            pe_set_error("cannot size the cooperative solver grid");
            return fail(PE_ERR_DEVICE);
        }
        w->gs_max_blocks = cus * per_cu;
        w->gs_blocks = default_gs_blocks(n, w->gs_max_blocks);
        int jp_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&jp_cu, (const void*)k_jp_colour, JP_THREADS, 0) != hipSuccess ||
            jp_cu < 1) {
            pe_set_error("cannot size the cooperative colouring grid");
            return fail(PE_ERR_DEVICE);
        }
        // as many workgroups as the Gauss-Seidel sweep: each round ends in a grid barrier, and fewer,
        // fuller workgroups measured faster (10k box rain at rest, ms per step with 16 / 32 / 64 / 128 /
        // 256 workgroups: 1.22 / 1.19 / 1.20 / 1.22 / 1.31; 50k shape mix 5.90 / 5.84 / 5.86 / 5.91 / 5.98)
        w->jp_blocks = std::max(1, std::min(cus * jp_cu, default_gs_blocks(n, cus * jp_cu)));
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) { pe_set_error("init sync failed"); return fail(PE_ERR_DEVICE); }
    *out = w;
    return PE_OK;
}

int pe_world_add_joints(pe_world* w, int32_t n, const pe_joint_desc* jd) {
    if (int rc = check_world(w)) return rc;
    if (n < 0 || (n > 0 && !jd) || (int64_t)w->joints.size() + n > 65536) { pe_set_error("pe_world_add_joints: invalid argument"); return PE_ERR_INVALID; }
    for (int k = 0; k < n; k++) {
        const pe_joint_desc& s = jd[k];
        if (s.type < PE_JOINT_BALL || s.type > PE_JOINT_SLIDER || s.body_a < 0 || s.body_a >= w->n || s.body_b < 0 ||
            s.body_b >= w->n || s.body_a == s.body_b) {
            pe_set_error("pe_world_add_joints: joint " + std::to_string(k) + " has an unknown type or invalid bodies");
            return PE_ERR_INVALID;
        }
        for (int x : {s.body_a, s.body_b})
            if (w->finite_life[x] || w->removed[x]) {
                pe_set_error("pe_world_add_joints: body " + std::to_string(x) + " has a finite lifetime or was removed");
                return PE_ERR_INVALID;
            }
    }
    for (int k = 0; k < n; k++) {
        const pe_joint_desc& s = jd[k];
        w->has_joint[s.body_a] = w->has_joint[s.body_b] = 1;
        JointDef j{};
        j.type = s.type; j.a = s.body_a; j.b = s.body_b;
        j.flags = (s.use_limits ? JF_LIMITS : 0) | (s.use_motor ? JF_MOTOR : 0);
        for (int c = 0; c < 3; c++) {
            j.anchor_a[c] = (float)s.anchor_a[c]; j.anchor_b[c] = (float)s.anchor_b[c];
            j.axis_a[c] = (float)s.axis_a[c]; j.axis_b[c] = (float)s.axis_b[c];
        }
        j.min_angle = (float)s.min_angle; j.max_angle = (float)s.max_angle; j.target_speed = (float)s.target_speed;
        w->joints.push_back(j);
    }
    // greedy colouring in joint order: joint k takes the lowest colour no earlier joint sharing a
    // dynamic body with it holds
    const int nj = (int)w->joints.size();
    std::vector<int> col(nj, 0), order;
    std::vector<std::vector<int>> body_cols(w->n);
    int ncol = 0;
    for (int k = 0; k < nj; k++) {
        std::vector<char> used(ncol + 1, 0);
        for (int x : {w->joints[k].a, w->joints[k].b})
            if (!w->kinematic[x]) for (int c : body_cols[x]) used[c] = 1;
        int c = 0;
        while (used[c]) c++;
        col[k] = c;
        ncol = std::max(ncol, c + 1);
        for (int x : {w->joints[k].a, w->joints[k].b})
            if (!w->kinematic[x]) body_cols[x].push_back(c);
    }
    std::vector<int32_t> jorder, jcs;
    for (int c = 0; c < ncol; c++) {
        jcs.push_back((int32_t)jorder.size());
        for (int k = 0; k < nj; k++) if (col[k] == c) jorder.push_back(k);
    }
    jcs.push_back((int32_t)jorder.size());
    PE_HIP(hipStreamSynchronize(w->stream));
    for (void*& p : w->joint_mem) if (p) { hipFree(p); p = nullptr; }
    const size_t bytes[4] = {nj * sizeof(JointDef), nj * sizeof(JointRow), nj * sizeof(int32_t), jcs.size() * sizeof(int32_t)};
    const void* src[4] = {w->joints.data(), nullptr, jorder.data(), jcs.data()};
    for (int k = 0; k < 4; k++) {
        PE_HIP(hipMalloc(&w->joint_mem[k], std::max<size_t>(bytes[k], 16)));
        if (src[k] && bytes[k]) PE_HIP(hipMemcpy(w->joint_mem[k], src[k], bytes[k], hipMemcpyHostToDevice));
    }
    w->d.joints = (const JointDef*)w->joint_mem[0];
    w->d.jrows = (JointRow*)w->joint_mem[1];
    w->d.jorder = (const int32_t*)w->joint_mem[2];
    w->d.jcs = (const int32_t*)w->joint_mem[3];
    w->d.njoints = nj;
    w->d.njcol = ncol;
    return PE_OK;
}

// Marks the bodies in ignore pairs (COLL_PARTNERS in their fourth shape parameter, unless they ignore
// every collision) and refreshes their AABB records.
__global__ void k_mark_partners(Dev d, const uint64_t* keys, int nkeys) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nkeys; k += gridDim.x * blockDim.x) {
        const int i = (int)(keys[k] >> 32), j = (int)(uint32_t)keys[k];
        for (int x : {i, j}) {
            float& f = d.b.at(F_SP + 3, x);
            if (f != COLL_IGNORE) f = COLL_PARTNERS;
            d.b.abr(x)[7] = f;
        }
    }
}

int pe_world_add_ignore_pairs(pe_world* w, int32_t n, const int32_t* pairs) {
    if (int rc = check_world(w)) return rc;
    if (n < 0 || (n > 0 && !pairs)) { pe_set_error("pe_world_add_ignore_pairs: invalid argument"); return PE_ERR_INVALID; }
    for (int k = 0; k < n; k++) {
        const int a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a < 0 || a >= w->n || b < 0 || b >= w->n || a == b) { pe_set_error("pe_world_add_ignore_pairs: invalid body index"); return PE_ERR_INVALID; }
        w->ignore_keys.push_back((uint64_t)std::min(a, b) << 32 | (uint32_t)std::max(a, b));
    }
    std::sort(w->ignore_keys.begin(), w->ignore_keys.end());
    w->ignore_keys.erase(std::unique(w->ignore_keys.begin(), w->ignore_keys.end()), w->ignore_keys.end());
    PE_HIP(hipStreamSynchronize(w->stream));
    if (w->ignore_mem) { hipFree(w->ignore_mem); w->ignore_mem = nullptr; }
    const size_t bytes = w->ignore_keys.size() * sizeof(uint64_t);
    PE_HIP(hipMalloc(&w->ignore_mem, std::max<size_t>(bytes, 16)));
    if (bytes) PE_HIP(hipMemcpy(w->ignore_mem, w->ignore_keys.data(), bytes, hipMemcpyHostToDevice));
    w->d.sap.ignore = (const uint64_t*)w->ignore_mem;
    w->d.sap.nignore = (int)w->ignore_keys.size();
    if (!w->ignore_keys.empty())
        hipLaunchKernelGGL(k_mark_partners, dim3(w->grid((int)w->ignore_keys.size(), 256)), dim3(256), 0, w->stream, w->d,
                           (const uint64_t*)w->ignore_mem, (int)w->ignore_keys.size());
    PE_HIP(hipGetLastError());
    PE_HIP(hipStreamSynchronize(w->stream));
    return PE_OK;
}

static int read_counts(pe_world* w, DevCounts& c);

int pe_world_raycast(pe_world* w, int32_t nrays, const float* rays, int32_t nignore, const int32_t* ignore,
                     pe_ray_hit* out) {
    if (int rc = check_world(w)) return rc;
    if (nrays < 0 || nignore < 0 || (nrays > 0 && (!rays || !out)) || (nignore > 0 && !ignore) || nrays > (1 << 24)) {
        pe_set_error("pe_world_raycast: invalid argument");
        return PE_ERR_INVALID;
    }
    if (nrays == 0) return PE_OK;
    std::vector<int32_t> ig(ignore, ignore + nignore);
    std::sort(ig.begin(), ig.end());
    ig.erase(std::unique(ig.begin(), ig.end()), ig.end());
    pe_clear_stale_error();
    const size_t in_b = (size_t)nrays * RAY_IN * 4, ig_b = std::max<size_t>(ig.size(), 1) * 4, out_b = (size_t)nrays * RAY_OUT * 4;
    if (w->ray_bytes < in_b + ig_b + out_b + 512) {
        if (w->ray_mem) hipFree(w->ray_mem);
        w->ray_mem = nullptr;
        w->ray_bytes = 0;
        PE_HIP(hipMalloc(&w->ray_mem, in_b + ig_b + out_b + 512));
        w->ray_bytes = in_b + ig_b + out_b + 512;
    }
    char* base = (char*)w->ray_mem;
    float* d_rays = (float*)base;
    int32_t* d_ig = (int32_t*)(base + ((in_b + 255) & ~(size_t)255));
    float* d_out = (float*)(base + ((in_b + 255) & ~(size_t)255) + ((ig_b + 255) & ~(size_t)255));
    PE_HIP(hipMemcpyAsync(d_rays, rays, in_b, hipMemcpyHostToDevice, w->stream));
    if (!ig.empty()) PE_HIP(hipMemcpyAsync(d_ig, ig.data(), ig.size() * 4, hipMemcpyHostToDevice, w->stream));
    hipLaunchKernelGGL(k_raycast, dim3(nrays), dim3(64), 0, w->stream, w->d, (const float*)d_rays, (const int32_t*)d_ig,
                       (int)ig.size(), d_out);
    PE_HIP(hipGetLastError());
    std::vector<float> res((size_t)nrays * RAY_OUT);
    PE_HIP(hipMemcpyAsync(res.data(), d_out, out_b, hipMemcpyDeviceToHost, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    for (int k = 0; k < nrays; k++) {
        const float* e = &res[(size_t)k * RAY_OUT];
        pe_ray_hit& h = out[k];
        h.hit = e[0] != 0.0f ? 1 : 0;
        std::memcpy(&h.body, &e[1], 4);
        h.distance = e[2];
        for (int c = 0; c < 3; c++) { h.point[c] = e[3 + c]; h.normal[c] = e[6 + c]; }
    }
    return PE_OK;
}

int pe_world_set_sleep(pe_world* w, double lin_vel2, double ang_vel2, double time) {
    if (int rc = check_world(w)) return rc;
    w->d.st.lin2 = (float)lin_vel2;
    w->d.st.ang2 = (float)ang_vel2;
    w->d.st.tsleep = (float)time;
    if (lin_vel2 != 0.0 || ang_vel2 != 0.0) w->d.status_on = 1;
    return PE_OK;
}

int pe_world_remove_body(pe_world* w, int32_t body) {
    if (int rc = check_world(w)) return rc;
    if (body < 0 || body >= w->n) { pe_set_error("pe_world_remove_body: body index out of range"); return PE_ERR_INVALID; }
    if (w->has_joint[body]) { pe_set_error("pe_world_remove_body: the body has joints"); return PE_ERR_INVALID; }
    w->removed[body] = 1;
    w->d.status_on = 1;
    pe_clear_stale_error();
    hipLaunchKernelGGL(k_remove_body, dim3(1), dim3(1), 0, w->stream, w->d, (int)body);
    PE_HIP(hipGetLastError());
    return PE_OK;
}

int pe_world_get_body_status(pe_world* w, int32_t* flags, int32_t* static_count, int32_t* dynamic_count,
                             float* sleep_time) {
    if (int rc = check_world(w)) return rc;
    PE_HIP(hipStreamSynchronize(w->stream));
    const size_t bytes = (size_t)w->n * 4;
    if (flags) PE_HIP(hipMemcpy(flags, w->d.st.flags, bytes, hipMemcpyDeviceToHost));
    if (static_count) PE_HIP(hipMemcpy(static_count, w->d.st.scnt, bytes, hipMemcpyDeviceToHost));
    if (dynamic_count) PE_HIP(hipMemcpy(dynamic_count, w->d.st.dcnt, bytes, hipMemcpyDeviceToHost));
    if (sleep_time) PE_HIP(hipMemcpy(sleep_time, w->d.st.sleep_time, bytes, hipMemcpyDeviceToHost));
    return PE_OK;
}

int pe_world_set_collision_reports(pe_world* w, int32_t n, const int32_t* bodies, int32_t on) {
    if (int rc = check_world(w)) return rc;
    if (n < 0 || (n > 0 && !bodies)) { pe_set_error("pe_world_set_collision_reports: invalid argument"); return PE_ERR_INVALID; }
    for (int k = 0; k < n; k++)
        if (bodies[k] < 0 || bodies[k] >= w->n) { pe_set_error("pe_world_set_collision_reports: body index out of range"); return PE_ERR_INVALID; }
    PE_HIP(hipStreamSynchronize(w->stream));
    std::vector<int32_t> rep(w->n);
    PE_HIP(hipMemcpy(rep.data(), w->d.st.report, (size_t)w->n * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) rep[bodies[k]] = on ? 1 : 0;
    PE_HIP(hipMemcpy(w->d.st.report, rep.data(), (size_t)w->n * 4, hipMemcpyHostToDevice));
    w->d.report_on = 0;
    for (int32_t r : rep) w->d.report_on |= r;
    return PE_OK;
}

int pe_world_get_collision_events(pe_world* w, int32_t cap, pe_collision_event* out, int32_t* count) {
    if (int rc = check_world(w)) return rc;
    if (cap < 0 || (cap > 0 && !out) || !count) { pe_set_error("pe_world_get_collision_events: invalid argument"); return PE_ERR_INVALID; }
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    const int np = std::min(c.npairs, w->d.pcap);
    int total = 0;
    if (w->d.report_on && np > 0) {
        std::vector<int32_t> flag(np);
        std::vector<float> ev((size_t)np * EVW);
        PE_HIP(hipMemcpy(flag.data(), w->d.ev_flag, (size_t)np * 4, hipMemcpyDeviceToHost));
        PE_HIP(hipMemcpy(ev.data(), w->d.events, ev.size() * 4, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; p++) {
            if (!flag[p]) continue;
            if (total < cap) {
                const float* e = &ev[(size_t)p * EVW];
                pe_collision_event& o = out[total];
                std::memcpy(&o.body_a, &e[0], 4);
                std::memcpy(&o.body_b, &e[1], 4);
                for (int k = 0; k < 3; k++) { o.position[k] = e[2 + k]; o.normal[k] = e[5 + k]; o.velocity[k] = e[9 + k]; }
                o.depth = e[8];
            }
            total++;
        }
    }
    *count = total;
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
static int read_counts(pe_world* w, DevCounts& c) {
    PE_HIP(hipMemcpyAsync(&c, w->d.cnt, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    PE_HIP(hipStreamSynchronize(w->stream));
    return PE_OK;
}
int pe_world_step_async(pe_world* w, int32_t nsteps, pe_stream_t stream) {
    if (int rc = check_world(w)) return rc;
    if (nsteps < 0) { pe_set_error("nsteps must be >= 0"); return PE_ERR_INVALID; }
    hipStream_t cs = (hipStream_t)stream;
    // the caller's stream first waits for the world's pending work ...
    PE_HIP(hipEventRecord(w->order_ev[0], w->stream));
    PE_HIP(hipStreamWaitEvent(cs, w->order_ev[0], 0));
    hipStream_t keep = w->stream;
    w->stream = cs;
    int rc = PE_OK;
    for (int s = 0; s < nsteps && rc == PE_OK; s++) rc = w->step();
    w->stream = keep;
    // ... and the world's later calls wait for the steps enqueued on it
    PE_HIP(hipEventRecord(w->order_ev[1], cs));
    PE_HIP(hipStreamWaitEvent(w->stream, w->order_ev[1], 0));
    return rc;
}
int pe_world_get_cache(pe_world* w, float* distances, int32_t cap, int32_t* prev_npairs) {
    if (int rc = check_world(w)) return rc;
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    if (prev_npairs) *prev_npairs = c.prev_npairs;
    const int m = std::min(cap, c.prev_npairs);
    if (distances && m > 0) PE_HIP(hipMemcpy(distances, w->d.ps.cd, (size_t)m * 8 * sizeof(float), hipMemcpyDeviceToHost));
    return PE_OK;
}
int pe_world_set_cache(pe_world* w, const float* distances, int32_t prev_npairs) {
    if (int rc = check_world(w)) return rc;
    if (prev_npairs < 0 || prev_npairs > w->d.pcap || (prev_npairs > 0 && !distances)) {
        pe_set_error("pe_world_set_cache: pair count beyond the world's pair capacity");
        return PE_ERR_INVALID;
    }
    DevCounts c;
    if (int rc = read_counts(w, c)) return rc;
    c.prev_npairs = prev_npairs;
    if (prev_npairs > 0)
        PE_HIP(hipMemcpy(w->d.ps.cd, distances, (size_t)prev_npairs * 8 * sizeof(float), hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(w->d.cnt, &c, sizeof(c), hipMemcpyHostToDevice));
    return PE_OK;
}
int pe_world_integrate(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    pe_clear_stale_error();
    w->launch_integrate();
    PE_HIP(hipGetLastError());
    return PE_OK;
}
int pe_world_broadphase(pe_world* w, int32_t* npairs) {
    if (int rc = check_world(w)) return rc;
    pe_clear_stale_error();
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
    pe_clear_stale_error();
    w->launch_narrowphase();
    PE_HIP(hipGetLastError());
    return PE_OK;
}
int pe_world_solve(pe_world* w) {
    if (int rc = check_world(w)) return rc;
    pe_clear_stale_error();
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

int pe_world_set_colouring(pe_world* w, int32_t mode) {
    if (int rc = check_world(w)) return rc;
    if (mode < 0 || mode > 2 || (mode == 1 && w->n > CLAIM_LDS)) {
        pe_set_error("pe_world_set_colouring: mode 0, 1 (at most 12288 bodies) or 2");
        return PE_ERR_INVALID;
    }
    w->colouring = mode;
    return PE_OK;
}

int pe_world_get_records(pe_world* w, float* aabb, float* body, float* solver) {
    if (int rc = check_world(w)) return rc;
    const size_t n = (size_t)w->n;
    PE_HIP(hipStreamSynchronize(w->stream));
    if (aabb) PE_HIP(hipMemcpy(aabb, w->d.b.ab, n * 8 * sizeof(float), hipMemcpyDeviceToHost));
    if (body) PE_HIP(hipMemcpy(body, w->d.b.cb, n * CBW * sizeof(float), hipMemcpyDeviceToHost));
    if (solver) PE_HIP(hipMemcpy(solver, w->d.b.sb, n * SBW * sizeof(float), hipMemcpyDeviceToHost));
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
    counts[5] = c.jp_rounds;
    return PE_OK;
}

}  // extern "C"

// ====================================================================== stage entry points
// The three stage seams of the reference's World (broadphase, narrowphase, constraint solver) on
// a caller's device buffers and stream, so a host that keeps the reference's World can replace one
// stage at a time.  A pe_stage is a pe_world used as workspace: each call loads the caller's
// records into its body and pair stores, runs the same kernels as pe_world_step, and writes the
// results back to the caller's buffers; nothing is synchronised (the caller waits on its stream).
namespace {

// Collision records of the caller ([n][CBW]: basis, origin, AABB scale, shape parameters, shape
// type and kinematic flag as int bits, friction, restitution) into the body store; the shape type
// and kinematic flag also into the int fields the pair dispatch and the colouring read.
__global__ void k_stage_bodies(Dev d, const float* __restrict__ body, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* s = reinterpret_cast<const float4*>(body + (size_t)i * CBW);
        float4* o = reinterpret_cast<float4*>(d.b.cbr(i));
#pragma unroll
        for (int k = 0; k < CBW / 4; k++) o[k] = s[k];
        d.b.iat(I_TYPE, i) = __float_as_int(body[(size_t)i * CBW + CB_TYPE]);
        d.b.iat(I_KIN, i) = __float_as_int(body[(size_t)i * CBW + CB_KIN]) != 0 ? 1 : 0;
    }
}
// Solver records of the caller ([n][SBW]: inverse mass, kinematic flag, world inverse inertia,
// velocities; the temp velocities are ignored) into the body store, with the temp velocities set
// as SequentialImpulseSolver::setupSolver's clearTempVelocity sets them (the velocities, zero for a
// kinematic body).
__global__ void k_stage_solver_in(Dev d, const float* __restrict__ sb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float r[SBW];
        ld_rec<SBW / 4>(sb + (size_t)i * SBW, r);
        const bool kin = r[SB_KIN] != 0.0f;
        for (int k = 0; k < 3; k++) {
            r[SB_TV + k] = kin ? 0.0f : r[SB_V + k];
            r[SB_TW + k] = kin ? 0.0f : r[SB_W + k];
        }
        r[SB_KIN] = kin ? 1.0f : 0.0f;
        if (kin) r[SB_IM] = 0.0f;
        st_rec<SBW / 4>(d.b.sbr(i), r);
        d.b.iat(I_KIN, i) = kin ? 1 : 0;
    }
}
// The solved velocities (syncTempVelocity) into the caller's solver records.
__global__ void k_stage_solver_out(Dev d, float* __restrict__ sb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (d.b.iat(I_KIN, i)) continue;
        const float* t = d.b.sbr(i);
        float* o = sb + (size_t)i * SBW;
        for (int k = 0; k < 3; k++) {
            o[SB_V + k] = t[SB_TV + k];
            o[SB_W + k] = t[SB_TW + k];
        }
    }
}
// Pairs (and, for the solver, point counts and contact slots; for the narrowphase, the distances
// of the slots that survive from the previous call) into the pair store.
__global__ void k_stage_pairs(Dev d, const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, int np, int prev,
                              const int32_t* __restrict__ npts, const float* __restrict__ points) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.cnt->npairs = np;
        d.cnt->prev_npairs = prev;
        d.cnt->nshape = 0;
        d.cnt->nsat = 0;
        d.cnt->npoints = 0;
        d.cnt->overflow = 0;
    }
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        d.ps.pa[p] = pa[p];
        d.ps.pb[p] = pb[p];
        if (npts) d.ps.npts[p] = npts[p];
        if (!points) continue;
        for (int k = 0; k < 8; k++) {
            const float* c = points + ((size_t)p * 8 + k) * CF;
            const size_t s = (size_t)p * 8 + k;
            if (npts) {
                if (k < npts[p])
                    for (int f = 0; f < CW; f++) d.ps.cp[s * CW + f] = c[f];
                d.ps.cd[s] = c[12];
            } else if (p < prev) {
                d.ps.cd[s] = c[12];
            }
        }
    }
}
// Contact results of the narrowphase into the caller's slots ([np][8][CF]: valid slots in cache
// order, invalid ones with the position at PE_REAL_MAX and their distance kept), point counts and
// swap flags.
__global__ void k_stage_contacts_out(Dev d, int np, float* __restrict__ points, int32_t* __restrict__ npts,
                                     int32_t* __restrict__ swap) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const int m = d.ps.npts[p];
        if (npts) npts[p] = m;
        if (swap) swap[p] = d.ps.swap[p] & 1;
        for (int k = 0; k < 8; k++) {
            const size_t s = (size_t)p * 8 + k;
            float* o = points + s * CF;
            for (int f = 0; f < CW; f++) o[f] = k < m ? d.ps.cp[s * CW + f] : (f < 3 ? REAL_MAX : 0.0f);
            o[12] = d.ps.cd[s];
        }
    }
}
__global__ void k_stage_impulses_out(Dev d, int np, float* __restrict__ imp) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const int m = d.ps.npts[p];
        for (int k = 0; k < 8; k++)
            for (int q = 0; q < 3; q++)
                imp[((size_t)p * 8 + k) * 3 + q] = k < m ? d.ps.rows[((size_t)p * 8 + k) * RW + R_LN + q] : 0.0f;
    }
}
// result[0] = pairs the sweep found (more than the capacity when it overflowed), result[1] = the
// next axis, result[2] = capacity-error bits.
__global__ void k_stage_sap_result(SapView s, int32_t* result) {
    result[0] = (int32_t)s.cnt->nchunk_pairs;
    result[1] = s.cnt->axis;
    result[2] = s.cnt->overflow;
}
__global__ void k_stage_result(Dev d, int32_t* result, int which) {
    result[0] = which == 0 ? d.cnt->overflow : d.cnt->ncolors;
    if (which == 1) result[1] = d.cnt->overflow;
}

// A sweep over a caller's AABB records ([n][8]) with the given working memory and pair outputs.
SapView sap_over_records(const SapView& work, int n, const float* aabb, int32_t* pa, int32_t* pb, int cap) {
    SapView s = work;
    s.n = n;
    s.amin = aabb;
    s.amax = aabb + 4;
    s.fs = 1;
    s.bs = 8;
    s.rec = aabb;
    s.pa = pa;
    s.pb = pb;
    s.pcap = cap;
    return s;
}

int stage_grid(int work, int block) { return std::max(1, std::min((work + block - 1) / block, 4096)); }

}  // namespace

struct pe_stage {
    pe_world* w = nullptr;
    int max_bodies = 0, max_pairs = 0;
    ~pe_stage() { delete w; }
};

extern "C" {

int pe_stage_create(int32_t max_bodies, int32_t max_pairs, pe_stage** out) {
    if (!out || max_bodies < 1 || max_pairs < 1) { pe_set_error("pe_stage_create: invalid argument"); return PE_ERR_INVALID; }
    // the workspace is a world of max_bodies unit spheres (their state is overwritten by each call)
    std::vector<int32_t> type(max_bodies, PE_SHAPE_SPHERE), kin(max_bodies, 1);
    std::vector<double> param(4 * (size_t)max_bodies, 0.0), mass(max_bodies, 1.0), zero3(3 * (size_t)max_bodies, 0.0),
        rot(9 * (size_t)max_bodies, 0.0), mu(max_bodies, 0.5), e(max_bodies, 0.0);
    for (int i = 0; i < max_bodies; i++) {
        param[4 * (size_t)i] = 1.0;
        rot[9 * (size_t)i] = rot[9 * (size_t)i + 4] = rot[9 * (size_t)i + 8] = 1.0;
    }
    pe_scene_desc sc{};
    sc.n = max_bodies;
    sc.shape_type = type.data();
    sc.shape_param = param.data();
    sc.mass = mass.data();
    sc.position = zero3.data();
    sc.rotation = rot.data();
    sc.linear_velocity = zero3.data();
    sc.angular_velocity = zero3.data();
    sc.friction = mu.data();
    sc.restitution = e.data();
    sc.kinematic = kin.data();
    pe_world_params prm{1.0 / 60.0, {0.0, 0.0, 0.0}, 10};
    auto* S = new pe_stage();
    if (int rc = pe_world_create(&sc, &prm, max_pairs, &S->w)) { delete S; return rc; }
    S->w->has_shapes = true;
    S->max_bodies = max_bodies;
    S->max_pairs = max_pairs;
    *out = S;
    return PE_OK;
}

int pe_stage_destroy(pe_stage* s) {
    delete s;
    return PE_OK;
}

int pe_stage_set_meshes(pe_stage* s, int32_t n_meshes, const pe_mesh_desc* meshes, const int32_t* concave,
                        int32_t prepared) {
    if (!s || n_meshes < 0 || (n_meshes > 0 && (!meshes || !concave))) { pe_set_error("invalid argument"); return PE_ERR_INVALID; }
    HostMeshes hm;
    for (int k = 0; k < n_meshes; k++)
        if (hm.add(meshes[k], concave[k] != 0, nullptr, nullptr, prepared != 0) < 0) return PE_ERR_INVALID;
    pe_world* w = s->w;
    PE_HIP(hipStreamSynchronize(w->stream));
    w->meshes.release();
    if (int rc = w->meshes.upload(hm, w->stream)) return rc;
    PE_HIP(hipStreamSynchronize(w->stream));
    w->d.b.ms = w->meshes.view();
    return PE_OK;
}

int pe_stage_broadphase(pe_stage* s, int32_t n, const float* aabb, int32_t axis, int32_t* pair_a, int32_t* pair_b,
                        int32_t cap, int32_t* result, pe_stream_t stream) {
    if (!s || n < 0 || n > s->max_bodies || axis < 0 || axis > 2 || cap < 0 || !result || (n > 0 && !aabb) ||
        (cap > 0 && (!pair_a || !pair_b))) {
        pe_set_error("pe_stage_broadphase: invalid argument");
        return PE_ERR_INVALID;
    }
    pe_clear_stale_error();
    hipStream_t st = (hipStream_t)stream;
    const SapView v = sap_over_records(s->w->d.sap, n, aabb, pair_a, pair_b, cap);
    hipLaunchKernelGGL(k_sap_begin, dim3(1), dim3(1), 0, st, v, (int)axis);
    if (n >= 2) sap_launch(v, st);   // fewer than 2 bodies: no pairs, the axis stays (broad_phase_sweep_and_prune.cpp:8)
    hipLaunchKernelGGL(k_stage_sap_result, dim3(1), dim3(1), 0, st, v, result);
    PE_HIP(hipGetLastError());
    return PE_OK;
}

int pe_stage_narrowphase(pe_stage* s, int32_t n, const float* body, int32_t npairs, const int32_t* pair_a,
                         const int32_t* pair_b, int32_t prev_npairs, float* points, int32_t* npoints, int32_t* swap,
                         int32_t* result, pe_stream_t stream) {
    if (!s || n < 1 || n > s->max_bodies || npairs < 0 || npairs > s->max_pairs || prev_npairs < 0 ||
        prev_npairs > s->max_pairs || !body || !result || (npairs > 0 && (!pair_a || !pair_b || !points))) {
        pe_set_error("pe_stage_narrowphase: invalid argument");
        return PE_ERR_INVALID;
    }
    pe_clear_stale_error();
    pe_world* w = s->w;
    hipStream_t keep = w->stream;
    w->stream = (hipStream_t)stream;
    w->n = n;
    w->d.b.n = n;
    hipLaunchKernelGGL(k_stage_bodies, dim3(stage_grid(n, 256)), dim3(256), 0, w->stream, w->d, body, (int)n);
    hipLaunchKernelGGL(k_stage_pairs, dim3(stage_grid(npairs, 256)), dim3(256), 0, w->stream, w->d, pair_a, pair_b,
                       (int)npairs, (int)prev_npairs, (const int32_t*)nullptr, (const float*)points);
    w->launch_narrowphase();
    if (npairs > 0)
        hipLaunchKernelGGL(k_stage_contacts_out, dim3(stage_grid(npairs, 128)), dim3(128), 0, w->stream, w->d,
                           (int)npairs, points, npoints, swap);
    hipLaunchKernelGGL(k_stage_result, dim3(1), dim3(1), 0, w->stream, w->d, result, 0);
    w->stream = keep;
    PE_HIP(hipGetLastError());
    return PE_OK;
}

int pe_stage_solve(pe_stage* s, int32_t n, const float* body, float* solver_body, int32_t npairs, const int32_t* pair_a,
                   const int32_t* pair_b, const int32_t* npoints, const float* points, float dt, int32_t iterations,
                   float* impulses, int32_t* result, pe_stream_t stream) {
    if (!s || n < 1 || n > s->max_bodies || npairs < 0 || npairs > s->max_pairs || !body || !solver_body || !result ||
        iterations < 0 || (npairs > 0 && (!pair_a || !pair_b || !npoints || !points))) {
        pe_set_error("pe_stage_solve: invalid argument");
        return PE_ERR_INVALID;
    }
    pe_clear_stale_error();
    pe_world* w = s->w;
    hipStream_t keep = w->stream;
    w->stream = (hipStream_t)stream;
    w->n = n;
    w->d.b.n = n;
    w->d.dt = dt;
    w->d.iterations = iterations;
    w->gs_blocks = default_gs_blocks(n, w->gs_max_blocks);
    hipLaunchKernelGGL(k_stage_bodies, dim3(stage_grid(n, 256)), dim3(256), 0, w->stream, w->d, body, (int)n);
    hipLaunchKernelGGL(k_stage_solver_in, dim3(stage_grid(n, 256)), dim3(256), 0, w->stream, w->d, (const float*)solver_body, (int)n);
    hipLaunchKernelGGL(k_stage_pairs, dim3(stage_grid(npairs, 256)), dim3(256), 0, w->stream, w->d, pair_a, pair_b,
                       (int)npairs, 0, npoints, points);
    int rc = w->launch_solve();
    if (rc == PE_OK) {
        hipLaunchKernelGGL(k_stage_solver_out, dim3(stage_grid(n, 256)), dim3(256), 0, w->stream, w->d, solver_body, (int)n);
        if (impulses && npairs > 0)
            hipLaunchKernelGGL(k_stage_impulses_out, dim3(stage_grid(npairs, 128)), dim3(128), 0, w->stream, w->d,
                               (int)npairs, impulses);
        hipLaunchKernelGGL(k_stage_result, dim3(1), dim3(1), 0, w->stream, w->d, result, 1);
    }
    w->stream = keep;
    if (rc != PE_OK) return rc;
    PE_HIP(hipGetLastError());
    return PE_OK;
}

}  // extern "C"
