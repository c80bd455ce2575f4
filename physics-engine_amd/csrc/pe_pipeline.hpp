// Per-body and per-pair stage routines shared by the single-world kernels (pe_world.hip) and
// the fused per-world kernels (pe_batch.hip).  Both pipelines therefore compute every world
// with exactly the same arithmetic.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_prims.hpp"

namespace pe {

// Pair-indexed storage: pairs, the 8 contact slots per pair, solver rows per slot.
struct PairStore {
    int32_t* pa;      // body index of A
    int32_t* pb;      // body index of B
    float* cp;        // [CF][cs]
    int32_t* npts;
    int32_t* swap;
    float* rows;      // [RF][cs]
    size_t cs;        // slot stride = pair capacity * 8
};

__device__ __forceinline__ f3 ld3(const BodyView& b, int f, int i) { return f3{b.at(f, i), b.at(f + 1, i), b.at(f + 2, i)}; }
__device__ __forceinline__ void st3(const BodyView& b, int f, int i, f3 v) { b.at(f, i) = v.x; b.at(f + 1, i) = v.y; b.at(f + 2, i) = v.z; }
__device__ __forceinline__ m33 ld9(const BodyView& b, int f, int i) {
    m33 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.m[k] = b.at(f + k, i);
    return r;
}
__device__ __forceinline__ void st9(const BodyView& b, int f, int i, const m33& r) {
#pragma unroll
    for (int k = 0; k < 9; k++) b.at(f + k, i) = r.m[k];
}

// ------------------------------------------------------------------ stage 0
// World::updateObjectStatus (world.cpp:35-119) with zero sleep thresholds (a body never
// sleeps), World::applyExternalForce (world.cpp:121-127), and the solver's clearTempVelocity
// (sequential_impulse_solver.cpp:17-20) for body i.
__device__ __forceinline__ void integrate_body(const BodyView& b, int i, float dt, const float g[3]) {
    const int type = b.iat(I_TYPE, i);
    const bool kin = b.iat(I_KIN, i) != 0;
    f3 pos = ld3(b, F_POS, i);
    m33 R = ld9(b, F_ROT, i);
    const float sp[4] = {b.at(F_SP, i), b.at(F_SP + 1, i), b.at(F_SP + 2, i), b.at(F_SP + 3, i)};
    if (!kin) {
        f3 v = ld3(b, F_V, i), w = ld3(b, F_W, i);
        integrate_pose(pos, R, v, w, dt);
        const m33 L = ld9(b, F_LINV, i);
        st9(b, F_IW, i, mmt(mm(R, L), R));
        st3(b, F_POS, i, pos);
        st9(b, F_ROT, i, R);
        // RigidBody::applyDamping (rigidbody.cpp:201-205)
        const float pl = powf(1.0f - b.at(F_LD, i), dt), pw = powf(1.0f - b.at(F_AD, i), dt);
        v = mul(v, pl);
        w = mul(w, pw);
        // RigidBody::applyForce with force = gravity * mass (rigidbody.cpp:193-199)
        const float m = b.at(F_MASS, i), im = b.at(F_IM, i);
        const f3 f{g[0] * m, g[1] * m, g[2] * m};
        v = add(v, mul(mul(f, im), dt));
        st3(b, F_V, i, v);
        st3(b, F_W, i, w);
        st3(b, F_TV, i, v);
        st3(b, F_TW, i, w);
    } else {
        st3(b, F_TV, i, f3{0, 0, 0});
        st3(b, F_TW, i, f3{0, 0, 0});
    }
    f3 mn, mx;
    shape_aabb(type, sp, R, pos, mn, mx);
    st3(b, F_AMIN, i, mn);
    st3(b, F_AMAX, i, mx);
    b.at(F_SCALE, i) = aabb_scale(mn, mx, pos);
}

// AABB, scale and world inverse inertia of the current pose (RigidBody::setTransform).
__device__ __forceinline__ void refresh_body(const BodyView& b, int i) {
    const f3 pos = ld3(b, F_POS, i);
    const m33 R = ld9(b, F_ROT, i);
    const float sp[4] = {b.at(F_SP, i), b.at(F_SP + 1, i), b.at(F_SP + 2, i), b.at(F_SP + 3, i)};
    f3 mn, mx;
    shape_aabb(b.iat(I_TYPE, i), sp, R, pos, mn, mx);
    st3(b, F_AMIN, i, mn);
    st3(b, F_AMAX, i, mx);
    b.at(F_SCALE, i) = aabb_scale(mn, mx, pos);
    st9(b, F_IW, i, mmt(mm(R, ld9(b, F_LINV, i)), R));
}

// ------------------------------------------------------------------ stage 1
// BroadPhaseBase::validateCollisionPair (broad_phase_base.cpp:5-19)
__device__ __forceinline__ bool bp_valid(const BodyView& b, int i, int j) {
    if (b.iat(I_KIN, i) && b.iat(I_KIN, j)) return false;
    if (b.at(F_AMIN, i) > b.at(F_AMAX, j) || b.at(F_AMIN, j) > b.at(F_AMAX, i)) return false;
    if (b.at(F_AMIN + 1, i) > b.at(F_AMAX + 1, j) || b.at(F_AMIN + 1, j) > b.at(F_AMAX + 1, i)) return false;
    if (b.at(F_AMIN + 2, i) > b.at(F_AMAX + 2, j) || b.at(F_AMIN + 2, j) > b.at(F_AMAX + 2, i)) return false;
    return true;
}

// ------------------------------------------------------------------ stage 2
__device__ __forceinline__ void load_pose(const BodyView& b, int i, Pose& P) {
    P.R = ld9(b, F_ROT, i);
    P.p = ld3(b, F_POS, i);
    P.scale = b.at(F_SCALE, i);
}

// SimpleNarrowPhase::calcContactResults (simple_narrow_phase.cpp:7-41) for pair slot p.
// Slots that existed in the previous step (p < prev_np) keep their distances, because the
// reference reuses its ContactResult objects and clearContactPoints only invalidates
// positions (contact_result.cpp:158-164); newer slots start from a fresh ContactResult.
static __device__ __noinline__ void narrow_pair(const BodyView& b, const PairStore& ps, int p, int prev_np) {
    const int ia = ps.pa[p], ib = ps.pb[p];
    Cache c;
    const size_t slot0 = (size_t)p * 8;
    const size_t CS = ps.cs;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        c.wx[k] = REAL_MAX; c.wy[k] = REAL_MAX; c.wz[k] = REAL_MAX;
        c.nx[k] = 0; c.ny[k] = 0; c.nz[k] = 0;
        c.ax[k] = 0; c.ay[k] = 0; c.az[k] = 0;
        c.bx[k] = 0; c.by[k] = 0; c.bz[k] = 0;
        c.d[k] = p < prev_np ? ps.cp[C_D * CS + slot0 + k] : REAL_MAX;
    }
    PairCtx x;
    load_pose(b, ia, x.A);
    load_pose(b, ib, x.B);
    x.swap = false;
    {
        const float sa = x.A.scale, sb = x.B.scale;
        const float th = (sa < sb ? sa : sb) * 0.02f;
        x.same_th2 = th * th;
    }
    const int ta = b.iat(I_TYPE, ia), tb = b.iat(I_TYPE, ib);
    const float spa[3] = {b.at(F_SP, ia), b.at(F_SP + 1, ia), b.at(F_SP + 2, ia)};
    const float spb[3] = {b.at(F_SP, ib), b.at(F_SP + 1, ib), b.at(F_SP + 2, ib)};
    if (ta == SHAPE_BOX && tb == SHAPE_BOX) box_box(c, x, spa, spb);
    else if (ta == SHAPE_BOX && tb == SHAPE_SPHERE) box_sphere(c, x, true, spa, spb[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_BOX) box_sphere(c, x, false, spb, spa[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) sphere_sphere(c, x, spa[0], spb[0]);
    int rank[8];
    const int np = sort_rank(c, rank);
    ps.npts[p] = np;
    ps.swap[p] = x.swap ? 1 : 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const size_t s = slot0 + rank[k];
        float* cp = ps.cp;
        cp[(C_WPOS + 0) * CS + s] = c.wx[k]; cp[(C_WPOS + 1) * CS + s] = c.wy[k]; cp[(C_WPOS + 2) * CS + s] = c.wz[k];
        cp[(C_N + 0) * CS + s] = c.nx[k]; cp[(C_N + 1) * CS + s] = c.ny[k]; cp[(C_N + 2) * CS + s] = c.nz[k];
        cp[(C_LA + 0) * CS + s] = c.ax[k]; cp[(C_LA + 1) * CS + s] = c.ay[k]; cp[(C_LA + 2) * CS + s] = c.az[k];
        cp[(C_LB + 0) * CS + s] = c.bx[k]; cp[(C_LB + 1) * CS + s] = c.by[k]; cp[(C_LB + 2) * CS + s] = c.bz[k];
        cp[C_D * CS + s] = c.d[k];
    }
}

// ------------------------------------------------------------------ stage 3
// FrictionContactConstraint::initSequentialImpulse (friction_contact_constraint.cpp:9-73) for
// every point of pair p.
__device__ __forceinline__ void setup_pair(const BodyView& b, const PairStore& ps, int p, float dt) {
    const int n = ps.npts[p];
    if (n == 0) return;
    const size_t CS = ps.cs;
    const int ia = ps.pa[p], ib = ps.pb[p];
    const bool ka = b.iat(I_KIN, ia) != 0, kb = b.iat(I_KIN, ib) != 0;
    const m33 RA = ld9(b, F_ROT, ia), RB = ld9(b, F_ROT, ib);
    m33 Ia, Ib;
    if (ka) { for (int k = 0; k < 9; k++) Ia.m[k] = 0.0f; } else Ia = ld9(b, F_IW, ia);
    if (kb) { for (int k = 0; k < 9; k++) Ib.m[k] = 0.0f; } else Ib = ld9(b, F_IW, ib);
    const float ims = (ka ? 0.0f : b.at(F_IM, ia)) + (kb ? 0.0f : b.at(F_IM, ib));
    const f3 va = ld3(b, F_V, ia), wa = ld3(b, F_W, ia), vb = ld3(b, F_V, ib), wb = ld3(b, F_W, ib);
    // ContactResult::setObjects (contact_result.cpp:80-86)
    const float mu = sqrtf(b.at(F_MU, ia) * b.at(F_MU, ib));
    const float e = sqrtf(b.at(F_E, ia) * b.at(F_E, ib));
    float* R = ps.rows;
    const float* cp = ps.cp;
    for (int k = 0; k < n; k++) {
        const size_t s = (size_t)p * 8 + k;
        const f3 la{cp[(C_LA) * CS + s], cp[(C_LA + 1) * CS + s], cp[(C_LA + 2) * CS + s]};
        const f3 lb{cp[(C_LB) * CS + s], cp[(C_LB + 1) * CS + s], cp[(C_LB + 2) * CS + s]};
        const f3 nn{cp[(C_N) * CS + s], cp[(C_N + 1) * CS + s], cp[(C_N + 2) * CS + s]};
        const Row r = make_row(RA, RB, la, lb, nn, cp[C_D * CS + s], ims, Ia, Ib, va, wa, vb, wb, mu, e, dt);
        R[(R_RA) * CS + s] = r.ra.x; R[(R_RA + 1) * CS + s] = r.ra.y; R[(R_RA + 2) * CS + s] = r.ra.z;
        R[(R_RB) * CS + s] = r.rb.x; R[(R_RB + 1) * CS + s] = r.rb.y; R[(R_RB + 2) * CS + s] = r.rb.z;
        R[(R_N) * CS + s] = r.n.x; R[(R_N + 1) * CS + s] = r.n.y; R[(R_N + 2) * CS + s] = r.n.z;
        R[(R_T0) * CS + s] = r.t0.x; R[(R_T0 + 1) * CS + s] = r.t0.y; R[(R_T0 + 2) * CS + s] = r.t0.z;
        R[(R_T1) * CS + s] = r.t1.x; R[(R_T1 + 1) * CS + s] = r.t1.y; R[(R_T1 + 2) * CS + s] = r.t1.z;
        R[R_RHS * CS + s] = r.rhs; R[R_MN * CS + s] = r.mn; R[R_MT0 * CS + s] = r.mt0; R[R_MT1 * CS + s] = r.mt1;
        R[R_MU * CS + s] = r.mu; R[R_LN * CS + s] = 0.0f; R[R_LT0 * CS + s] = 0.0f; R[R_LT1 * CS + s] = 0.0f;
    }
}

__device__ __forceinline__ Row load_row(const float* R, size_t CS, size_t s) {
    Row r;
    r.ra = f3{R[(R_RA) * CS + s], R[(R_RA + 1) * CS + s], R[(R_RA + 2) * CS + s]};
    r.rb = f3{R[(R_RB) * CS + s], R[(R_RB + 1) * CS + s], R[(R_RB + 2) * CS + s]};
    r.n = f3{R[(R_N) * CS + s], R[(R_N + 1) * CS + s], R[(R_N + 2) * CS + s]};
    r.t0 = f3{R[(R_T0) * CS + s], R[(R_T0 + 1) * CS + s], R[(R_T0 + 2) * CS + s]};
    r.t1 = f3{R[(R_T1) * CS + s], R[(R_T1 + 1) * CS + s], R[(R_T1 + 2) * CS + s]};
    r.rhs = R[R_RHS * CS + s]; r.mn = R[R_MN * CS + s]; r.mt0 = R[R_MT0 * CS + s]; r.mt1 = R[R_MT1 * CS + s];
    r.mu = R[R_MU * CS + s];
    return r;
}

// RigidBody::applyTempImpulse (rigidbody.cpp:166-170), with the body's inverse mass and world
// inverse inertia already in registers.
__device__ __forceinline__ void apply_temp(f3& tv, f3& tw, float im, const m33& Iw, f3 r, f3 J) {
    tv = add(tv, mul(J, im));
    tw = add(tw, mv(Iw, cross(r, J)));
}

// FrictionContactConstraint::iterateSequentialImpulse (friction_contact_constraint.cpp:75-108)
// over all rows of pair p.  All loads of a group of up to 4 rows and of both bodies are issued
// before the dependent chain, so a pair costs a few memory round trips, not one per row.
__device__ __forceinline__ void iterate_pair(const BodyView& b, const PairStore& ps, int p) {
    const size_t CS = ps.cs;
    const int ia = ps.pa[p], ib = ps.pb[p];
    const int n = ps.npts[p];
    const bool ka = b.iat(I_KIN, ia) != 0, kb = b.iat(I_KIN, ib) != 0;
    f3 tva = ld3(b, F_TV, ia), twa = ld3(b, F_TW, ia), tvb = ld3(b, F_TV, ib), twb = ld3(b, F_TW, ib);
    const m33 IwA = ld9(b, F_IW, ia), IwB = ld9(b, F_IW, ib);
    const float imA = b.at(F_IM, ia), imB = b.at(F_IM, ib);
    float* R = ps.rows;
    for (int k0 = 0; k0 < n; k0 += 4) {
        Row rw[4];
        float ln[4], lt0[4], lt1[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                const size_t s = (size_t)p * 8 + k0 + u;
                rw[u] = load_row(R, CS, s);
                ln[u] = R[R_LN * CS + s]; lt0[u] = R[R_LT0 * CS + s]; lt1[u] = R[R_LT1 * CS + s];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                const f3 J = iterate_row(rw[u], ln[u], lt0[u], lt1[u], tva, twa, tvb, twb);
                if (!ka) apply_temp(tva, twa, imA, IwA, rw[u].ra, J);
                if (!kb) apply_temp(tvb, twb, imB, IwB, rw[u].rb, neg(J));
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                const size_t s = (size_t)p * 8 + k0 + u;
                R[R_LN * CS + s] = ln[u]; R[R_LT0 * CS + s] = lt0[u]; R[R_LT1 * CS + s] = lt1[u];
            }
        }
    }
    if (!ka) { st3(b, F_TV, ia, tva); st3(b, F_TW, ia, twa); }
    if (!kb) { st3(b, F_TV, ib, tvb); st3(b, F_TW, ib, twb); }
}

// SequentialImpulseSolver::solve's syncTempVelocity (sequential_impulse_solver.cpp:71-80)
__device__ __forceinline__ void sync_body(const BodyView& b, int i) {
    if (b.iat(I_KIN, i)) return;
    st3(b, F_V, i, ld3(b, F_TV, i));
    st3(b, F_W, i, ld3(b, F_TW, i));
}

// ------------------------------------------------------------------ constraint-graph colouring
// Colours pairs so that no two pairs of one colour share a dynamic body (kinematic bodies never
// change, rigidbody.cpp:167, so they impose no order).  Colour c is a maximal independent set
// built in sub-rounds: every candidate claims its dynamic bodies with atomicMin of its pair
// index; a candidate that holds all its claims joins the colour and takes its bodies; a
// candidate touching a taken body waits for a later colour.  The colour classes depend only on
// the pair set, not on list order, so sweeping a colour in parallel equals the sequential
// Gauss-Seidel sweep over (colour, any order within the colour).
// claim[] holds, per body, INT_MAX (free), -1 (taken by this colour) or the lowest claim.
namespace colour {
constexpr int FREE = 0x7fffffff;
constexpr int TAKEN = -1;
}

template <bool LDS>
__device__ __forceinline__ int claim_get(int* claim, int i) {
    if (LDS) return claim[i];
    return __hip_atomic_load(&claim[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Appends the lanes of the calling wave with pred set to a list whose length is *counter:
// one atomic per wave.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(bool pred, int* counter) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if ((int)(threadIdx.x & (WAVE - 1)) == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, WAVE);
    return pred ? base + (int)__popcll(m & lanemask_lt()) : -1;
}

// bufs: three lists of pair capacity; bufs[0] holds the nact active pairs on entry.
// claim is indexed by (body - off).  Writes the solve order and colour starts; returns colours.
// The lists are compacted with atomics, so the order inside a colour varies from run to run;
// the colour classes do not, and neither does any result (pairs of a colour are independent).
template <bool LDS>
__device__ int colour_pairs(const BodyView& b, const PairStore& ps, int* bufs[3], int nact, int* order, int* cs,
                            int* claim, int off, int* wsum) {
    using namespace colour;
    __shared__ int cnt[3];   // winners (order), deferred, staying candidates
    const int tid = threadIdx.x, B = blockDim.x;
    (void)wsum;
    int ia = 0, ncol = 0;
    if (tid == 0) cnt[0] = 0;
    __syncthreads();
    while (nact > 0) {
        int* cand = bufs[ia];
        int* nxt = bufs[(ia + 1) % 3];
        int* def = bufs[(ia + 2) % 3];
        int ncand = nact;
        const int col_begin = cnt[0];
        if (tid == 0) { cs[ncol] = col_begin; cnt[1] = 0; }
        __syncthreads();
        while (ncand > 0) {
            for (int q = tid; q < ncand; q += B) {
                const int p = cand[q];
                const int a = ps.pa[p], c = ps.pb[p];
                if (!b.iat(I_KIN, a)) atomicMin(&claim[a - off], p);
                if (!b.iat(I_KIN, c)) atomicMin(&claim[c - off], p);
            }
            if (tid == 0) cnt[2] = 0;
            __syncthreads();
            const int sub_begin = cnt[0];
            for (int base = 0; base < ncand; base += B) {
                const int q = base + tid;
                bool win = false, blk = false, stay = false;
                int p = -1;
                if (q < ncand) {
                    p = cand[q];
                    const int a = ps.pa[p], c = ps.pb[p];
                    const bool ka = b.iat(I_KIN, a) != 0, kc = b.iat(I_KIN, c) != 0;
                    const int ca = ka ? 0 : claim_get<LDS>(claim, a - off);
                    const int cc = kc ? 0 : claim_get<LDS>(claim, c - off);
                    blk = (!ka && ca == TAKEN) || (!kc && cc == TAKEN);
                    win = !blk && (ka || ca == p) && (kc || cc == p);
                    stay = !blk && !win;
                }
                const int iw = wave_append(win, &cnt[0]);
                const int ib = wave_append(blk, &cnt[1]);
                const int is = wave_append(stay, &cnt[2]);
                if (win) order[iw] = p;
                if (blk) def[ib] = p;
                if (stay) nxt[is] = p;
            }
            __syncthreads();
            const int sub_end = cnt[0], nstay = cnt[2];
            // clear this sub-round's claims (taken bodies stay taken)
            for (int q = tid; q < ncand; q += B) {
                const int p = cand[q];
                const int a = ps.pa[p], c = ps.pb[p];
                if (!b.iat(I_KIN, a) && claim_get<LDS>(claim, a - off) != TAKEN) claim[a - off] = FREE;
                if (!b.iat(I_KIN, c) && claim_get<LDS>(claim, c - off) != TAKEN) claim[c - off] = FREE;
            }
            __syncthreads();
            for (int q = sub_begin + tid; q < sub_end; q += B) {
                const int p = order[q];
                const int a = ps.pa[p], c = ps.pb[p];
                if (!b.iat(I_KIN, a)) claim[a - off] = TAKEN;
                if (!b.iat(I_KIN, c)) claim[c - off] = TAKEN;
            }
            __syncthreads();
            int* t = cand; cand = nxt; nxt = t;
            ncand = nstay;
        }
        const int col_end = cnt[0];
        for (int q = col_begin + tid; q < col_end; q += B) {
            const int p = order[q];
            claim[ps.pa[p] - off] = FREE;
            claim[ps.pb[p] - off] = FREE;
        }
        const int ndef = cnt[1];
        __syncthreads();
        ia = (ia + 2) % 3;
        nact = ndef;
        ncol++;
    }
    if (tid == 0) cs[ncol] = cnt[0];
    __syncthreads();
    return ncol;
}

}  // namespace pe
