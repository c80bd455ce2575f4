// Per-body and per-pair stage routines shared by the single-world kernels (pe_world.hip) and
// the fused per-world kernels (pe_batch.hip).  Both pipelines therefore compute every world
// with exactly the same arithmetic.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"

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
    const int np = sort_points(c);
    ps.npts[p] = np;
    ps.swap[p] = x.swap ? 1 : 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const size_t s = slot0 + k;
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

// RigidBody::applyTempImpulse (rigidbody.cpp:166-170)
__device__ __forceinline__ void apply_temp(const BodyView& b, int i, f3& tv, f3& tw, f3 r, f3 J) {
    tv = add(tv, mul(J, b.at(F_IM, i)));
    tw = add(tw, mv(ld9(b, F_IW, i), cross(r, J)));
}

// FrictionContactConstraint::iterateSequentialImpulse (friction_contact_constraint.cpp:75-108)
// over all rows of pair p.
__device__ __forceinline__ void iterate_pair(const BodyView& b, const PairStore& ps, int p) {
    const size_t CS = ps.cs;
    const int ia = ps.pa[p], ib = ps.pb[p];
    const bool ka = b.iat(I_KIN, ia) != 0, kb = b.iat(I_KIN, ib) != 0;
    f3 tva = ld3(b, F_TV, ia), twa = ld3(b, F_TW, ia), tvb = ld3(b, F_TV, ib), twb = ld3(b, F_TW, ib);
    const int n = ps.npts[p];
    float* R = ps.rows;
    for (int k = 0; k < n; k++) {
        const size_t s = (size_t)p * 8 + k;
        const Row r = load_row(R, CS, s);
        float ln = R[R_LN * CS + s], lt0 = R[R_LT0 * CS + s], lt1 = R[R_LT1 * CS + s];
        const f3 J = iterate_row(r, ln, lt0, lt1, tva, twa, tvb, twb);
        R[R_LN * CS + s] = ln; R[R_LT0 * CS + s] = lt0; R[R_LT1 * CS + s] = lt1;
        if (!ka) apply_temp(b, ia, tva, twa, r.ra, J);
        if (!kb) apply_temp(b, ib, tvb, twb, r.rb, neg(J));
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

}  // namespace pe
