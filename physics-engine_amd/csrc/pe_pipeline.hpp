// Per-body and per-pair stage routines shared by the single-world kernels (pe_world.hip) and
// the fused per-world kernels (pe_batch.hip).  Both pipelines therefore compute every world
// with exactly the same arithmetic.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_prims.hpp"
#include "pe_shapes.hpp"

namespace pe {

// Pair-indexed storage: pairs, the 8 contact slots per pair, solver rows per slot.
struct PairStore {
    int32_t* pa;      // body index of A
    int32_t* pb;      // body index of B
    float* cp;        // [cap*8][CW] contact slot records
    float* cd;        // [cap*8] contact distances
    int32_t* npts;
    int32_t* swap;
    float* rows;      // [cap*8][RW] solver rows
    size_t cs;        // slot capacity = pair capacity * 8
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
// Record access in float4 units (records are 16-byte aligned).
template <int N4> __device__ __forceinline__ void ld_rec(const float* src, float* dst) {
    const float4* s = reinterpret_cast<const float4*>(src);
#pragma unroll
    for (int k = 0; k < N4; k++) { const float4 v = s[k]; dst[4 * k] = v.x; dst[4 * k + 1] = v.y; dst[4 * k + 2] = v.z; dst[4 * k + 3] = v.w; }
}
template <int N4> __device__ __forceinline__ void st_rec(float* dst, const float* src) {
    float4* d = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int k = 0; k < N4; k++) d[k] = make_float4(src[4 * k], src[4 * k + 1], src[4 * k + 2], src[4 * k + 3]);
}

// Writes the collision-body, solver-body and AABB records of body i.
__device__ __forceinline__ void write_records(const BodyView& b, int i, int type, bool kin, const float sp[4], const m33& R, f3 pos,
                                              float scale, f3 mn, f3 mx, const m33& Iw, float im, f3 v, f3 w, f3 tv, f3 tw,
                                              float mu, float e) {
    float c[CBW];
#pragma unroll
    for (int k = 0; k < 9; k++) c[CB_R + k] = R.m[k];
    c[CB_P] = pos.x; c[CB_P + 1] = pos.y; c[CB_P + 2] = pos.z;
    c[CB_SCALE] = scale;
    c[CB_SP] = sp[0]; c[CB_SP + 1] = sp[1]; c[CB_SP + 2] = sp[2];
    c[CB_TYPE] = __int_as_float(type);
    c[CB_KIN] = __int_as_float(kin ? 1 : 0);
    c[CB_MU] = mu;
    c[CB_E] = e;
    st_rec<CBW / 4>(b.cbr(i), c);
    float s[SBW];
    s[SB_TV] = tv.x; s[SB_TV + 1] = tv.y; s[SB_TV + 2] = tv.z;
    s[SB_TW] = tw.x; s[SB_TW + 1] = tw.y; s[SB_TW + 2] = tw.z;
    s[SB_IM] = im;
    s[SB_KIN] = kin ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) s[SB_IW + k] = Iw.m[k];
    s[SB_V] = v.x; s[SB_V + 1] = v.y; s[SB_V + 2] = v.z;
    s[SB_W] = w.x; s[SB_W + 1] = w.y; s[SB_W + 2] = w.z;
    s[23] = 0.0f;
    st_rec<SBW / 4>(b.sbr(i), s);
    const float a[8] = {mn.x, mn.y, mn.z, kin ? 1.0f : 0.0f, mx.x, mx.y, mx.z, sp[3]};   // sp[3]: COLL_*
    st_rec<2>(b.abr(i), a);
}

// ------------------------------------------------------------------ stage 0
// Sleeping, lifetimes and neighbour counts of a single world (World::updateObjectStatus,
// world.cpp:35-119; RigidBody::step's lifetime, rigidbody.cpp:215-218; the narrowphase's neighbour
// counts, simple_narrow_phase.cpp:43-62): per body, structure of arrays.
constexpr int ST_SLEEP = 1, ST_REMOVED = 2;
struct StatusView {
    int32_t* flags;      // ST_SLEEP | ST_REMOVED
    float* sleep_time;   // seconds below the sleep thresholds
    float* life;         // lifetime (s)
    float* last;         // time lived (s)
    int32_t* scnt;       // static neighbours of the last narrowphase
    int32_t* dcnt;       // dynamic neighbours of the last narrowphase
    int32_t* report;     // 1: contacts are reported (collision callbacks)
    float lin2, ang2, tsleep;   // World::setSleepLinVel2Threshold / AngVel2 / Time
};
__device__ __forceinline__ float sqnorm3(f3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }

// A body leaves the sweep: empty AABB (min +inf, max -inf; sorts last, no candidates, not in the
// axis variance) and the ignore-collision flag.
__device__ __forceinline__ void retire_body(const BodyView& b, int i) {
    const float inf = __builtin_inff();
    st3(b, F_AMIN, i, f3{inf, inf, inf});
    st3(b, F_AMAX, i, f3{-inf, -inf, -inf});
    b.at(F_SP + 3, i) = COLL_IGNORE;
    const float a[8] = {inf, inf, inf, 0.0f, -inf, -inf, -inf, COLL_IGNORE};
    st_rec<2>(b.abr(i), a);
}

// World::updateObjectStatus (world.cpp:35-119), World::applyExternalForce (world.cpp:121-127),
// and the solver's clearTempVelocity (sequential_impulse_solver.cpp:17-20) for body i.  Without
// STATUS no body sleeps (the reference's default thresholds are zero) or ends its lifetime; with it,
// the sequential loop's rules (world.cpp:85-117): a sleeping body wakes when a velocity reaches its
// threshold scaled by (static + 1) / (static + dynamic + 1), and is neither moved nor damped; an
// awake body lives dt longer (and leaves the world at its lifetime), is moved and damped, and falls
// asleep, velocities zeroed, after tsleep seconds below the scaled thresholds.
// Returns the body's new AABB record (min, kinematic flag, max, collision flag; min +inf and max
// -inf for a body that has left the world), as written to the body's AABB record.
// REC: the records hold the body's pose, velocities and AABB between steps (the batched and served
// worlds): they are read from and written to the records only, not to the field-major store, whose
// copies of those fields then keep the values the world was created with.
struct Box3 {
    f3 mn, mx;
    float kin, coll;
};
template <bool STATUS = false, bool REC = false>
__device__ __forceinline__ Box3 integrate_body(const BodyView& b, int i, float dt, const float g[3],
                                               const StatusView* st = nullptr) {
    const float inf = __int_as_float(0x7f800000);
    const Box3 gone{f3{inf, inf, inf}, f3{-inf, -inf, -inf}, 0.0f, COLL_IGNORE};
    const int type = b.iat(I_TYPE, i);
    const bool kin = b.iat(I_KIN, i) != 0;
    int flags = 0;
    if constexpr (STATUS) {
        flags = st->flags[i];
        if (flags & ST_REMOVED) return gone;
    }
    // Every field the step reads, in one round of loads (the mass properties of a kinematic body
    // are loaded and not used): the loads are not left behind the branches that use them.
    f3 pos, v, w;
    m33 R;
    if constexpr (REC) {
        const float4* c = reinterpret_cast<const float4*>(b.cbr(i));   // basis, origin: floats 0..11
        const float4 c0 = c[0], c1 = c[1], c2 = c[2];
        R = m33{{c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x}};
        pos = f3{c2.y, c2.z, c2.w};
        const float4* q = reinterpret_cast<const float4*>(b.sbr(i));   // v, w: floats 17..22
        const float4 q4 = q[4], q5 = q[5];
        v = f3{q4.y, q4.z, q4.w};
        w = f3{q5.x, q5.y, q5.z};
    } else {
        pos = ld3(b, F_POS, i);
        R = ld9(b, F_ROT, i);
        v = ld3(b, F_V, i);
        w = ld3(b, F_W, i);
    }
    const float sp[4] = {b.at(F_SP, i), b.at(F_SP + 1, i), b.at(F_SP + 2, i), b.at(F_SP + 3, i)};
    const m33 L = ld9(b, F_LINV, i);
    const float ldamp = b.at(F_LD, i), adamp = b.at(F_AD, i), mass = b.at(F_MASS, i), imass = b.at(F_IM, i);
    const float mu = b.at(F_MU, i), e = b.at(F_E, i);
    m33 Iw;
    float im = 0.0f;
    f3 tv{0, 0, 0}, tw{0, 0, 0};   // kinematic: temp velocities stay zero (rigidbody.cpp:72-80)
    if (!kin) {
        bool awake = true;
        float ratio = 0.0f;
        if constexpr (STATUS) {
            const int sc = st->scnt[i], dc = st->dcnt[i];
            ratio = (float)(sc + 1) / (float)(dc + sc + 1);
            if (flags & ST_SLEEP) {
                awake = false;
                if (sqnorm3(v) >= st->lin2 * ratio || sqnorm3(w) >= st->ang2 * ratio) {
                    flags &= ~ST_SLEEP;
                    st->sleep_time[i] = 0.0f;
                }
            } else {
                const float lived = st->last[i] + dt;
                st->last[i] = lived;
                if (lived >= st->life[i]) {
                    st->flags[i] = flags | ST_REMOVED;
                    retire_body(b, i);
                    return gone;
                }
            }
        }
        if (awake) {
            integrate_pose(pos, R, v, w, dt);
            if constexpr (!REC) {
                st3(b, F_POS, i, pos);
                st9(b, F_ROT, i, R);
            }
        }
        Iw = mmt(mm(R, L), R);
        if (awake) {
            // RigidBody::applyDamping (rigidbody.cpp:201-205).  With zero damping the factor is
            // pow(1, dt) = 1 exactly and the product leaves every component as it is, so the power
            // is only evaluated for a damped body.
            if (ldamp != 0.0f) v = mul(v, powf(1.0f - ldamp, dt));
            if (adamp != 0.0f) w = mul(w, powf(1.0f - adamp, dt));
        }
        if constexpr (STATUS) {
            if (awake) {
                if (sqnorm3(v) < st->lin2 * ratio && sqnorm3(w) < st->ang2 * ratio) {
                    const float t = st->sleep_time[i] + dt;
                    st->sleep_time[i] = t;
                    if (t >= st->tsleep) {
                        flags |= ST_SLEEP;
                        v = f3{0.0f, 0.0f, 0.0f};
                        w = f3{0.0f, 0.0f, 0.0f};
                    }
                } else {
                    st->sleep_time[i] = 0.0f;
                }
            }
            st->flags[i] = flags;
            st->scnt[i] = 0;
            st->dcnt[i] = 0;
        }
        // RigidBody::applyForce with force = gravity * mass (rigidbody.cpp:193-199)
        im = imass;
        const f3 f{g[0] * mass, g[1] * mass, g[2] * mass};
        v = add(v, mul(mul(f, im), dt));
        if constexpr (!REC) {
            st3(b, F_V, i, v);
            st3(b, F_W, i, w);
        }
        tv = v;
        tw = w;
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) Iw.m[k] = 0.0f;
    }
    f3 mn, mx;
    if (type <= SHAPE_CYLINDER) {
        shape_aabb(type, sp, R, pos, mn, mx);
    } else if (kin) {
        // a kinematic body's pose does not change in a step, so neither does its mesh AABB
        if constexpr (REC) {
            const float4* a = reinterpret_cast<const float4*>(b.abr(i));
            const float4 a0 = a[0], a1 = a[1];
            mn = f3{a0.x, a0.y, a0.z};
            mx = f3{a1.x, a1.y, a1.z};
        } else {
            mn = ld3(b, F_AMIN, i);
            mx = ld3(b, F_AMAX, i);
        }
    } else if (type == SHAPE_COMPOUND) {
        compound_aabb(b.ms, __float_as_int(sp[0]), R, pos, mn, mx);
    } else {
        mesh_aabb(b.ms, __float_as_int(sp[0]), R, pos, mn, mx);
    }
    if constexpr (!REC) {
        st3(b, F_AMIN, i, mn);
        st3(b, F_AMAX, i, mx);
    }
    write_records(b, i, type, kin, sp, R, pos, aabb_scale(mn, mx, pos), mn, mx, Iw, im, v, w, tv, tw, mu, e);
    return Box3{mn, mx, kin ? 1.0f : 0.0f, sp[3]};
}

// AABB, scale, world inverse inertia and records of the current pose (RigidBody::setTransform).
__device__ __forceinline__ void refresh_body(const BodyView& b, int i) {
    const int type = b.iat(I_TYPE, i);
    const bool kin = b.iat(I_KIN, i) != 0;
    const f3 pos = ld3(b, F_POS, i);
    const m33 R = ld9(b, F_ROT, i);
    const float sp[4] = {b.at(F_SP, i), b.at(F_SP + 1, i), b.at(F_SP + 2, i), b.at(F_SP + 3, i)};
    f3 mn, mx;
    if (type <= SHAPE_CYLINDER) shape_aabb(type, sp, R, pos, mn, mx);
    else if (type == SHAPE_COMPOUND) compound_aabb(b.ms, __float_as_int(sp[0]), R, pos, mn, mx);
    else mesh_aabb(b.ms, __float_as_int(sp[0]), R, pos, mn, mx);
    st3(b, F_AMIN, i, mn);
    st3(b, F_AMAX, i, mx);
    m33 Iw;
    if (kin) { for (int k = 0; k < 9; k++) Iw.m[k] = 0.0f; }
    else Iw = mmt(mm(R, ld9(b, F_LINV, i)), R);
    const f3 v = ld3(b, F_V, i), w = ld3(b, F_W, i);
    write_records(b, i, type, kin, sp, R, pos, aabb_scale(mn, mx, pos), mn, mx, Iw, kin ? 0.0f : b.at(F_IM, i), v, w,
                  kin ? f3{0, 0, 0} : v, kin ? f3{0, 0, 0} : w, b.at(F_MU, i), b.at(F_E, i));
}

// ------------------------------------------------------------------ stage 1
// BroadPhaseBase::validateCollisionPair (broad_phase_base.cpp:5-19)
// The same test on AABB records already loaded (min xyz, kinematic flag, max xyz, collision flag);
// the same test as sap_pair_ok (pe_sap.hpp; batched worlds have no ignore-pair lists).
__device__ __forceinline__ bool bp_valid_rec(const float (&A)[8], const float (&B)[8]) {
    if (A[7] == COLL_IGNORE || B[7] == COLL_IGNORE) return false;
    if (A[3] != 0.0f && B[3] != 0.0f) return false;
    if (A[0] > B[4] || B[0] > A[4]) return false;
    if (A[1] > B[5] || B[1] > A[5]) return false;
    if (A[2] > B[6] || B[2] > A[6]) return false;
    return true;
}
__device__ __forceinline__ bool bp_valid(const BodyView& b, int i, int j) {
    float A[8], B[8];
    ld_rec<2>(b.abr(i), A);
    ld_rec<2>(b.abr(j), B);
    return bp_valid_rec(A, B);
}

// ------------------------------------------------------------------ stage 2
struct CollBody {
    Pose P;
    int type;
    bool kin;
    float sp[3];
    float mu, e;
};
__device__ __forceinline__ CollBody load_coll(const BodyView& b, int i) {
    float c[CBW];
    ld_rec<CBW / 4>(b.cbr(i), c);
    CollBody r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.P.R.m[k] = c[CB_R + k];
    r.P.p = f3{c[CB_P], c[CB_P + 1], c[CB_P + 2]};
    r.P.scale = c[CB_SCALE];
    r.sp[0] = c[CB_SP]; r.sp[1] = c[CB_SP + 1]; r.sp[2] = c[CB_SP + 2];
    r.type = __float_as_int(c[CB_TYPE]);
    r.kin = __float_as_int(c[CB_KIN]) != 0;
    r.mu = c[CB_MU];
    r.e = c[CB_E];
    return r;
}

// Wave-aggregated append: the active lanes with pred take consecutive slots of *counter (one
// atomic per wave); returns the lane's slot or -1.
__device__ __forceinline__ int wave_append(bool pred, int* counter) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if ((int)(threadIdx.x & (WAVE - 1)) == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, WAVE);
    return pred ? base + (int)__popcll(m & lanemask_lt()) : -1;
}

// True when both bodies of pair p are boxes or spheres.
__device__ __forceinline__ bool pair_is_simple(const BodyView& b, const PairStore& ps, int p) {
    return b.iat(I_TYPE, ps.pa[p]) <= SHAPE_SPHERE && b.iat(I_TYPE, ps.pb[p]) <= SHAPE_SPHERE;
}
// True when pair p is solved by a whole wave (narrow_pair's G > 1 branch): the pairs of the
// convex separating-axis routine (and its clipper) -- cylinder and cylinder, box and convex mesh,
// convex mesh and convex mesh (edge-pair axes in the thousands), box or convex mesh and concave
// mesh (one test per overlapped triangle) -- and a sphere or cylinder with a convex or concave
// mesh (the faces its BVH walk finds are tested on the wave's lanes).
__device__ __forceinline__ bool pair_on_wave(const BodyView& b, const PairStore& ps, int p) {
    const int ta = b.iat(I_TYPE, ps.pa[p]), tb = b.iat(I_TYPE, ps.pb[p]);
    const int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
    return (lo == SHAPE_CYLINDER && hi == SHAPE_CYLINDER) || (lo == SHAPE_BOX && hi == SHAPE_CONVEX) ||
           (lo == SHAPE_CONVEX && hi == SHAPE_CONVEX) ||
           ((lo == SHAPE_BOX || lo == SHAPE_CONVEX) && hi == SHAPE_CONCAVE) ||
           ((lo == SHAPE_SPHERE || lo == SHAPE_CYLINDER) && (hi == SHAPE_CONVEX || hi == SHAPE_CONCAVE));
}

// The polyhedral view of a box, cylinder or convex-mesh body.
__device__ __forceinline__ Poly body_poly(const MeshSet& ms, const CollBody& B) {
    if (B.type == SHAPE_BOX) return poly_box(B.sp);
    if (B.type == SHAPE_CYLINDER) return poly_cyl(&ms, B.sp[0], B.sp[1]);
    return poly_mesh(&ms, __float_as_int(B.sp[0]));
}
// NarrowPhaseBase::getAlgorithm (narrow_phase_base.cpp:20-30) for the pairs that involve a
// cylinder, a convex mesh or a concave mesh and are not separating-axis pairs, with the role and
// swap flag each algorithm's processCollision gives the two bodies; the swap flag is cleared again
// afterwards, as processCollision does.  G = 1: one lane (the thread-per-pair kernels, where only
// box or sphere with cylinder arrive, and compound children); G = WAVE: the sphere and cylinder
// pairs with a mesh that pair_on_wave sends to a whole wave.
template <int G = 1>
static __device__ __noinline__ void narrow_shapes(const MeshSet& ms, Cache& c, PairCtx& x, const CollBody& A,
                                                  const CollBody& B) {
    const int ta = A.type, tb = B.type;
    auto is = [&](int p, int q) { return (ta == p && tb == q) || (ta == q && tb == p); };
    if (is(SHAPE_BOX, SHAPE_CYLINDER)) {                     // box_cylinder_collision_algorithm.cpp:13-47
        const CollBody& bx = ta == SHAPE_BOX ? A : B;
        const CollBody& cy = ta == SHAPE_BOX ? B : A;
        x.swap = ta == SHAPE_BOX;
        box_cylinder(c, x, bx.P, bx.sp, cy.P, cy.sp[0], cy.sp[1]);
    } else if (is(SHAPE_SPHERE, SHAPE_CYLINDER)) {           // sphere_cylinder_collision_algorithm.cpp:8-30
        const CollBody& sp = ta == SHAPE_SPHERE ? A : B;
        const CollBody& cy = ta == SHAPE_SPHERE ? B : A;
        x.swap = ta == SHAPE_SPHERE;
        sphere_cylinder(c, x, sp.P, sp.sp[0], cy.P, cy.sp[0], cy.sp[1]);
    } else if (is(SHAPE_SPHERE, SHAPE_CONVEX)) {             // sphere_convex_collision_algorithm.cpp:11-34
        const CollBody& sp = ta == SHAPE_SPHERE ? A : B;
        const CollBody& mb = ta == SHAPE_SPHERE ? B : A;
        x.swap = ta == SHAPE_CONVEX;
        sphere_mesh<G>(c, x, ms, __float_as_int(mb.sp[0]), mb.P, sp.P, sp.sp[0]);
    } else if (is(SHAPE_CYLINDER, SHAPE_CONVEX)) {           // cylinder_convex_collision_algorithm.cpp:11-51
        const CollBody& cy = ta == SHAPE_CYLINDER ? A : B;
        const CollBody& mb = ta == SHAPE_CYLINDER ? B : A;
        x.swap = ta == SHAPE_CONVEX;
        cylinder_mesh<G>(c, x, ms, __float_as_int(mb.sp[0]), mb.P, cy.P, cy.sp[0], cy.sp[1]);
    } else if ((ta == SHAPE_CONCAVE) != (tb == SHAPE_CONCAVE)) {
        // concave_{box,sphere,cylinder,convex}_collision_algorithm.cpp: the other body is A
        const CollBody& mb = ta == SHAPE_CONCAVE ? A : B;
        const CollBody& ob = ta == SHAPE_CONCAVE ? B : A;
        const int mesh = __float_as_int(mb.sp[0]);
        x.swap = ta == SHAPE_CONCAVE;
        if (ob.type == SHAPE_SPHERE) sphere_mesh<G>(c, x, ms, mesh, mb.P, ob.P, ob.sp[0]);
        else if (ob.type == SHAPE_CYLINDER) cylinder_mesh<G>(c, x, ms, mesh, mb.P, ob.P, ob.sp[0], ob.sp[1]);
    }
    x.swap = false;
}

// A compound's child k as a narrowphase participant: the child's shape at its world pose; the
// body's material, kinematic flag and AABB scale.
__device__ __forceinline__ CollBody child_coll(const MeshSet& ms, const CollBody& body, int k) {
    CollBody c = body;
    float sp[3];
    m33 Rc;
    f3 pc;
    c.type = child_pose(ms, k, body.P.R, body.P.p, sp, Rc, pc);
    c.P.R = Rc;
    c.P.p = pc;
    c.sp[0] = sp[0]; c.sp[1] = sp[1]; c.sp[2] = sp[2];
    return c;
}
// NarrowPhaseBase::getAlgorithm (narrow_phase_base.cpp:20-30) and processCollision for two shapes
// of which neither is a compound, every algorithm on one thread (compound children: rare, so the
// separating-axis routine runs here on a lane of its own rather than a wave).
static __device__ __noinline__ void child_dispatch(const MeshSet& ms, Cache& c, PairCtx& x, const CollBody& A,
                                                   const CollBody& B, float ref_scale) {
    const int ta = A.type, tb = B.type;
    const int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
    if (ta == SHAPE_BOX && tb == SHAPE_BOX) box_box(c, x, A.P, B.P, A.sp, B.sp);
    else if (ta == SHAPE_BOX && tb == SHAPE_SPHERE) box_sphere(c, x, true, A.P, B.P, A.sp, B.sp[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_BOX) box_sphere(c, x, false, A.P, B.P, B.sp, A.sp[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) sphere_sphere(c, x, A.P, B.P, A.sp[0], B.sp[0]);
    else if ((lo == SHAPE_CYLINDER && hi == SHAPE_CYLINDER) || (lo == SHAPE_BOX && hi == SHAPE_CONVEX) ||
             (lo == SHAPE_CONVEX && hi == SHAPE_CONVEX)) {
        convex_convex<1>(c, x, body_poly(ms, A), body_poly(ms, B), A.P, B.P, ref_scale);
    } else if ((lo == SHAPE_BOX || lo == SHAPE_CONVEX) && hi == SHAPE_CONCAVE) {
        // concave_convex_collision_algorithm.cpp: the convex shape is A
        const CollBody& mb = ta == SHAPE_CONCAVE ? A : B;
        const CollBody& ob = ta == SHAPE_CONCAVE ? B : A;
        x.swap = ta == SHAPE_CONCAVE;
        concave_poly<1>(c, x, ms, __float_as_int(mb.sp[0]), mb.P, body_poly(ms, ob), ob.P, ref_scale);
        x.swap = false;
    } else {
        narrow_shapes(ms, c, x, A, B);
    }
}
// CompoundCompoundCollisionAlgorithm::processCollision (compound_compound_collision_algorithm.cpp:
// 8-48): a compound B trades places with A, the result's objects too (returns true then); each child
// of A, posed by A's transform, meets B or each child of B.
static __device__ __noinline__ bool narrow_compound(const MeshSet& ms, Cache& c, PairCtx& x, const CollBody& A0,
                                                    const CollBody& B0, float ref_scale) {
    const bool sw = B0.type == SHAPE_COMPOUND;
    const CollBody& A = sw ? B0 : A0;
    const CollBody& B = sw ? A0 : B0;
    if (sw) {
        const Pose t = x.A;
        x.A = x.B;
        x.B = t;
    }
    const int2 ka = ms.krange[__float_as_int(A.sp[0])];
    for (int k = 0; k < ka.y; k++) {
        const CollBody ca = child_coll(ms, A, ka.x + k);
        if (B.type == SHAPE_COMPOUND) {
            const int2 kb = ms.krange[__float_as_int(B.sp[0])];
            for (int q = 0; q < kb.y; q++) child_dispatch(ms, c, x, ca, child_coll(ms, B, kb.x + q), ref_scale);
        } else {
            child_dispatch(ms, c, x, ca, B, ref_scale);
        }
    }
    return sw;
}

// SimpleNarrowPhase::calcContactResults (simple_narrow_phase.cpp:7-41) for pair slot p.
// Slots that existed in the previous step (p < prev_np) keep their distances, because the
// reference reuses its ContactResult objects and clearContactPoints only invalidates
// positions (contact_result.cpp:158-164); newer slots start from a fresh ContactResult.
// SHAPES = false handles box and sphere pairs only (the kernel then carries no polyhedral
// code and keeps its register budget); SHAPES = true handles every pair type.
// G = 64 (a whole wave on one pair, every lane holding the same values) is for the pairs
// pair_on_wave selects: the separating-axis loops, or the mesh faces a BVH walk finds, are split
// over the lanes, and lane 0 writes.
// COMPOUND = true is for the pairs with a compound body (one thread each).
template <bool SHAPES, int G = 1, bool COMPOUND = false>
static __device__ __forceinline__ void narrow_pair(const BodyView& b, const PairStore& ps, int p, int prev_np) {
    const int ia = ps.pa[p], ib = ps.pb[p];
    // The shape kernels keep the contact cache in LDS: their algorithms take it by reference across
    // calls that are not inlined, which would otherwise put it in scratch memory.  One row per lane
    // in the thread-per-pair kernels (at most 64 threads per workgroup), one per wave when a wave
    // works on one pair (at most 4 waves per workgroup; every lane writes the same values).
    constexpr bool LDS_CACHE = SHAPES;
    __shared__ Cache s_cache[LDS_CACHE ? (G == 1 ? 64 : 4) : 1];   // rows of odd length: distinct banks
    Cache c_reg;
    Cache& c = LDS_CACHE ? s_cache[G == 1 ? threadIdx.x % 64 : (threadIdx.x / WAVE) % 4] : c_reg;
    const size_t slot0 = (size_t)p * 8;
    float dprev[8];
    if (p < prev_np) ld_rec<2>(ps.cd + slot0, dprev);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        c.wx[k] = REAL_MAX; c.wy[k] = REAL_MAX; c.wz[k] = REAL_MAX;
        c.nx[k] = 0; c.ny[k] = 0; c.nz[k] = 0;
        c.ax[k] = 0; c.ay[k] = 0; c.az[k] = 0;
        c.bx[k] = 0; c.by[k] = 0; c.bz[k] = 0;
        c.d[k] = p < prev_np ? dprev[k] : REAL_MAX;
    }
    c.trunc = 0;
    const CollBody A = load_coll(b, ia), B = load_coll(b, ib);
    PairCtx x;
    x.A = A.P;
    x.B = B.P;
    x.swap = false;
    {
        const float sa = x.A.scale, sb = x.B.scale;
        const float th = (sa < sb ? sa : sb) * 0.02f;
        x.same_th2 = th * th;
    }
    const int ta = A.type, tb = B.type;
    bool swapped = false;
    if (COMPOUND) {
        swapped = narrow_compound(b.ms, c, x, A, B, (x.A.scale + x.B.scale) * 0.05f);
    } else if (G > 1) {
        const float ref_scale = (x.A.scale + x.B.scale) * 0.05f;
        if (ta == SHAPE_SPHERE || tb == SHAPE_SPHERE || (ta == SHAPE_CYLINDER) != (tb == SHAPE_CYLINDER)) {
            narrow_shapes<G>(b.ms, c, x, A, B);   // a sphere or a cylinder with a mesh
        } else if (ta == SHAPE_CONCAVE || tb == SHAPE_CONCAVE) {
            // concave_convex_collision_algorithm.cpp: the convex body is A (narrow_shapes)
            const CollBody& mb = ta == SHAPE_CONCAVE ? A : B;
            const CollBody& ob = ta == SHAPE_CONCAVE ? B : A;
            x.swap = ta == SHAPE_CONCAVE;
            concave_poly<G>(c, x, b.ms, __float_as_int(mb.sp[0]), mb.P, body_poly(b.ms, ob), ob.P, ref_scale);
            x.swap = false;
        } else {
            convex_convex<G>(c, x, body_poly(b.ms, A), body_poly(b.ms, B), A.P, B.P, ref_scale);
        }
    } else if (ta == SHAPE_BOX && tb == SHAPE_BOX) box_box(c, x, A.P, B.P, A.sp, B.sp);
    else if (ta == SHAPE_BOX && tb == SHAPE_SPHERE) box_sphere(c, x, true, A.P, B.P, A.sp, B.sp[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_BOX) box_sphere(c, x, false, A.P, B.P, B.sp, A.sp[0]);
    else if (ta == SHAPE_SPHERE && tb == SHAPE_SPHERE) sphere_sphere(c, x, A.P, B.P, A.sp[0], B.sp[0]);
    else if (SHAPES) narrow_shapes(b.ms, c, x, A, B);
    int rank[8];
    const int np = sort_rank(c, rank);
    if (G > 1 && Grp<G>::rank() != 0) return;
    ps.npts[p] = np;
    if (COMPOUND && swapped) {   // the result's objects: the compound first
        ps.pa[p] = ib;
        ps.pb[p] = ia;
    }
    // bit 0: the swap flag; bit 1: contact points were lost to a full clip buffer (a capacity error)
    ps.swap[p] = (x.swap ? 1 : 0) | (c.trunc ? 2 : 0);
    float dsorted[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
#pragma unroll
        for (int q = 0; q < 8; q++) if (rank[q] == k) dsorted[k] = c.d[q];
    }
    st_rec<2>(ps.cd + slot0, dsorted);
    // only the valid points are read later; write them at their sorted positions
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (rank[k] >= np) continue;
        const float rec[CW] = {c.wx[k], c.wy[k], c.wz[k], c.nx[k], c.ny[k], c.nz[k],
                               c.ax[k], c.ay[k], c.az[k], c.bx[k], c.by[k], c.bz[k]};
        st_rec<CW / 4>(ps.cp + (slot0 + rank[k]) * CW, rec);
    }
}

// ------------------------------------------------------------------ stage 3
struct SolvBody {
    m33 Iw;
    float im;
    bool kin;
    f3 tv, tw, v, w;
};
__device__ __forceinline__ SolvBody load_solv(const BodyView& b, int i) {
    float s[SBW];
    ld_rec<SBW / 4>(b.sbr(i), s);
    SolvBody r;
    r.tv = f3{s[SB_TV], s[SB_TV + 1], s[SB_TV + 2]};
    r.tw = f3{s[SB_TW], s[SB_TW + 1], s[SB_TW + 2]};
    r.im = s[SB_IM];
    r.kin = s[SB_KIN] != 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) r.Iw.m[k] = s[SB_IW + k];
    r.v = f3{s[SB_V], s[SB_V + 1], s[SB_V + 2]};
    r.w = f3{s[SB_W], s[SB_W + 1], s[SB_W + 2]};
    return r;
}

// FrictionContactConstraint::initSequentialImpulse (friction_contact_constraint.cpp:9-73) for
// point k of pair p, given the pair's bodies and pair coefficients; the row goes to out.
// The pair-level quantities: getKinematicInvMass (rigidbody.cpp:48-62, zero when kinematic)
// summed, and ContactResult::setObjects's combined coefficients (contact_result.cpp:80-86).
struct PairCoef {
    float ims, mu, e;
};
__device__ __forceinline__ PairCoef pair_coef(const CollBody& CA, const CollBody& CB, const SolvBody& A, const SolvBody& B) {
    return PairCoef{A.im + B.im, sqrtf(CA.mu * CB.mu), sqrtf(CA.e * CB.e)};
}
__device__ __forceinline__ void setup_point(const PairStore& ps, int p, int k, const CollBody& CA, const CollBody& CB,
                                            const SolvBody& A, const SolvBody& B, const PairCoef& q, float dt, float* out) {
    const size_t s = (size_t)p * 8 + k;
    float c[CW];
    ld_rec<CW / 4>(ps.cp + s * CW, c);
    const f3 nn{c[C_N], c[C_N + 1], c[C_N + 2]};
    const f3 la{c[C_LA], c[C_LA + 1], c[C_LA + 2]};
    const f3 lb{c[C_LB], c[C_LB + 1], c[C_LB + 2]};
    const Row r = make_row(CA.P.R, CB.P.R, la, lb, nn, ps.cd[s], q.ims, A.Iw, B.Iw, A.v, A.w, B.v, B.w, q.mu, q.e, dt);
    float o[RW];
    o[R_RA] = r.ra.x; o[R_RA + 1] = r.ra.y; o[R_RA + 2] = r.ra.z;
    o[R_RB] = r.rb.x; o[R_RB + 1] = r.rb.y; o[R_RB + 2] = r.rb.z;
    o[R_N] = r.n.x; o[R_N + 1] = r.n.y; o[R_N + 2] = r.n.z;
    o[R_T0] = r.t0.x; o[R_T0 + 1] = r.t0.y; o[R_T0 + 2] = r.t0.z;
    o[R_T1] = r.t1.x; o[R_T1 + 1] = r.t1.y; o[R_T1 + 2] = r.t1.z;
    o[R_RHS] = r.rhs; o[R_MN] = r.mn; o[R_MT0] = r.mt0; o[R_MT1] = r.mt1; o[R_MU] = r.mu;
    o[R_LN] = 0.0f; o[R_LT0] = 0.0f; o[R_LT1] = 0.0f; o[23] = 0.0f;
    st_rec<RW / 4>(out, o);
}
// Every point of pair p, to the row store (row k of the pair at slot p * 8 + k).
__device__ __forceinline__ void setup_pair(const BodyView& b, const PairStore& ps, int p, float dt) {
    const int n = ps.npts[p];
    if (n == 0) return;
    const int ia = ps.pa[p], ib = ps.pb[p];
    const CollBody CA = load_coll(b, ia), CB = load_coll(b, ib);
    const SolvBody A = load_solv(b, ia), B = load_solv(b, ib);
    const PairCoef q = pair_coef(CA, CB, A, B);
    for (int k = 0; k < n; k++) setup_point(ps, p, k, CA, CB, A, B, q, dt, ps.rows + ((size_t)p * 8 + k) * RW);
}
// Point k of pair p alone, with the same arithmetic as setup_pair; the rows of a pair can so be
// set up by different threads.
// ia, ib: the pair's bodies, when the caller holds them (one dependent load fewer).
__device__ __forceinline__ void setup_row(const BodyView& b, const PairStore& ps, int p, int ia, int ib, int k, float dt) {
    const CollBody CA = load_coll(b, ia), CB = load_coll(b, ib);
    const SolvBody A = load_solv(b, ia), B = load_solv(b, ib);
    setup_point(ps, p, k, CA, CB, A, B, pair_coef(CA, CB, A, B), dt, ps.rows + ((size_t)p * 8 + k) * RW);
}
__device__ __forceinline__ void setup_row(const BodyView& b, const PairStore& ps, int p, int k, float dt) {
    setup_row(b, ps, p, ps.pa[p], ps.pb[p], k, dt);
}

// RigidBody::applyTempImpulse (rigidbody.cpp:166-170), with the body's inverse mass and world
// inverse inertia already in registers.
__device__ __forceinline__ void apply_temp(f3& tv, f3& tw, float im, const m33& Iw, f3 r, f3 J) {
    tv = add(tv, mul(J, im));
    tw = add(tw, mv(Iw, cross(r, J)));
}

// FrictionContactConstraint::iterateSequentialImpulse (friction_contact_constraint.cpp:75-108)
// over all rows of pair p: both bodies and all rows are loaded before the dependent chain.
__device__ __forceinline__ void iterate_pair(const BodyView& b, const PairStore& ps, int p) {
    const int ia = ps.pa[p], ib = ps.pb[p];
    const int n = ps.npts[p];
    SolvBody A = load_solv(b, ia), B = load_solv(b, ib);
    float* R = ps.rows;
    for (int k0 = 0; k0 < n; k0 += 4) {
        float rr[4][RW];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k0 + u < n) ld_rec<RW / 4>(R + ((size_t)p * 8 + k0 + u) * RW, rr[u]);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                const float* o = rr[u];
                Row r;
                r.ra = f3{o[R_RA], o[R_RA + 1], o[R_RA + 2]};
                r.rb = f3{o[R_RB], o[R_RB + 1], o[R_RB + 2]};
                r.n = f3{o[R_N], o[R_N + 1], o[R_N + 2]};
                r.t0 = f3{o[R_T0], o[R_T0 + 1], o[R_T0 + 2]};
                r.t1 = f3{o[R_T1], o[R_T1 + 1], o[R_T1 + 2]};
                r.rhs = o[R_RHS]; r.mn = o[R_MN]; r.mt0 = o[R_MT0]; r.mt1 = o[R_MT1]; r.mu = o[R_MU];
                float ln = o[R_LN], lt0 = o[R_LT0], lt1 = o[R_LT1];
                const f3 J = iterate_row(r, ln, lt0, lt1, A.tv, A.tw, B.tv, B.tw);
                if (!A.kin) apply_temp(A.tv, A.tw, A.im, A.Iw, r.ra, J);
                if (!B.kin) apply_temp(B.tv, B.tw, B.im, B.Iw, r.rb, neg(J));
                rr[u][R_LN] = ln; rr[u][R_LT0] = lt0; rr[u][R_LT1] = lt1;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                float4* dst = reinterpret_cast<float4*>(R + ((size_t)p * 8 + k0 + u) * RW) + 5;   // floats 20..23
                *dst = make_float4(rr[u][R_LN], rr[u][R_LT0], rr[u][R_LT1], 0.0f);
            }
        }
    }
    if (!A.kin) {
        float4* d = reinterpret_cast<float4*>(b.sbr(ia));
        d[0] = make_float4(A.tv.x, A.tv.y, A.tv.z, A.tw.x);
        d[1].x = A.tw.y; d[1].y = A.tw.z;
    }
    if (!B.kin) {
        float4* d = reinterpret_cast<float4*>(b.sbr(ib));
        d[0] = make_float4(B.tv.x, B.tv.y, B.tv.z, B.tw.x);
        d[1].x = B.tw.y; d[1].y = B.tw.z;
    }
}

// iterate_pair split around the barrier before the pair's colour: gs_prefetch loads what no other
// thread changes during the sweeps -- the pair's bodies, their world inverse inertias and its
// first four rows (a pair's rows, accumulated impulses included, are only written by the thread
// that solves it, every iteration the same thread) -- and gs_solve, after the barrier, loads the
// bodies' temp velocities, which earlier colours changed, and solves the pair with the same
// arithmetic in the same order.
struct GsPair {
    int p, ia, ib, n;
    m33 Ia, Ib;
    float rr[4][RW];
};
__device__ __forceinline__ void gs_prefetch(const BodyView& b, const PairStore& ps, int p, GsPair& G) {
    G.p = p;
    G.ia = ps.pa[p];
    G.ib = ps.pb[p];
    G.n = ps.npts[p];
    const float* sa = b.sbr(G.ia);
    const float* sb = b.sbr(G.ib);
#pragma unroll
    for (int k = 0; k < 9; k++) { G.Ia.m[k] = sa[SB_IW + k]; G.Ib.m[k] = sb[SB_IW + k]; }
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (u < G.n) ld_rec<RW / 4>(ps.rows + ((size_t)p * 8 + u) * RW, G.rr[u]);
}
__device__ __forceinline__ void gs_solve(const BodyView& b, const PairStore& ps, GsPair& G) {
    const int p = G.p, n = G.n;
    float* R = ps.rows;
    const float4* qa = reinterpret_cast<const float4*>(b.sbr(G.ia));
    const float4* qb = reinterpret_cast<const float4*>(b.sbr(G.ib));
    const float4 a0 = qa[0], a1 = qa[1], b0 = qb[0], b1 = qb[1];
    f3 tva{a0.x, a0.y, a0.z}, twa{a0.w, a1.x, a1.y}, tvb{b0.x, b0.y, b0.z}, twb{b0.w, b1.x, b1.y};
    const float ima = a1.z, imb = b1.z;
    const bool ka = a1.w != 0.0f, kb = b1.w != 0.0f;
    for (int k0 = 0; k0 < n; k0 += 4) {
        if (k0 > 0) {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (k0 + u < n) ld_rec<RW / 4>(R + ((size_t)p * 8 + k0 + u) * RW, G.rr[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k0 + u < n) {
                const float* o = G.rr[u];
                Row r;
                r.ra = f3{o[R_RA], o[R_RA + 1], o[R_RA + 2]};
                r.rb = f3{o[R_RB], o[R_RB + 1], o[R_RB + 2]};
                r.n = f3{o[R_N], o[R_N + 1], o[R_N + 2]};
                r.t0 = f3{o[R_T0], o[R_T0 + 1], o[R_T0 + 2]};
                r.t1 = f3{o[R_T1], o[R_T1 + 1], o[R_T1 + 2]};
                r.rhs = o[R_RHS]; r.mn = o[R_MN]; r.mt0 = o[R_MT0]; r.mt1 = o[R_MT1]; r.mu = o[R_MU];
                float ln = o[R_LN], lt0 = o[R_LT0], lt1 = o[R_LT1];
                const f3 J = iterate_row(r, ln, lt0, lt1, tva, twa, tvb, twb);
                if (!ka) apply_temp(tva, twa, ima, G.Ia, r.ra, J);
                if (!kb) apply_temp(tvb, twb, imb, G.Ib, r.rb, neg(J));
                G.rr[u][R_LN] = ln; G.rr[u][R_LT0] = lt0; G.rr[u][R_LT1] = lt1;
                float4* dst = reinterpret_cast<float4*>(R + ((size_t)p * 8 + k0 + u) * RW) + 5;   // floats 20..23
                *dst = make_float4(ln, lt0, lt1, 0.0f);
            }
        }
    }
    if (!ka) {
        float4* d = reinterpret_cast<float4*>(b.sbr(G.ia));
        d[0] = make_float4(tva.x, tva.y, tva.z, twa.x);
        d[1].x = twa.y; d[1].y = twa.z;
    }
    if (!kb) {
        float4* d = reinterpret_cast<float4*>(b.sbr(G.ib));
        d[0] = make_float4(tvb.x, tvb.y, tvb.z, twb.x);
        d[1].x = twb.y; d[1].y = twb.z;
    }
}

// Solver state held in LDS for the Gauss-Seidel sweeps: per body (local index) one 64-byte record
// of the temp velocities, the inverse mass (negative for a kinematic body) and the world inverse
// inertia, read with four 128-bit loads; per pair its local bodies and first row (row counts are
// differences).
struct LdsSolve {
    float4* bs;         // [n][4]: (tv xyz, tw x), (tw yz, im or -1: kinematic, Iw0), (Iw1..4), (Iw5..8)
    int* pab;           // [npairs] local A | local B << 16
    int* rs;            // [npairs + 1] first row of each pair
    int n;
};
__device__ __forceinline__ void lds_load_body(const LdsSolve& L, const BodyView& b, int gi, int li) {
    const float* s = b.sbr(gi);
    float r[SBW];
    ld_rec<SBW / 4>(s, r);
    float4* q = L.bs + 4 * li;
    q[0] = make_float4(r[SB_TV], r[SB_TV + 1], r[SB_TV + 2], r[SB_TW]);
    q[1] = make_float4(r[SB_TW + 1], r[SB_TW + 2], r[SB_KIN] != 0.0f ? -1.0f : r[SB_IM], r[SB_IW]);
    q[2] = make_float4(r[SB_IW + 1], r[SB_IW + 2], r[SB_IW + 3], r[SB_IW + 4]);
    q[3] = make_float4(r[SB_IW + 5], r[SB_IW + 6], r[SB_IW + 7], r[SB_IW + 8]);
}
// iterate_pair on the LDS copy of the bodies (same arithmetic, same order); the pair's rows are
// at rows (the row store), RW floats apart.
// Loads the first two solver rows of local pair pl (those it has) into cur and nxt, ahead of
// iterate_pair_lds: the sweep issues them one colour early so their latency overlaps the barrier,
// and a pair of one or two rows then solves without waiting on memory.
__device__ __forceinline__ void prefetch_pair_rows(const LdsSolve& L, int pl, const float* rows, float (&cur)[RW],
                                                   float (&nxt)[RW]) {
    const int nr = L.rs[pl + 1] - L.rs[pl];
    if (nr > 0) ld_rec<RW / 4>(rows, cur);
    if (nr > 1) ld_rec<RW / 4>(rows + RW, nxt);
}

// One Gauss-Seidel pass over the rows of local pair pl on the LDS body state; cur and nxt hold
// its first two rows (prefetch_pair_rows); row k + 2 is loaded into the buffer row k leaves, while
// row k + 1 is solved.
__device__ __forceinline__ void iterate_pair_lds(const LdsSolve& L, int pl, float* rows, float (&cur)[RW],
                                                 float (&nxt)[RW]) {
    const int ab = L.pab[pl];
    const int a = ab & 0xffff, c = ab >> 16;
    const int nr = L.rs[pl + 1] - L.rs[pl];
    const float4 a0 = L.bs[4 * a], a1 = L.bs[4 * a + 1], a2 = L.bs[4 * a + 2], a3 = L.bs[4 * a + 3];
    const float4 c0 = L.bs[4 * c], c1 = L.bs[4 * c + 1], c2 = L.bs[4 * c + 2], c3 = L.bs[4 * c + 3];
    f3 tva{a0.x, a0.y, a0.z}, twa{a0.w, a1.x, a1.y};
    f3 tvb{c0.x, c0.y, c0.z}, twb{c0.w, c1.x, c1.y};
    const float ima = a1.z, imb = c1.z;
    const m33 Ia{{a1.w, a2.x, a2.y, a2.z, a2.w, a3.x, a3.y, a3.z, a3.w}};
    const m33 Ib{{c1.w, c2.x, c2.y, c2.z, c2.w, c3.x, c3.y, c3.z, c3.w}};
    // One row of the chain from buffer o; row k's accumulated impulses go back to the row store.
    auto row = [&](const float (&o)[RW], int k) {
        Row r;
        r.ra = f3{o[R_RA], o[R_RA + 1], o[R_RA + 2]};
        r.rb = f3{o[R_RB], o[R_RB + 1], o[R_RB + 2]};
        r.n = f3{o[R_N], o[R_N + 1], o[R_N + 2]};
        r.t0 = f3{o[R_T0], o[R_T0 + 1], o[R_T0 + 2]};
        r.t1 = f3{o[R_T1], o[R_T1 + 1], o[R_T1 + 2]};
        r.rhs = o[R_RHS]; r.mn = o[R_MN]; r.mt0 = o[R_MT0]; r.mt1 = o[R_MT1]; r.mu = o[R_MU];
        float ln = o[R_LN], lt0 = o[R_LT0], lt1 = o[R_LT1];
        const f3 J = iterate_row(r, ln, lt0, lt1, tva, twa, tvb, twb);
        if (ima >= 0.0f) apply_temp(tva, twa, ima, Ia, r.ra, J);
        if (imb >= 0.0f) apply_temp(tvb, twb, imb, Ib, r.rb, neg(J));
        reinterpret_cast<float4*>(rows + (size_t)k * RW)[5] = make_float4(ln, lt0, lt1, 0.0f);
    };
    // Rows alternate between cur and nxt (no copies).
    for (int k = 0; k < nr; k += 2) {
        row(cur, k);
        if (k + 2 < nr) ld_rec<RW / 4>(rows + (size_t)(k + 2) * RW, cur);
        if (k + 1 >= nr) break;
        row(nxt, k + 1);
        if (k + 3 < nr) ld_rec<RW / 4>(rows + (size_t)(k + 3) * RW, nxt);
    }
    if (ima >= 0.0f) {
        L.bs[4 * a] = make_float4(tva.x, tva.y, tva.z, twa.x);
        L.bs[4 * a + 1] = make_float4(twa.y, twa.z, ima, a1.w);
    }
    if (imb >= 0.0f) {
        L.bs[4 * c] = make_float4(tvb.x, tvb.y, tvb.z, twb.x);
        L.bs[4 * c + 1] = make_float4(twb.y, twb.z, imb, c1.w);
    }
}

// SequentialImpulseSolver::solve's syncTempVelocity (sequential_impulse_solver.cpp:71-80)
__device__ __forceinline__ void sync_body(const BodyView& b, int i) {
    if (b.iat(I_KIN, i)) return;
    const float* s = b.sbr(i);
    st3(b, F_V, i, f3{s[SB_TV], s[SB_TV + 1], s[SB_TV + 2]});
    st3(b, F_W, i, f3{s[SB_TW], s[SB_TW + 1], s[SB_TW + 2]});
}
// The same for a world whose records hold its velocities (integrate_body<.., true>): into the
// solver record's v, w.
__device__ __forceinline__ void sync_body_rec(const BodyView& b, int i) {
    if (b.iat(I_KIN, i)) return;
    float* s = b.sbr(i);
    const f3 tv{s[SB_TV], s[SB_TV + 1], s[SB_TV + 2]}, tw{s[SB_TW], s[SB_TW + 1], s[SB_TW + 2]};
    s[SB_V] = tv.x; s[SB_V + 1] = tv.y; s[SB_V + 2] = tv.z;
    s[SB_W] = tw.x; s[SB_W + 1] = tw.y; s[SB_W + 2] = tw.z;
}

// ------------------------------------------------------------------ constraint-graph colouring
// Colours pairs so that no two pairs of one colour share a dynamic body (kinematic bodies never
// change, rigidbody.cpp:167, so they impose no order).  Colour c is a maximal independent set
// built in sub-rounds: every candidate claims its dynamic bodies with atomicMin of its priority
// (colour::prio of its index in the world); a candidate that holds all its claims joins the
// colour and takes its bodies; a candidate touching a taken body waits for a later colour.  The colour classes depend only on
// the pair set, not on list order, so sweeping a colour in parallel equals the sequential
// Gauss-Seidel sweep over (colour, any order within the colour).
// claim[] holds, per body, INT_MAX (free), -1 (taken by this colour) or the lowest claim.
namespace colour {
constexpr int FREE = 0x7fffffff;
constexpr int TAKEN = -1;
// A pair's claim priority (the lowest wins), unique within its world: its contact-point count
// first, more points colouring earlier, then a fixed bijection of its index in the world onto
// [0, 2^28) (odd multiplications modulo 2^28 and right xor-shifts, each invertible there).  A
// colour's pass lasts as long as its pair with the most rows (points), so gathering those pairs in
// the first colours shortens the sum of the passes (about a fifth for settled box piles).  Claims
// by plain index would make long chains (pairs come out of the sweep sorted by body, so neighbours
// in the constraint graph have neighbouring indices and each sub-round frees only the next link);
// mixed indices settle a colour in a few sub-rounds (70 instead of 738 for the 10k box rain at
// rest).  Pair indices are below 2^28 (a world's pair capacity is capped there).
constexpr int PRIO_INDEX_BITS = 28;
__host__ __device__ inline int prio(int p, int npts) {
    const uint32_t m = (1u << PRIO_INDEX_BITS) - 1u;
    uint32_t x = (uint32_t)p & m;
    x = (x * 0x9E3779B1u) & m;
    x ^= x >> 15;
    x = (x * 0x2C1B3C6Du) & m;
    x ^= x >> 12;
    x = (x * 0x297A2D39u) & m;
    x ^= x >> 15;
    const int k = npts < 1 ? 1 : (npts > 8 ? 8 : npts);
    return (int)((uint32_t)(8 - k) << PRIO_INDEX_BITS | x);
}
}

template <bool LDS>
__device__ __forceinline__ int claim_get(int* claim, int i) {
    if (LDS) return claim[i];
    return __hip_atomic_load(&claim[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// bufs: three lists of pair capacity; bufs[0] holds the nact active pairs on entry (pair store
// indices; pbase is the world's first, so p - pbase is the index in the world).
// claim is indexed by (body - off).  Writes the solve order and colour starts; returns colours.
// The lists are compacted with atomics, so the order inside a colour varies from run to run;
// the colour classes do not, and neither does any result (pairs of a colour are independent).
template <bool LDS>
__device__ int colour_pairs(const BodyView& b, const PairStore& ps, int* bufs[3], int nact, int* order, int* cs,
                            int* claim, int off, int pbase, int* wsum) {
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
                const int pr = prio(p - pbase, ps.npts[p]);
                if (!b.iat(I_KIN, a)) atomicMin(&claim[a - off], pr);
                if (!b.iat(I_KIN, c)) atomicMin(&claim[c - off], pr);
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
                    const int pr = prio(p - pbase, ps.npts[p]);
                    blk = (!ka && ca == TAKEN) || (!kc && cc == TAKEN);
                    win = !blk && (ka || ca == pr) && (kc || cc == pr);
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


// colour_pairs on LDS data for a group of worlds solved together: pairs by local index, their
// bodies by local index (pab: A | B << 16), kinematic bodies have a negative inverse mass in
// bs[6 n + i]; pri[p] is the pair's claim priority, colour::prio of its index in its own world.
// Bodies of different worlds are disjoint, so the colours of the group are the union of each
// world's own colours (the same sub-rounds, the same claims) and every world's colour classes are
// those it gets alone.  All lists live in LDS; cs needs ncolours + 1 slots.
__device__ int colour_group(const int* pab, const float* imk, const int* pri, int* bufs[3], int nact, int* order,
                            int* cs, int* claim) {
    using namespace colour;
    __shared__ int cnt[3];
    const int tid = threadIdx.x, B = blockDim.x;
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
                const int a = pab[p] & 0xffff, c = pab[p] >> 16;
                if (imk[a] >= 0.0f) atomicMin(&claim[a], pri[p]);
                if (imk[c] >= 0.0f) atomicMin(&claim[c], pri[p]);
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
                    const int a = pab[p] & 0xffff, c = pab[p] >> 16;
                    const bool ka = imk[a] < 0.0f, kc = imk[c] < 0.0f;
                    const int ca = ka ? 0 : claim[a];
                    const int cc = kc ? 0 : claim[c];
                    blk = (!ka && ca == TAKEN) || (!kc && cc == TAKEN);
                    win = !blk && (ka || ca == pri[p]) && (kc || cc == pri[p]);
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
            for (int q = tid; q < ncand; q += B) {
                const int p = cand[q];
                const int a = pab[p] & 0xffff, c = pab[p] >> 16;
                if (imk[a] >= 0.0f && claim[a] != TAKEN) claim[a] = FREE;
                if (imk[c] >= 0.0f && claim[c] != TAKEN) claim[c] = FREE;
            }
            __syncthreads();
            for (int q = sub_begin + tid; q < sub_end; q += B) {
                const int p = order[q];
                const int a = pab[p] & 0xffff, c = pab[p] >> 16;
                if (imk[a] >= 0.0f) claim[a] = TAKEN;
                if (imk[c] >= 0.0f) claim[c] = TAKEN;
            }
            __syncthreads();
            int* t = cand; cand = nxt; nxt = t;
            ncand = nstay;
        }
        const int col_end = cnt[0];
        for (int q = col_begin + tid; q < col_end; q += B) {
            const int p = order[q];
            claim[pab[p] & 0xffff] = FREE;
            claim[pab[p] >> 16] = FREE;
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
