// Joint constraints in the sequential-impulse solve: BallJointConstraint, HingeJointConstraint,
// SliderJointConstraint (src/rigid/constraint/constraint/{ball,hinge,slider}_joint_constraint.cpp).
// initSequentialImpulse builds a JointRow from the two bodies' poses and mass properties after the
// integrate stage; iterateSequentialImpulse runs once per Gauss-Seidel iteration after the contacts
// (sequential_impulse_solver.cpp:66-68).  The arithmetic follows the reference's FP32 Eigen
// evaluation order, as the contact rows do.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_pipeline.hpp"

namespace pe {

enum JointType : int { JT_BALL = 0, JT_HINGE = 1, JT_SLIDER = 2 };
constexpr int JF_LIMITS = 1, JF_MOTOR = 2;

// A joint as World::addConstraint received it: bodies, body-local anchors and axes, hinge limits
// and motor (the setters of {ball,hinge,slider}_joint_constraint.h).
struct JointDef {
    int type, a, b, flags;
    float anchor_a[3], anchor_b[3], axis_a[3], axis_b[3];
    float min_angle, max_angle, target_speed, pad;
};

// The state initSequentialImpulse leaves (the members of the three constraint classes).
struct JointRow {
    f3 ra, rb, wax_a, wt_a0, wt_a1, wt_b0;
    m33 jmj_ball, jmj_rot;
    f3 rhs_ball, rhs_rot;
    float rhs_h[2], jmj_h[2], rhs_pos[2], jmj_pos[2];
    float rhs_motor, rhs_limit, jmj_ml;
    int limit_exceeded;
};

// Constraint::getSkewSymmetricMatrix (constraint.cpp:9-13)
PE_HD m33 skew(f3 v) { return m33{{0.0f, -v.z, v.y, v.z, 0.0f, -v.x, -v.y, v.x, 0.0f}}; }
PE_HD m33 madd(const m33& A, const m33& B) {
    m33 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.m[k] = A.m[k] + B.m[k];
    return r;
}
// Eigen's 3x3 inverse (InverseImpl.h, cofactor form): det = c0 m00 + (c1 m10 + c2 m20), entry (i, j) =
// cofactor(j, i) / det computed as cofactor * (1 / det).
PE_HD m33 inverse3(const m33& M) {
    auto m = [&](int i, int j) { return M.m[3 * i + j]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
    };
    const float c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
    const float det = c0 * m(0, 0) + (c1 * m(1, 0) + c2 * m(2, 0));
    const float inv = 1.0f / det;
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[3 * i + j] = cof(j, i) * inv;
    return r;
}
// (I (ima + imb) + rxA Ia rxA^T + rxB Ib rxB^T)^-1: the ball part's effective mass (ball_joint_constraint.cpp:13-16)
PE_HD m33 ball_jmj(float ims, f3 ra, const m33& Ia, f3 rb, const m33& Ib) {
    m33 S;
#pragma unroll
    for (int k = 0; k < 9; k++) S.m[k] = (k % 4 == 0 ? 1.0f : 0.0f) * ims;
    const m33 xa = skew(ra), xb = skew(rb);
    return inverse3(madd(madd(S, mmt(mm(xa, Ia), xa)), mmt(mm(xb, Ib), xb)));
}
// RigidBody::applyTempAngularImpulse (rigidbody.cpp:178-181)
PE_HD void apply_temp_angular(f3& tw, bool kin, const m33& Iw, f3 J) {
    if (!kin) tw = add(tw, mv(Iw, J));
}

// initSequentialImpulse of joint j: the poses (collision records) and mass properties (solver
// records, zero inverse mass and inertia for a kinematic body) of its two bodies.
__device__ __forceinline__ void joint_setup(const BodyView& b, const JointDef& j, JointRow& r, float dt) {
    const CollBody CA = load_coll(b, j.a), CB = load_coll(b, j.b);
    const SolvBody A = load_solv(b, j.a), B = load_solv(b, j.b);
    const float kerp = 0.2f;
    const f3 anc_a{j.anchor_a[0], j.anchor_a[1], j.anchor_a[2]}, anc_b{j.anchor_b[0], j.anchor_b[1], j.anchor_b[2]};
    const f3 ax_a{j.axis_a[0], j.axis_a[1], j.axis_a[2]}, ax_b{j.axis_b[0], j.axis_b[1], j.axis_b[2]};
    r.ra = mv(CA.P.R, anc_a);
    r.rb = mv(CB.P.R, anc_b);
    const float ims = A.im + B.im;
    const f3 rel = sub(xf_apply(CA.P, anc_a), xf_apply(CB.P, anc_b));
    r.limit_exceeded = 0;
    if (j.type == JT_BALL || j.type == JT_HINGE) {   // ball_joint_constraint.cpp:6-19, hinge_joint_constraint.cpp:25-34
        r.jmj_ball = ball_jmj(ims, r.ra, A.Iw, r.rb, B.Iw);
        r.rhs_ball = mul(mv(r.jmj_ball, rel), -kerp / dt);
    }
    f3 wax_b{0, 0, 0}, wt_b1{0, 0, 0};
    if (j.type == JT_HINGE || j.type == JT_SLIDER) {
        r.wax_a = mv(CA.P.R, ax_a);
        wax_b = mv(CB.P.R, ax_b);
        f3 t0, t1;
        ortho_units(ax_a, t0, t1);
        r.wt_a0 = mv(CA.P.R, t0);
        r.wt_a1 = mv(CA.P.R, t1);
        if (j.type == JT_SLIDER || (j.flags & JF_LIMITS)) {
            ortho_units(ax_b, t0, t1);
            r.wt_b0 = mv(CB.P.R, t0);
            wt_b1 = mv(CB.P.R, t1);
        }
    }
    if (j.type == JT_HINGE) {   // hinge_joint_constraint.cpp:36-70
        const m33 Is = madd(A.Iw, B.Iw);
        r.jmj_h[0] = 1.0f / dot(r.wt_a0, mv(Is, r.wt_a0));
        r.jmj_h[1] = 1.0f / dot(r.wt_a1, mv(Is, r.wt_a1));
        const f3 u = cross(r.wax_a, wax_b);
        r.rhs_h[0] = r.jmj_h[0] * (dot(u, r.wt_a0) * kerp / dt);
        r.rhs_h[1] = r.jmj_h[1] * (dot(u, r.wt_a1) * kerp / dt);
        if (j.flags & (JF_MOTOR | JF_LIMITS)) r.jmj_ml = 1.0f / dot(r.wax_a, mv(Is, r.wax_a));
        if (j.flags & JF_MOTOR) r.rhs_motor = r.jmj_ml * -j.target_speed;
        if (j.flags & JF_LIMITS) {
            const float ac = dot(r.wt_a0, r.wt_b0);
            const float as = dot(cross(r.wt_a0, r.wt_b0), r.wax_a);
            const float angle = atan2f(as, ac);
            if (angle < j.min_angle) {
                r.limit_exceeded = 1;
                r.rhs_limit = -r.jmj_ml * (j.min_angle - angle) * kerp / dt;
            } else if (angle > j.max_angle) {
                r.limit_exceeded = 1;
                r.rhs_limit = -r.jmj_ml * (j.max_angle - angle) * kerp / dt;
            }
        }
    }
    if (j.type == JT_SLIDER) {   // slider_joint_constraint.cpp:24-53
        auto jmj_dir = [&](f3 d) {
            const f3 xa = cross(r.ra, d), xb = cross(r.rb, d);
            return 1.0f / (ims + dot(mv(A.Iw, xa), xa) + dot(mv(B.Iw, xb), xb));
        };
        r.jmj_pos[0] = jmj_dir(r.wt_a0);
        r.jmj_pos[1] = jmj_dir(r.wt_a1);
        // the reference scales the second row by the first row's effective mass (:45)
        r.rhs_pos[0] = r.jmj_pos[0] * (dot(rel, r.wt_a0) * (-kerp / dt));
        r.rhs_pos[1] = r.jmj_pos[0] * (dot(rel, r.wt_a1) * (-kerp / dt));
        r.jmj_rot = inverse3(madd(A.Iw, B.Iw));
        r.rhs_rot = mul(mv(r.jmj_rot, add(cross(r.wt_a0, r.wt_b0), cross(r.wt_a1, wt_b1))), kerp / dt);
    }
}

// iterateSequentialImpulse of joint j on the bodies' temp velocities (solver records); each impulse
// reads the angular velocities the ones before it left, as the reference reads them through
// references to the bodies' members.
__device__ __forceinline__ void joint_iterate(const BodyView& b, const JointDef& j, const JointRow& r) {
    SolvBody A = load_solv(b, j.a), B = load_solv(b, j.b);
    if (j.type == JT_BALL || j.type == JT_HINGE) {
        const f3 va = add(A.tv, cross(A.tw, r.ra));
        const f3 vb = add(B.tv, cross(B.tw, r.rb));
        const f3 imp = sub(r.rhs_ball, mv(r.jmj_ball, sub(va, vb)));
        if (!A.kin) apply_temp(A.tv, A.tw, A.im, A.Iw, r.ra, imp);
        if (!B.kin) apply_temp(B.tv, B.tw, B.im, B.Iw, r.rb, neg(imp));
    }
    if (j.type == JT_HINGE) {
        const f3 dw = sub(A.tw, B.tw);
        const float h0 = r.rhs_h[0] - r.jmj_h[0] * dot(r.wt_a0, dw);
        const float h1 = r.rhs_h[1] - r.jmj_h[1] * dot(r.wt_a1, dw);
        const f3 iv = add(mul(r.wt_a0, h0), mul(r.wt_a1, h1));
        apply_temp_angular(A.tw, A.kin, A.Iw, iv);
        apply_temp_angular(B.tw, B.kin, B.Iw, neg(iv));
        if (j.flags & JF_MOTOR) {
            const float m = r.rhs_motor - r.jmj_ml * dot(r.wax_a, sub(A.tw, B.tw));
            apply_temp_angular(A.tw, A.kin, A.Iw, mul(r.wax_a, m));
            apply_temp_angular(B.tw, B.kin, B.Iw, mul(neg(r.wax_a), m));
        }
        if ((j.flags & JF_LIMITS) && r.limit_exceeded) {
            const float delta = dot(r.wax_a, sub(A.tw, B.tw));
            const float nd = delta * r.rhs_limit > 0 ? 0.0f : delta;   // only resist
            const float l = r.rhs_limit - r.jmj_ml * nd;
            apply_temp_angular(A.tw, A.kin, A.Iw, mul(r.wax_a, l));
            apply_temp_angular(B.tw, B.kin, B.Iw, mul(neg(r.wax_a), l));
        }
    }
    if (j.type == JT_SLIDER) {
        const f3 va = add(A.tv, cross(A.tw, r.ra));
        const f3 vb = add(B.tv, cross(B.tw, r.rb));
        const float p0 = r.rhs_pos[0] - r.jmj_pos[0] * dot(r.wt_a0, sub(va, vb));
        const float p1 = r.rhs_pos[1] - r.jmj_pos[1] * dot(r.wt_a1, sub(va, vb));
        const f3 pv = add(mul(r.wt_a0, p0), mul(r.wt_a1, p1));
        if (!A.kin) apply_temp(A.tv, A.tw, A.im, A.Iw, r.ra, pv);
        if (!B.kin) apply_temp(B.tv, B.tw, B.im, B.Iw, r.rb, neg(pv));
        const f3 rot = neg(mv(r.jmj_rot, sub(A.tw, B.tw)));
        apply_temp_angular(A.tw, A.kin, A.Iw, rot);
        apply_temp_angular(B.tw, B.kin, B.Iw, neg(rot));
    }
    if (!A.kin) {
        float4* d = reinterpret_cast<float4*>(b.sbr(j.a));
        d[0] = make_float4(A.tv.x, A.tv.y, A.tv.z, A.tw.x);
        d[1].x = A.tw.y; d[1].y = A.tw.z;
    }
    if (!B.kin) {
        float4* d = reinterpret_cast<float4*>(b.sbr(j.b));
        d[0] = make_float4(B.tv.x, B.tv.y, B.tv.z, B.tw.x);
        d[1].x = B.tw.y; d[1].y = B.tw.z;
    }
}

}  // namespace pe
