// CPU ORACLE -- test infrastructure only.
//
// A plain C++ restatement of tcy2002/Physics-Engine's per-step rigid-body pipeline
// (integrate -> sweep-and-prune -> narrowphase + contact cache -> sequential impulse),
// templated on the real type so the same code runs in FP32 (the configs' precision)
// and FP64 (the reference default and the residual oracle).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker.  The product (physics-engine_amd/) never links it.
//
// Arithmetic order follows the reference's Eigen 3.4 evaluation exactly, so that a
// g++ build of this file (no FMA contraction) is bit-identical to the reference
// built with the same compiler:
//   FP32: every 3-term sum (dot, M*v, M^T*v, M*N, trace) is p0 + (p1 + p2)
//   FP64: dot / M^T*v / M^T*N are (p0 + p1) + p2; M*v, M*N, M*N^T are
//         (p0 + p1) + p2 in rows 0 and 1 and p0 + (p1 + p2) in row 2
//   quaternion product: the SSE form of Eigen/src/Geometry/arch/Geometry_SIMD.h
//   quaternion normalise: sqrt((x^2 + z^2) + (y^2 + w^2))
// These were pinned against the vendored Eigen by brute force on random inputs.
//
// Parity status per function is listed in the table at the end of this header.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <utility>
#include <memory>
#include <set>

namespace pe_oracle {

enum ShapeType { ST_BOX = 0, ST_SPHERE = 1, ST_CYLINDER = 2, ST_CONVEX = 3, ST_CONCAVE = 4, ST_COMPOUND = 5 };

template <class R> struct Lim;
template <> struct Lim<float>  { static float  max() { return 1e30f; }  static float  min() { return -1e30f; } };
template <> struct Lim<double> { static double max() { return 1e100; } static double min() { return -1e100; } };

// ---------------------------------------------------------------- vector / matrix
template <class R> struct V3 {
    R v[3];
    V3() : v{0, 0, 0} {}
    V3(R a, R b, R c) : v{a, b, c} {}
    R& operator[](int i) { return v[i]; }
    const R& operator[](int i) const { return v[i]; }
    R x() const { return v[0]; } R y() const { return v[1]; } R z() const { return v[2]; }
};
template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
template <class R> inline V3<R> operator-(const V3<R>& a) { return {-a[0], -a[1], -a[2]}; }
template <class R> inline V3<R> operator*(const V3<R>& a, R s) { return {a[0] * s, a[1] * s, a[2] * s}; }
template <class R> inline V3<R> operator*(R s, const V3<R>& a) { return {s * a[0], s * a[1], s * a[2]}; }
template <class R> inline V3<R> operator/(const V3<R>& a, R s) { return {a[0] / s, a[1] / s, a[2] / s}; }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// 3-term sums in Eigen's order (see header comment)
inline float  sum_dot(float a, float b, float c) { return a + (b + c); }
inline double sum_dot(double a, double b, double c) { return (a + b) + c; }
inline float  sum_row(int, float a, float b, float c) { return a + (b + c); }
inline double sum_row(int row, double a, double b, double c) { return row < 2 ? (a + b) + c : a + (b + c); }

template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return sum_dot(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
template <class R> inline R sqnorm(const V3<R>& a) { return dot(a, a); }
template <class R> inline R norm(const V3<R>& a) { return std::sqrt(sqnorm(a)); }
template <class R> inline V3<R> normalized(const V3<R>& a) {
    R n = sqnorm(a);
    if (n > R(0)) { R s = std::sqrt(n); return {a[0] / s, a[1] / s, a[2] / s}; }
    return a;
}

template <class R> struct M3 {
    R m[3][3];  // m[row][col]
    M3() { std::memset(m, 0, sizeof(m)); }
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = R(1); return r; }
    R operator()(int i, int j) const { return m[i][j]; }
    R& operator()(int i, int j) { return m[i][j]; }
    V3<R> col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
};
template <class R> inline V3<R> mv(const M3<R>& A, const V3<R>& v) {
    V3<R> r;
    for (int i = 0; i < 3; i++) r[i] = sum_row(i, A(i, 0) * v[0], A(i, 1) * v[1], A(i, 2) * v[2]);
    return r;
}
template <class R> inline V3<R> mtv(const M3<R>& A, const V3<R>& v) {
    V3<R> r;
    for (int i = 0; i < 3; i++) r[i] = sum_dot(A(0, i) * v[0], A(1, i) * v[1], A(2, i) * v[2]);
    return r;
}
template <class R> inline M3<R> mm(const M3<R>& A, const M3<R>& B) {
    M3<R> r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        r(i, j) = sum_row(i, A(i, 0) * B(0, j), A(i, 1) * B(1, j), A(i, 2) * B(2, j));
    return r;
}
template <class R> inline M3<R> mtm(const M3<R>& A, const M3<R>& B) {  // A^T * B
    M3<R> r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        r(i, j) = sum_dot(A(0, i) * B(0, j), A(1, i) * B(1, j), A(2, i) * B(2, j));
    return r;
}
template <class R> inline M3<R> mmt(const M3<R>& A, const M3<R>& B) {  // A * B^T
    M3<R> r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        r(i, j) = sum_row(i, A(i, 0) * B(j, 0), A(i, 1) * B(j, 1), A(i, 2) * B(j, 2));
    return r;
}
// Eigen 3x3 inverse (Eigen/src/LU/InverseImpl.h, cofactor form)
template <class R> inline R cof(const M3<R>& m, int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
}
template <class R> inline M3<R> inverse3(const M3<R>& m) {
    R c0 = cof(m, 0, 0), c1 = cof(m, 1, 0), c2 = cof(m, 2, 0);
    R det = sum_dot(c0 * m(0, 0), c1 * m(1, 0), c2 * m(2, 0));
    R inv = R(1) / det;
    M3<R> r;
    r(0, 1) = 0;
    R c01 = cof(m, 0, 1) * inv, c11 = cof(m, 1, 1) * inv, c02 = cof(m, 0, 2) * inv;
    r(1, 2) = cof(m, 2, 1) * inv;
    r(2, 1) = cof(m, 1, 2) * inv;
    r(2, 2) = cof(m, 2, 2) * inv;
    r(1, 0) = c01; r(1, 1) = c11; r(2, 0) = c02;
    r(0, 0) = c0 * inv; r(0, 1) = c1 * inv; r(0, 2) = c2 * inv;
    return r;
}

// ---------------------------------------------------------------- quaternion (x,y,z,w)
template <class R> struct Quat { R x, y, z, w; };
// Eigen quaternionbase_assign_impl<Other,3,3> (Shoemake)
template <class R> inline Quat<R> quat_from_matrix(const M3<R>& m) {
    Quat<R> q;
    R t = m(0, 0) + (m(1, 1) + m(2, 2));
    if (t > R(0)) {
        t = std::sqrt(t + R(1));
        q.w = R(0.5) * t;
        t = R(0.5) / t;
        q.x = (m(2, 1) - m(1, 2)) * t;
        q.y = (m(0, 2) - m(2, 0)) * t;
        q.z = (m(1, 0) - m(0, 1)) * t;
    } else {
        int i = 0;
        if (m(1, 1) > m(0, 0)) i = 1;
        if (m(2, 2) > m(i, i)) i = 2;
        int j = (i + 1) % 3, k = (j + 1) % 3;
        R c[3];
        t = std::sqrt(m(i, i) - m(j, j) - m(k, k) + R(1));
        c[i] = R(0.5) * t;
        t = R(0.5) / t;
        q.w = (m(k, j) - m(j, k)) * t;
        c[j] = (m(j, i) + m(i, j)) * t;
        c[k] = (m(k, i) + m(i, k)) * t;
        q.x = c[0]; q.y = c[1]; q.z = c[2];
    }
    return q;
}
inline Quat<float> quat_mul(const Quat<float>& a, const Quat<float>& b) {
    Quat<float> r;
    r.x = (a.x * b.w - a.z * b.y) + (a.y * b.z + a.w * b.x);
    r.y = (a.y * b.w - a.x * b.z) + (a.z * b.x + a.w * b.y);
    r.z = (a.z * b.w - a.y * b.x) + (a.x * b.y + a.w * b.z);
    r.w = (a.w * b.w - a.x * b.x) - (a.z * b.z + a.y * b.y);
    return r;
}
inline Quat<double> quat_mul(const Quat<double>& a, const Quat<double>& b) {
    Quat<double> r;
    r.x = (a.w * b.x + a.y * b.z) - (a.z * b.y - a.x * b.w);
    r.y = (a.w * b.y + a.y * b.w) + (a.z * b.x - a.x * b.z);
    r.z = (a.w * b.z - a.y * b.x) + (a.z * b.w + a.x * b.y);
    r.w = (a.w * b.w - a.y * b.y) - (a.z * b.z + a.x * b.x);
    return r;
}
template <class R> inline void quat_normalize(Quat<R>& q) {
    R n = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
    if (n > R(0)) { q.x /= n; q.y /= n; q.z /= n; q.w /= n; }
}
template <class R> inline M3<R> quat_to_matrix(const Quat<R>& q) {
    const R tx = R(2) * q.x, ty = R(2) * q.y, tz = R(2) * q.z;
    const R twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const R txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const R tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    M3<R> r;
    r(0, 0) = R(1) - (tyy + tzz); r(0, 1) = txy - twz; r(0, 2) = txz + twy;
    r(1, 0) = txy + twz; r(1, 1) = R(1) - (txx + tzz); r(1, 2) = tyz - twx;
    r(2, 0) = txz - twy; r(2, 1) = tyz + twx; r(2, 2) = R(1) - (txx + tyy);
    return r;
}

// ---------------------------------------------------------------- transform
template <class R> struct Xf {
    M3<R> basis; V3<R> origin;
    V3<R> apply(const V3<R>& p) const { return mv(basis, p) + origin; }             // transform.cpp:2-4
    V3<R> inv_apply(const V3<R>& p) const { return mtv(basis, p - origin); }        // transform.cpp:73-76
};

// ---------------------------------------------------------------- bodies
template <class R> struct Poly;
template <class R> struct Body {
    int type = ST_BOX;
    R size[3] = {0, 0, 0};        // box full size
    R half[3] = {0, 0, 0};        // box half size (size / 2)
    R radius = 0, height = 0;     // sphere / cylinder
    R mass = 1, inv_mass = 1;
    M3<R> local_inertia = M3<R>::identity(), local_inv_inertia = M3<R>::identity();
    M3<R> world_inv_inertia = M3<R>::identity();
    Xf<R> xf{M3<R>::identity(), V3<R>()};
    V3<R> v, w, tv, tw;
    V3<R> aabb_min, aabb_max;
    R mu = R(0.5), e = R(0.5), lin_damp = 0, ang_damp = 0;
    bool kinematic = false, ignore = false;
    const Poly<R>* poly = nullptr;   // mesh of a box / cylinder / convex / concave body
    // World::updateObjectStatus bookkeeping (rigidbody.h:36, 72-83): asleep, seconds below the
    // sleep thresholds, lifetime and time lived, contacts with static / dynamic neighbours in the
    // last narrowphase; removed once its lifetime is over (or World::removeRigidBody)
    bool sleep = false, removed = false;
    R sleep_time = 0, life_time = Lim<R>::max(), last_time = 0;
    int static_count = 0, dynamic_count = 0;
    bool report = false;             // has collision callbacks (RigidBody::addCollisionCallback)
    // a compound body (CompoundShape): its children as bodies whose xf is the local transform, their
    // mass ratios, and the unit-mass inertia CompoundShape::init computed
    std::vector<Body> kids;
    std::vector<R> ratios;
    M3<R> compound_inertia;
};

template <class R> M3<R> poly_inertia(const Poly<R>& p);
template <class R> void poly_aabb(const Poly<R>& p, const M3<R>& rot, const V3<R>& pos, V3<R>& mn, V3<R>& mx);
// shape local inertia at unit mass: box_shape.cpp:16-22, sphere_shape.cpp:5-12, cylinder_shape.cpp:23-31
template <class R> inline M3<R> shape_inertia(const Body<R>& b) {
    M3<R> I;
    if (b.type == ST_BOX) {
        const R x2 = b.size[0] * b.size[0], y2 = b.size[1] * b.size[1], z2 = b.size[2] * b.size[2];
        I(0, 0) = (y2 + z2) / R(12); I(1, 1) = (x2 + z2) / R(12); I(2, 2) = (x2 + y2) / R(12);
    } else if (b.type == ST_SPHERE) {
        const R i = R(2.0 / 5.0) * b.radius * b.radius;
        I(0, 0) = I(1, 1) = I(2, 2) = i;
    } else if (b.type == ST_CYLINDER) {
        const R r_2 = b.radius * b.radius, h_2 = b.height * b.height;
        const R axis = r_2 * R(0.5), diag = r_2 / R(4.0) + h_2 / R(12.0);
        I(0, 0) = diag; I(1, 1) = axis; I(2, 2) = diag;
    } else if (b.type == ST_COMPOUND) {
        I = b.compound_inertia;      // CompoundShape::init
    } else {
        I = poly_inertia(*b.poly);   // convex_mesh_shape.cpp:93-100 / concave_mesh_shape.cpp:9-10
    }
    return I;
}
// CompoundShape::init (compound_shape.cpp:24-47): the mass-ratio-weighted centroid of the children's
// origins; every child moved by minus it; the unit-mass inertia sum over the children of
// (R I_child R^T + |p|^2 1 - p p^T) * ratio / total.  Returns the centroid.
template <class R> inline V3<R> compound_init(Body<R>& b) {
    R total = 0;
    for (R r : b.ratios) total += r;
    V3<R> c;
    for (size_t k = 0; k < b.kids.size(); k++) c = c + b.kids[k].xf.origin * b.ratios[k];
    c = c / total;
    M3<R> I;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) I(i, j) = 0;
    for (size_t k = 0; k < b.kids.size(); k++) {
        Body<R>& kid = b.kids[k];
        const M3<R> si = shape_inertia(kid);
        const V3<R> p = kid.xf.origin - c;
        kid.xf.origin = p;
        const M3<R>& rot = kid.xf.basis;
        const R pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        M3<R> t;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) t(i, j) = (i == j ? R(1) : R(0)) * pp;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) t(i, j) -= p[i] * p[j];
        const M3<R> T = mm(rot, si);
        const R q = b.ratios[k] / total;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const R a = T(i, 0) * rot(j, 0) + (T(i, 1) * rot(j, 1) + T(i, 2) * rot(j, 2));
                I(i, j) += (a + t(i, j)) * q;
            }
    }
    b.compound_inertia = I;
    return c;
}

// RigidBody::setMass then setCollisionShape then setTransform (rigidbody.cpp:11-46)
template <class R> inline void body_set_mass_props(Body<R>& b, R mass) {
    const R eps = R(0.00001);
    b.mass = mass;
    b.inv_mass = mass > eps ? R(1.0) / mass : R(0);
    M3<R> I = shape_inertia(b);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) b.local_inertia(i, j) = I(i, j) * mass;
    b.local_inv_inertia = mass > eps ? inverse3(b.local_inertia) : M3<R>();
}
// RigidBody::updateWorldInertia (rigidbody.cpp:43-46): (R * L) * R^T.  In this expression the
// reference (FP64 as well as FP32) sums every entry of the second product as p0 + (p1 + p2);
// pinned against the compiled reference's own I_w values.
template <class R> inline void body_update_world_inertia(Body<R>& b) {
    const M3<R> T = mm(b.xf.basis, b.local_inv_inertia);
    const M3<R>& Rb = b.xf.basis;
    M3<R> r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r(i, j) = T(i, 0) * Rb(j, 0) + (T(i, 1) * Rb(j, 1) + T(i, 2) * Rb(j, 2));
    b.world_inv_inertia = r;
}

// Shape::getAABB
template <class R> inline void body_aabb(Body<R>& b) {
    const M3<R>& rot = b.xf.basis;
    const V3<R>& pos = b.xf.origin;
    if (b.type == ST_BOX) {  // box_shape.cpp:27-43
        V3<R> mn(Lim<R>::max(), Lim<R>::max(), Lim<R>::max()), mx(Lim<R>::min(), Lim<R>::min(), Lim<R>::min());
        for (int i = 0; i < 8; i++) {
            V3<R> p(i & 1 ? b.half[0] : -b.half[0], i & 2 ? b.half[1] : -b.half[1], i & 4 ? b.half[2] : -b.half[2]);
            V3<R> v = mv(rot, p);
            for (int k = 0; k < 3; k++) {
                mn[k] = mn[k] < v[k] ? mn[k] : v[k];
                mx[k] = mx[k] > v[k] ? mx[k] : v[k];
            }
        }
        b.aabb_min = mn + pos; b.aabb_max = mx + pos;
    } else if (b.type == ST_SPHERE) {  // sphere_shape.cpp:14-18
        V3<R> r(b.radius, b.radius, b.radius);
        b.aabb_min = pos - r; b.aabb_max = pos + r;
    } else if (b.type == ST_CYLINDER) {  // cylinder_shape.cpp:34-43
        V3<R> axis = rot.col(1);
        const R hh = b.height * R(0.5);
        V3<R> ext(std::abs(axis[0]) * hh, std::abs(axis[1]) * hh, std::abs(axis[2]) * hh);
        ext[1] += b.radius * std::sqrt(axis[0] * axis[0] + axis[2] * axis[2]);
        ext[0] += b.radius * std::sqrt(axis[1] * axis[1] + axis[2] * axis[2]);
        ext[2] += b.radius * std::sqrt(axis[0] * axis[0] + axis[1] * axis[1]);
        b.aabb_min = pos - ext; b.aabb_max = pos + ext;
    } else if (b.type == ST_COMPOUND) {  // compound_shape.cpp:49-58: the union of the children's
        V3<R> mn(Lim<R>::max(), Lim<R>::max(), Lim<R>::max()), mx(Lim<R>::min(), Lim<R>::min(), Lim<R>::min());
        for (const Body<R>& kid : b.kids) {
            Body<R> t = kid;
            t.xf = Xf<R>{mm(b.xf.basis, kid.xf.basis), mv(b.xf.basis, kid.xf.origin) + b.xf.origin};
            body_aabb(t);
            for (int k = 0; k < 3; k++) {
                mn[k] = mn[k] < t.aabb_min[k] ? mn[k] : t.aabb_min[k];
                mx[k] = mx[k] > t.aabb_max[k] ? mx[k] : t.aabb_max[k];
            }
        }
        b.aabb_min = mn; b.aabb_max = mx;
    } else {  // convex_mesh_shape.cpp:138-150 (the concave mesh inherits it)
        poly_aabb(*b.poly, rot, pos, b.aabb_min, b.aabb_max);
    }
}

// RigidBody::getAABBScale (rigidbody.cpp:119-123)
template <class R> inline R body_aabb_scale(const Body<R>& b) {
    const R scale = norm(b.aabb_max - b.aabb_min);
    const R diff = norm((b.aabb_max + b.aabb_min) - b.xf.origin * R(2));
    return (scale + diff) * R(0.5);
}

// RigidBody::step (rigidbody.cpp:207-243), PE_USE_QUATERNION branch
template <class R> inline void body_step(Body<R>& b, R dt) {
    if (b.kinematic) { body_aabb(b); return; }
    b.xf.origin = b.xf.origin + b.v * dt;
    Quat<R> q = quat_from_matrix(b.xf.basis);
    const V3<R> dr = (b.w * dt) * R(0.5);
    Quat<R> a{dr[0], dr[1], dr[2], R(0)};
    const Quat<R> dq = quat_mul(a, q);
    q = Quat<R>{q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w};
    quat_normalize(q);
    b.xf.basis = quat_to_matrix(q);
    body_update_world_inertia(b);
    body_aabb(b);
}
// RigidBody::applyDamping (rigidbody.cpp:201-205)
template <class R> inline void body_damping(Body<R>& b, R dt) {
    if (b.kinematic) return;
    const R pl = std::pow(R(1.0) - b.lin_damp, dt), pa = std::pow(R(1.0) - b.ang_damp, dt);
    b.v = b.v * pl; b.w = b.w * pa;
}
// World::applyExternalForce + RigidBody::applyForce (world.cpp:121-127, rigidbody.cpp:193-199)
// The torque term is I_w^-1 (0 * dt): adding it leaves w unchanged, so it is omitted.
template <class R> inline void body_gravity(Body<R>& b, const V3<R>& g, R dt) {
    if (b.kinematic) return;
    const V3<R> f = g * b.mass;
    b.v = b.v + (f * b.inv_mass) * dt;
}

// ---------------------------------------------------------------- contacts
template <class R> struct ContactPoint {
    V3<R> wpos{Lim<R>::max(), Lim<R>::max(), Lim<R>::max()};
    V3<R> normal{0, 1, 0};
    V3<R> la, lb;
    R dist = Lim<R>::max();
    V3<R> t0, t1;
    bool valid() const { return wpos[0] < Lim<R>::max() + R(0.1); }
};

// ContactPoint::getOrthoUnits (contact_result.cpp:8-31)
template <class R> inline void ortho_units(V3<R> n, V3<R>& t1, V3<R>& t2) {
    n = normalized(n);
    if (std::abs(n[2]) > R(0.7071)) {
        const R a = n[1] * n[1] + n[2] * n[2];
        const R k = R(1.0) / std::sqrt(a);
        t1[0] = 0; t1[1] = -n[2] * k; t1[2] = n[1] * k;
        t2[0] = a * k; t2[1] = -n[0] * t1[2]; t2[2] = n[0] * t1[1];
    } else {
        const R a = n[0] * n[0] + n[1] * n[1];
        const R k = R(1.0) / std::sqrt(a);
        t1[0] = -n[1] * k; t1[1] = n[0] * k; t1[2] = 0;
        t2[0] = -n[2] * t1[1]; t2[1] = n[2] * t1[0]; t2[2] = a * k;
    }
}

template <class R> struct ContactResult {
    int a = -1, b = -1;
    R mu = 0, e = 0;
    int npts = 0;
    bool swap = false;
    ContactPoint<R> pts[8];
};

// ---------------------------------------------------------------- world
template <class R> struct Joint;
// What World::execCollisionCallbacks (world.cpp:129-192) hands the callbacks of a contact result with
// points: A's callbacks get (A, B, pos + nor * depth, -nor, -vel), B's get (B, A, pos, nor, vel).
template <class R> struct CollisionEvent {
    int a, b;
    V3<R> pos, nor, vel;
    R depth;
};
template <class R> struct World {
    std::vector<Body<R>> bodies;
    std::vector<Joint<R>> joints;             // World::addConstraint order
    std::vector<int> joint_order;             // joint sweep order (colour order when use_color_order)
    std::set<std::pair<int, int>> ignore;     // body pairs (lo, hi) that never collide (addIgnoreCollisionId)
    V3<R> gravity{0, R(-9.8), 0};
    R dt = R(0.01);
    int iterations = 10;
    int axis = 0;                             // BroadPhaseSweepAndPrune::_target_axis
    std::vector<std::pair<int, int>> pairs;   // (A, B) body indices, in SAP order
    std::vector<ContactResult<R>> results;    // one per pair, persistent slots
    std::vector<int> solve_order;             // permutation of result indices for the GS sweep
    bool use_color_order = true;
    int ncolors = 0;
    std::vector<int> color_of;                // per result
    std::vector<std::shared_ptr<Poly<R>>> polys;   // meshes the bodies point at
    R sleep_lin2 = 0, sleep_ang2 = 0, sleep_time = 0;   // World::setSleep*Threshold (world.h:26-28)
    std::vector<CollisionEvent<R>> events;    // the last narrowphase's collision reports
};

// ContactResult::addContactPoint (contact_result.cpp:88-142)
template <class R> inline R same_point_threshold(const World<R>& w, const ContactResult<R>& cr) {
    const R sa = body_aabb_scale(w.bodies[cr.a]), sb = body_aabb_scale(w.bodies[cr.b]);
    return (sa < sb ? sa : sb) * R(0.02);
}
template <class R> inline void add_contact_point(const World<R>& w, ContactResult<R>& cr,
                                                 const V3<R>& world_normal, const V3<R>& world_pos, R depth) {
    V3<R> point_a = world_pos, point_b = world_pos, n = world_normal, point = world_pos;
    if (cr.swap) {
        n = -n;
        point_b = world_pos + world_normal * depth;
        point = point_b;
    } else {
        point_a = world_pos + world_normal * depth;
    }
    const Body<R>& A = w.bodies[cr.a];
    const Body<R>& B = w.bodies[cr.b];
    const V3<R> la = A.xf.inv_apply(point_a);
    const V3<R> lb = B.xf.inv_apply(point_b);
    auto make = [&](ContactPoint<R>& cp) {
        cp.wpos = point; cp.normal = n; cp.la = la; cp.lb = lb; cp.dist = depth;
        ortho_units(n, cp.t0, cp.t1);
    };
    // getExistingClosestPoint (contact_result.cpp:166-180)
    R th = same_point_threshold(w, cr);
    R min_dist = th * th;
    int near = -1;
    for (int i = 0; i < 8; i++) {
        if (!cr.pts[i].valid()) continue;
        const V3<R> d = cr.pts[i].lb - lb;
        const R dd = dot(d, d);
        if (dd < min_dist) { min_dist = dd; near = i; }
    }
    if (near >= 0) {
        if (depth < cr.pts[near].dist) make(cr.pts[near]);
        return;
    }
    for (int i = 0; i < 8; i++) {
        if (!cr.pts[i].valid()) { make(cr.pts[i]); return; }
    }
    R max_dist = Lim<R>::min();
    int idx = -1;
    for (int i = 0; i < 8; i++) {
        if (cr.pts[i].dist > max_dist) { max_dist = cr.pts[i].dist; idx = i; }
    }
    if (idx >= 0) make(cr.pts[idx]);
}
// ContactResult::sortContactPoints (contact_result.cpp:144-156): libstdc++ std::sort on 8
// elements is a plain insertion sort (stable).  Invalidated slots keep the distance they had,
// so a stale distance can sort ahead of valid points and truncate point_size.
template <class R> inline void sort_contact_points(ContactResult<R>& cr) {
    ContactPoint<R>* p = cr.pts;
    for (int i = 1; i < 8; i++) {
        ContactPoint<R> val = p[i];
        if (val.dist < p[0].dist) {
            for (int k = i; k > 0; k--) p[k] = p[k - 1];
            p[0] = val;
        } else {
            int k = i;
            while (val.dist < p[k - 1].dist) { p[k] = p[k - 1]; k--; }
            p[k] = val;
        }
    }
    cr.npts = 8;
    for (int i = 0; i < 8; i++) if (!p[i].valid()) { cr.npts = i; break; }
}
// ContactResult::clearContactPoints (contact_result.cpp:158-164)
template <class R> inline void clear_contact_points(ContactResult<R>& cr) {
    for (int i = 0; i < 8; i++) cr.pts[i].wpos = V3<R>(Lim<R>::max(), Lim<R>::max(), Lim<R>::max());
    cr.npts = 0;
    cr.swap = false;
}

// ---------------------------------------------------------------- box-box (box_box_collision_algorithm.cpp)
template <class R> inline int intersect_rect_quad2(const R h[2], R p[8], R ret[16]) {  // :20-71
    int nq = 4, nr = 0;
    R buffer[16];
    R* q = p;
    R* r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            R* pq = q;
            R* pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if (R(sign) * pq[dir] < h[dir]) {
                    pr[0] = pq[0]; pr[1] = pq[1]; pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                R* next_q = (i > 1) ? pq + 2 : q;
                if ((R(sign) * pq[dir] < h[dir]) ^ (R(sign) * next_q[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] + (next_q[1 - dir] - pq[1 - dir]) / (next_q[dir] - pq[dir]) *
                                                    (R(sign) * h[dir] - pq[dir]);
                    pr[dir] = R(sign) * h[dir];
                    pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                pq += 2;
            }
            q = r;
            r = q == ret ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret) std::memcpy(ret, q, nr * 2 * sizeof(R));
    return nr;
}

template <class R> inline void cull_points2(int n, R p[], int m, int i0, int i_ret[]) {  // :73-135
    const R PI = R(3.141592653589);
    int i;
    R a, cx, cy, q;
    if (n == 1) { cx = p[0]; cy = p[1]; }
    else if (n == 2) { cx = R(0.5) * (p[0] + p[2]); cy = R(0.5) * (p[1] + p[3]); }
    else {
        a = 0; cx = 0; cy = 0;
        for (i = 0; i < n - 1; i++) {
            const R p0 = p[i * 2], p1 = p[i * 2 + 1], p2 = p[i * 2 + 2], p3 = p[i * 2 + 3];
            q = p0 * p3 - p2 * p1;
            a += q;
            cx += q * (p0 + p2);
            cy += q * (p1 + p3);
        }
        const R p0 = p[0], p1 = p[1], pm1 = p[n * 2 - 1], pm2 = p[n * 2 - 2];
        q = pm2 * p1 - p0 * pm1;
        if (std::abs(a + q) > R(0.00001)) a = R(1.0) / (R(3.0) * (a + q));
        else a = Lim<R>::max();
        cx = a * (cx + q * (pm2 + p0));
        cy = a * (cy + q * (pm1 + p1));
    }
    R A[8]; int avail[8];
    for (i = 0; i < n; i++) { A[i] = std::atan2(p[i * 2 + 1] - cy, p[i * 2] - cx); avail[i] = 1; }
    avail[i0] = 0;
    i_ret[0] = i0;
    i_ret++;
    for (int j = 1; j < m; j++) {
        a = R(j) * (2 * PI / R(m)) + A[i0];
        if (a > PI) a -= 2 * PI;
        R max_diff = Lim<R>::max();
        *i_ret = i0;
        for (i = 0; i < n; i++) {
            if (avail[i]) {
                R diff = std::abs(A[i] - a);
                if (diff > PI) diff = 2 * PI - diff;
                if (diff < max_diff) { max_diff = diff; *i_ret = i; }
            }
        }
        avail[*i_ret] = 0;
        i_ret++;
    }
}

template <class R> inline bool box_box(const World<R>& w, ContactResult<R>& result,
                                       const Body<R>& ba, const Body<R>& bb) {  // :137-580
    const Xf<R>& trans_a = ba.xf;
    const Xf<R>& trans_b = bb.xf;
    const R eps = R(0.00001);
    V3<R> normal;
    R depth;
    int max_c = 4;
    const R fudge_factor = R(1.05);
    bool use_normalR = false;
    V3<R> p, pp, normalC(0, 0, 0), normalR;
    R A[3], B[3], Q11, Q12, Q13, Q21, Q22, Q23, Q31, Q32, Q33, s, s2, l;
    int i, j, invert_normal, code;
    const V3<R>& p1 = trans_a.origin;
    const V3<R>& p2 = trans_b.origin;
    const M3<R>& R1 = trans_a.basis;
    const M3<R>& R2 = trans_b.basis;
    p = p2 - p1;
    pp = mtv(R1, p);
    A[0] = ba.size[0] * R(0.5); A[1] = ba.size[1] * R(0.5); A[2] = ba.size[2] * R(0.5);
    B[0] = bb.size[0] * R(0.5); B[1] = bb.size[1] * R(0.5); B[2] = bb.size[2] * R(0.5);
    const M3<R> Rij = mtm(R1, R2);
    Q11 = std::abs(Rij(0, 0)); Q12 = std::abs(Rij(0, 1)); Q13 = std::abs(Rij(0, 2));
    Q21 = std::abs(Rij(1, 0)); Q22 = std::abs(Rij(1, 1)); Q23 = std::abs(Rij(1, 2));
    Q31 = std::abs(Rij(2, 0)); Q32 = std::abs(Rij(2, 1)); Q33 = std::abs(Rij(2, 2));
    s = Lim<R>::min();
    invert_normal = 0;
    code = 0;
#define PEO_TST(expr1, expr2, norm, cc)      \
    s2 = std::abs(expr1) - (expr2);          \
    if (s2 > 0) return false;                \
    if (s2 > s) {                            \
        s = s2; normalR = norm; use_normalR = true; \
        invert_normal = ((expr1) < 0); code = (cc); \
    }
    PEO_TST(pp[0], A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13, R1.col(0), 1)
    PEO_TST(pp[1], A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23, R1.col(1), 2)
    PEO_TST(pp[2], A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33, R1.col(2), 3)
    PEO_TST(dot(R2.col(0), p), A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0], R2.col(0), 4)
    PEO_TST(dot(R2.col(1), p), A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1], R2.col(1), 5)
    PEO_TST(dot(R2.col(2), p), A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2], R2.col(2), 6)
#undef PEO_TST
#define PEO_TST(expr1, expr2, n1, n2, n3, cc)                \
    s2 = std::abs(expr1) - (expr2);                          \
    if (s2 > eps) return false;                              \
    l = std::sqrt((n1) * (n1) + (n2) * (n2) + (n3) * (n3));  \
    if (l > eps) {                                           \
        s2 /= l;                                             \
        if (s2 * fudge_factor > s) {                         \
            s = s2; use_normalR = false;                     \
            normalC[0] = (n1) / l; normalC[1] = (n2) / l; normalC[2] = (n3) / l; \
            invert_normal = ((expr1) < 0); code = (cc);      \
        }                                                    \
    }
    Q11 += eps; Q12 += eps; Q13 += eps;
    Q21 += eps; Q22 += eps; Q23 += eps;
    Q31 += eps; Q32 += eps; Q33 += eps;
    PEO_TST(pp[2] * Rij(1, 0) - pp[1] * Rij(2, 0), (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), R(0), -Rij(2, 0), Rij(1, 0), 7)
    PEO_TST(pp[2] * Rij(1, 1) - pp[1] * Rij(2, 1), (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), R(0), -Rij(2, 1), Rij(1, 1), 8)
    PEO_TST(pp[2] * Rij(1, 2) - pp[1] * Rij(2, 2), (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), R(0), -Rij(2, 2), Rij(1, 2), 9)
    PEO_TST(pp[0] * Rij(2, 0) - pp[2] * Rij(0, 0), (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), Rij(2, 0), R(0), -Rij(0, 0), 10)
    PEO_TST(pp[0] * Rij(2, 1) - pp[2] * Rij(0, 1), (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), Rij(2, 1), R(0), -Rij(0, 1), 11)
    PEO_TST(pp[0] * Rij(2, 2) - pp[2] * Rij(0, 2), (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), Rij(2, 2), R(0), -Rij(0, 2), 12)
    PEO_TST(pp[1] * Rij(0, 0) - pp[0] * Rij(1, 0), (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -Rij(1, 0), Rij(0, 0), R(0), 13)
    PEO_TST(pp[1] * Rij(0, 1) - pp[0] * Rij(1, 1), (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -Rij(1, 1), Rij(0, 1), R(0), 14)
    PEO_TST(pp[1] * Rij(0, 2) - pp[0] * Rij(1, 2), (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -Rij(1, 2), Rij(0, 2), R(0), 15)
#undef PEO_TST
    if (!code) return false;
    if (use_normalR) normal = normalR;
    else normal = mv(R1, normalC);
    if (invert_normal) { normal[0] = -normal[0]; normal[1] = -normal[1]; normal[2] = -normal[2]; }
    depth = -s;

    if (code > 6) {  // edge-edge :281-348
        V3<R> pa = p1;
        R sign;
        for (j = 0; j < 3; j++) {
            sign = dot(normal, R1.col(j)) > 0 ? R(1.0) : R(-1.0);
            pa[0] += sign * A[j] * R1(0, j); pa[1] += sign * A[j] * R1(1, j); pa[2] += sign * A[j] * R1(2, j);
        }
        V3<R> pb = p2;
        for (j = 0; j < 3; j++) {
            sign = dot(normal, R2.col(j)) > 0 ? R(-1.0) : R(1.0);
            pb[0] += sign * B[j] * R2(0, j); pb[1] += sign * B[j] * R2(1, j); pb[2] += sign * B[j] * R2(2, j);
        }
        R alpha, beta;
        V3<R> ua, ub;
        {
            int quotient = (code - 7) / 3;
            ua = R1.col(quotient);
            ub = R2.col(quotient);
        }
        const V3<R> b_a = pb - pa;
        const R ua_ub = dot(ua, ub);
        const R q1 = dot(ua, b_a);
        const R q2 = -dot(ub, b_a);
        R d = 1 - ua_ub * ua_ub;
        if (d <= eps) { alpha = 0; beta = 0; }
        else { d = R(1.0) / d; alpha = (q1 + ua_ub * q2) * d; beta = (ua_ub * q1 + q2) * d; }
        pa[0] += ua[0] * alpha; pa[1] += ua[1] * alpha; pa[2] += ua[2] * alpha;
        pb[0] += ub[0] * beta; pb[1] += ub[1] * beta; pb[2] += ub[2] * beta;
        (void)pa;
        const V3<R> normVec = normal;
        add_contact_point(w, result, -normVec, pb + normVec * R(0), -depth + R(0) * 2);
        return true;
    }

    const R *Sa, *Sb;
    M3<R> Ra, Rb;
    V3<R> pa, pb;
    if (code <= 3) { result.swap = false; Ra = R1; Rb = R2; pa = p1; pb = p2; Sa = A; Sb = B; }
    else { result.swap = true; Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }
    V3<R> normal2, nr, anr;
    if (code <= 3) normal2 = normal;
    else normal2 = -normal;
    nr = mtv(Rb, normal2);
    anr[0] = std::abs(nr[0]); anr[1] = std::abs(nr[1]); anr[2] = std::abs(nr[2]);
    int lan_r, a1, a2;
    if (anr[1] > anr[0]) {
        if (anr[1] > anr[2]) { a1 = 0; lan_r = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lan_r = 2; }
    } else {
        if (anr[0] > anr[2]) { lan_r = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lan_r = 2; }
    }
    V3<R> center;
    {
        const R sb = Sb[lan_r];
        if (nr[lan_r] < 0) {
            for (int k = 0; k < 3; k++) center[k] = pb[k] - pa[k] + sb * Rb(k, lan_r);
        } else {
            for (int k = 0; k < 3; k++) center[k] = pb[k] - pa[k] - sb * Rb(k, lan_r);
        }
    }
    int codeN, code1, code2;
    codeN = code <= 3 ? code - 1 : code - 4;
    if (codeN == 0) { code1 = 1; code2 = 2; }
    else if (codeN == 1) { code1 = 0; code2 = 2; }
    else { code1 = 0; code2 = 1; }
    R quad[8];
    R c1, c2, m11, m12, m21, m22;
    c1 = dot(center, Ra.col(code1));
    c2 = dot(center, Ra.col(code2));
    m11 = dot(Ra.col(code1), Rb.col(a1));
    m12 = dot(Ra.col(code1), Rb.col(a2));
    m21 = dot(Ra.col(code2), Rb.col(a1));
    m22 = dot(Ra.col(code2), Rb.col(a2));
    {
        const R k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
        const R c1mk1 = c1 - k1, c2mk2 = c2 - k2;
        quad[0] = c1mk1 - k3; quad[1] = c2mk2 - k4; quad[2] = c1mk1 + k3; quad[3] = c2mk2 + k4;
        const R c1pk1 = c1 + k1, c2pk2 = c2 + k2;
        quad[4] = c1pk1 + k3; quad[5] = c2pk2 + k4; quad[6] = c1pk1 - k3; quad[7] = c2pk2 - k4;
    }
    R rect[2] = {Sa[code1], Sa[code2]};
    R ret[16];
    int n = intersect_rect_quad2(rect, quad, ret);
    if (n < 1) return false;
    R point[3 * 8];
    R dep[8];
    R det1 = R(1.f) / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    int c_num = 0;
    for (j = 0; j < n; j++) {
        R k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        R k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        for (int k = 0; k < 3; k++) point[c_num * 3 + k] = center[k] + k1 * Rb(k, a1) + k2 * Rb(k, a2);
        dep[c_num] = Sa[codeN] - dot(normal2, V3<R>(point[c_num * 3], point[c_num * 3 + 1], point[c_num * 3 + 2]));
        if (dep[c_num] >= 0) {
            ret[c_num * 2] = ret[j * 2];
            ret[c_num * 2 + 1] = ret[j * 2 + 1];
            c_num++;
        }
    }
    if (c_num < 1) return false;
    max_c = max_c < c_num ? max_c : c_num;
    max_c = max_c > 1 ? max_c : 1;
    const V3<R> normVec = normal;
    const V3<R> marginVec = normVec * R(0);
    if (c_num <= max_c) {
        for (j = 0; j < c_num; j++) {
            const V3<R> pw(point[j * 3 + 0] + pa[0], point[j * 3 + 1] + pa[1], point[j * 3 + 2] + pa[2]);
            if (code < 4) add_contact_point(w, result, -normVec, pw + marginVec, -dep[j] + R(0) * 2);
            else add_contact_point(w, result, normVec, pw - marginVec, -dep[j] + R(0) * 2);
        }
    } else {
        int i1 = 0;
        R max_depth = dep[0];
        for (i = 1; i < c_num; i++) if (dep[i] > max_depth) { max_depth = dep[i]; i1 = i; }
        int i_ret[8];
        cull_points2(c_num, ret, max_c, i1, i_ret);
        for (j = 0; j < max_c; j++) {
            const V3<R> pw(point[i_ret[j] * 3 + 0] + pa[0], point[i_ret[j] * 3 + 1] + pa[1], point[i_ret[j] * 3 + 2] + pa[2]);
            if (code < 4) add_contact_point(w, result, -normVec, pw + marginVec, -dep[i_ret[j]] + R(0) * 2);
            else add_contact_point(w, result, normVec, pw - marginVec, -dep[i_ret[j]] + R(0) * 2);
        }
    }
    return true;
}

// ---------------------------------------------------------------- sphere pairs
// box_sphere_collision_algorithm.cpp:30-80 (note: the +y face point uses half_extent.x())
template <class R> inline R sphere_penetration(const V3<R>& he, const V3<R>& ps, V3<R>& cp, V3<R>& normal) {
    R face_dist = he[0] - ps[0];
    R min_dist = face_dist;
    cp[0] = he[0];
    normal = V3<R>(1, 0, 0);
    face_dist = he[0] + ps[0];
    if (face_dist < min_dist) { min_dist = face_dist; cp = ps; cp[0] = -he[0]; normal = V3<R>(-1, 0, 0); }
    face_dist = he[1] - ps[1];
    if (face_dist < min_dist) { min_dist = face_dist; cp = ps; cp[1] = he[0]; normal = V3<R>(0, 1, 0); }
    face_dist = he[1] + ps[1];
    if (face_dist < min_dist) { min_dist = face_dist; cp = ps; cp[1] = -he[1]; normal = V3<R>(0, -1, 0); }
    face_dist = he[2] - ps[2];
    if (face_dist < min_dist) { min_dist = face_dist; cp = ps; cp[2] = he[2]; normal = V3<R>(0, 0, 1); }
    face_dist = he[2] + ps[2];
    if (face_dist < min_dist) { min_dist = face_dist; cp = ps; cp[2] = -he[2]; normal = V3<R>(0, 0, -1); }
    return min_dist;
}
// box_sphere_collision_algorithm.cpp:6-28, 82-140
template <class R> inline bool box_sphere(const World<R>& w, ContactResult<R>& result,
                                          const Body<R>& ba, const Body<R>& bb) {
    const bool a_is_box = ba.type == ST_BOX;
    const Body<R>& sph = a_is_box ? bb : ba;
    const Body<R>& box = a_is_box ? ba : bb;
    result.swap = a_is_box;
    const V3<R> center = sph.xf.origin;
    const R radius = sph.radius;
    const V3<R> he(box.size[0] / R(2.0), box.size[1] / R(2.0), box.size[2] / R(2.0));
    const V3<R> ps = box.xf.inv_apply(center);
    V3<R> cp = ps;
    for (int k = 0; k < 3; k++) {
        cp[k] = he[k] < cp[k] ? he[k] : cp[k];
        cp[k] = -he[k] > cp[k] ? -he[k] : cp[k];
    }
    V3<R> normal = ps - cp;
    const R dist2 = sqnorm(normal);
    bool ret = false;
    if (!(dist2 > radius * radius)) {
        R distance;
        if (dist2 <= R(0.00001) * R(0.00001)) {
            distance = -sphere_penetration(he, ps, cp, normal);
        } else {
            distance = norm(normal);
            normal = normal / distance;
        }
        normal = mv(box.xf.basis, normal);
        const V3<R> pt_on_box = box.xf.apply(cp);
        const R dist = distance - radius;
        add_contact_point(w, result, normal, pt_on_box - normal * R(0), dist + 2 * R(0));
        ret = true;
    }
    result.swap = false;
    return ret;
}
// sphere_sphere_collision_algorithm.cpp:23-37
template <class R> inline bool sphere_sphere(const World<R>& w, ContactResult<R>& result,
                                             const Body<R>& ba, const Body<R>& bb) {
    const V3<R> rel = ba.xf.origin - bb.xf.origin;
    const R dist = norm(rel);
    if (dist > ba.radius + bb.radius || std::abs(dist - R(0)) < R(0.00001)) return false;
    const V3<R> normal = rel / dist;
    const V3<R> pt_on_b = bb.xf.origin + normal * bb.radius;
    const R depth = dist - ba.radius - bb.radius;
    add_contact_point(w, result, normal, pt_on_b - normal * R(0), depth + 2 * R(0));
    return true;
}

}  // namespace pe_oracle
#include "pe_oracle_convex.hpp"
namespace pe_oracle {

// ---------------------------------------------------------------- stages
// World::updateObjectStatus + applyExternalForce (world.cpp:35-127), the sequential loop's
// semantics (world.cpp:85-117; the threaded loop erases from the array it is iterating, so the body
// after a removed one skips that step and the last is visited twice -- with one removal, of the last
// body of the array, the two agree).  A sleeping body wakes when a velocity reaches its threshold
// scaled by (static + 1) / (static + dynamic + 1) neighbour counts of the last narrowphase, and is not
// moved or damped; an awake body is stepped (RigidBody::step, which ends its lifetime,
// rigidbody.cpp:215-218), damped, and falls asleep (velocities zeroed) after sleep_time seconds below
// the scaled thresholds.  Gravity then acts on every body left, asleep or not.
template <class R> inline void stage_integrate(World<R>& w) {
    for (auto& b : w.bodies) {
        if (b.removed) continue;
        if (b.kinematic) { body_step(b, w.dt); continue; }
        const R ratio = R(b.static_count + 1) / R(b.dynamic_count + b.static_count + 1);
        if (b.sleep) {
            if (sqnorm(b.v) >= w.sleep_lin2 * ratio || sqnorm(b.w) >= w.sleep_ang2 * ratio) {
                b.sleep = false;
                b.sleep_time = 0;
            }
        } else {
            b.last_time += w.dt;
            if (b.last_time >= b.life_time) { b.removed = true; continue; }
            body_step(b, w.dt);
            body_damping(b, w.dt);
            if (sqnorm(b.v) < w.sleep_lin2 * ratio && sqnorm(b.w) < w.sleep_ang2 * ratio) {
                b.sleep_time += w.dt;
                if (b.sleep_time >= w.sleep_time) {
                    b.sleep = true;
                    b.v = V3<R>();
                    b.w = V3<R>();
                }
            } else {
                b.sleep_time = 0;
            }
        }
        b.static_count = 0;
        b.dynamic_count = 0;
    }
    for (auto& b : w.bodies) if (!b.removed) body_gravity(b, w.gravity, w.dt);
}

// BroadPhaseBase::validateCollisionPair (broad_phase_base.cpp:5-19)
template <class R> inline bool validate_pair(const Body<R>& a, const Body<R>& b) {
    if (a.ignore || b.ignore) return false;
    if (a.kinematic && b.kinematic) return false;
    if (a.aabb_min[0] > b.aabb_max[0] || b.aabb_min[0] > a.aabb_max[0]) return false;
    if (a.aabb_min[1] > b.aabb_max[1] || b.aabb_min[1] > a.aabb_max[1]) return false;
    if (a.aabb_min[2] > b.aabb_max[2] || b.aabb_min[2] > a.aabb_max[2]) return false;
    return true;
}
// BroadPhaseSweepAndPrune::calcCollisionPairs (broad_phase_sweep_and_prune.cpp:6-41).
// The reference sorts with std::sort (unspecified order among equal keys); this restatement
// breaks ties by body index, which is the order the device radix sort produces.
// Removed bodies are not in the world's array; the others keep their order.
template <class R> inline void stage_broadphase(World<R>& w) {
    std::vector<int> order;
    for (int i = 0; i < (int)w.bodies.size(); i++) if (!w.bodies[i].removed) order.push_back(i);
    const int n = (int)order.size();
    if (n < 2) return;  // pairs are left untouched, as in the reference
    w.pairs.clear();
    const int ax = w.axis;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return w.bodies[x].aabb_min[ax] < w.bodies[y].aabb_min[ax];
    });
    V3<R> s, s2;
    for (int i = 0; i < n; i++) {
        const Body<R>& b1 = w.bodies[order[i]];
        const V3<R> c = (b1.aabb_min + b1.aabb_max) * R(0.5);
        s = s + c;
        s2 = s2 + V3<R>(c[0] * c[0], c[1] * c[1], c[2] * c[2]);
        for (int j = i + 1; j < n; j++) {
            const Body<R>& b2 = w.bodies[order[j]];
            if (b2.aabb_min[ax] > b1.aabb_max[ax]) break;
            if (validate_pair(b1, b2) &&
                !w.ignore.count(std::make_pair(std::min(order[i], order[j]), std::max(order[i], order[j]))))
                w.pairs.emplace_back(order[i], order[j]);
        }
    }
    const V3<R> v = s2 - V3<R>(s[0] * s[0], s[1] * s[1], s[2] * s[2]) / R(n);
    int axis = 0;
    if (v[1] > v[0]) axis = 1;
    if (v[2] > v[axis]) axis = 2;
    w.axis = axis;
}

// SimpleNarrowPhase::calcContactResults, its second loop (simple_narrow_phase.cpp:43-62): a result with
// points counts, for each non-kinematic body of it, a static neighbour when the other body sleeps or
// is kinematic and a dynamic one otherwise.
template <class R> inline void count_neighbours(World<R>& w) {
    for (size_t i = 0; i < w.pairs.size(); i++) {
        const ContactResult<R>& cr = w.results[i];
        if (cr.npts == 0) continue;
        Body<R>& A = w.bodies[cr.a];
        Body<R>& B = w.bodies[cr.b];
        if (!B.kinematic) { if (A.sleep || A.kinematic) B.static_count++; else B.dynamic_count++; }
        if (!A.kinematic) { if (B.sleep || B.kinematic) A.static_count++; else A.dynamic_count++; }
    }
}
// World::execCollisionCallbacks (world.cpp:129-192): for each result with points whose bodies have
// callbacks, the mean contact position, the normalised sum of the normals, the mean distance and the
// mean relative velocity of the contact points (RigidBody::getLinearVelocityAtLocalPoint,
// rigidbody.cpp:138-140), in pair order.
template <class R> inline void collision_events(World<R>& w) {
    w.events.clear();
    for (size_t i = 0; i < w.pairs.size(); i++) {
        const ContactResult<R>& cr = w.results[i];
        if (cr.npts == 0) continue;
        const Body<R>& A = w.bodies[cr.a];
        const Body<R>& B = w.bodies[cr.b];
        if (!A.report && !B.report) continue;
        CollisionEvent<R> e{cr.a, cr.b, V3<R>(), V3<R>(), V3<R>(), R(0)};
        for (int j = 0; j < cr.npts; j++) {
            const ContactPoint<R>& c = cr.pts[j];
            e.pos = e.pos + c.wpos;
            e.nor = e.nor + c.normal;
            e.depth += c.dist;
            e.vel = e.vel + (A.v + cross(A.w, c.la));
            e.vel = e.vel - (B.v + cross(B.w, c.lb));
        }
        const R n = R(cr.npts);
        e.pos = e.pos / n;
        const R z = sqnorm(e.nor);
        if (z > R(0)) e.nor = e.nor / std::sqrt(z);
        e.depth /= n;
        e.vel = e.vel / n;
        w.events.push_back(e);
    }
}

// NarrowPhaseBase::getAlgorithm (narrow_phase_base.cpp:20-30) and each algorithm's processCollision
// for shapes A and B (bodies, or compound children posed in the world), adding to the result cr whose
// objects are cr.a / cr.b: which shape plays which role, and the swap flag it sets.
template <class R> inline void dispatch_pair(World<R>& w, ContactResult<R>& cr, const Body<R>& A, const Body<R>& B,
                                             R ref_scale) {
    const int ta = A.type, tb = B.type;
    auto is = [&](int x, int y) { return (ta == x && tb == y) || (ta == y && tb == x); };
    auto pick = [&](int t) -> const Body<R>& { return ta == t ? A : B; };
    auto other = [&](int t) -> const Body<R>& { return ta == t ? B : A; };
    auto part = [](const Body<R>& b) { return Part<R>{b.type == ST_CONCAVE ? ST_CONVEX : b.type, &b, b.poly}; };
    if (ta == ST_COMPOUND || tb == ST_COMPOUND) {
        // compound_compound_collision_algorithm.cpp:8-48: a compound B trades places with A (the result's
        // objects too); each child of A, posed by A's transform, against B or each child of B
        const Body<R>* pa = &A;
        const Body<R>* pb = &B;
        if (tb == ST_COMPOUND) {
            std::swap(pa, pb);
            std::swap(cr.a, cr.b);
        }
        for (const Body<R>& ka : pa->kids) {
            Body<R> ca = ka;
            ca.xf = xf_mul(pa->xf, ka.xf);
            if (pb->type == ST_COMPOUND) {
                for (const Body<R>& kb : pb->kids) {
                    Body<R> cb = kb;
                    cb.xf = xf_mul(pb->xf, kb.xf);
                    dispatch_pair(w, cr, ca, cb, ref_scale);
                }
            } else {
                dispatch_pair(w, cr, ca, *pb, ref_scale);
            }
        }
    } else if (ta == ST_BOX && tb == ST_BOX) box_box(w, cr, A, B);
    else if (is(ST_BOX, ST_SPHERE)) box_sphere(w, cr, A, B);
    else if (ta == ST_SPHERE && tb == ST_SPHERE) sphere_sphere(w, cr, A, B);
    else if (is(ST_BOX, ST_CYLINDER)) {            // box_cylinder_collision_algorithm.cpp:13-47
        cr.swap = ta == ST_BOX;
        box_cylinder(w, cr, pick(ST_BOX), pick(ST_CYLINDER));
        cr.swap = false;
    } else if (is(ST_SPHERE, ST_CYLINDER)) {       // sphere_cylinder_collision_algorithm.cpp:8-30
        cr.swap = ta == ST_SPHERE;
        sphere_cylinder(w, cr, pick(ST_SPHERE), pick(ST_CYLINDER));
        cr.swap = false;
    } else if ((ta == ST_CYLINDER && tb == ST_CYLINDER) || is(ST_BOX, ST_CONVEX) || (ta == ST_CONVEX && tb == ST_CONVEX)) {
        // cylinder_cylinder / box_convex / convex_convex: the SAT routine, no swap
        convex_convex(w, cr, part(A), part(B), A.xf, B.xf, ref_scale);
    } else if (is(ST_SPHERE, ST_CONVEX)) {         // sphere_convex_collision_algorithm.cpp:11-34
        cr.swap = ta == ST_CONVEX;
        sphere_mesh(w, cr, pick(ST_SPHERE), pick(ST_CONVEX), *pick(ST_CONVEX).poly);
        cr.swap = false;
    } else if (is(ST_CYLINDER, ST_CONVEX)) {       // cylinder_convex_collision_algorithm.cpp:11-51
        cr.swap = ta == ST_CONVEX;
        cylinder_mesh(w, cr, pick(ST_CONVEX), *pick(ST_CONVEX).poly, pick(ST_CYLINDER));
        cr.swap = false;
    } else if ((ta == ST_CONCAVE) != (tb == ST_CONCAVE)) {
        // concave_{box,sphere,cylinder,convex}_collision_algorithm.cpp: the other body is A
        const Body<R>& mb = pick(ST_CONCAVE);
        const Body<R>& ob = other(ST_CONCAVE);
        cr.swap = ta == ST_CONCAVE;
        if (ob.type == ST_SPHERE) sphere_mesh(w, cr, ob, mb, *mb.poly);
        else if (ob.type == ST_CYLINDER) cylinder_mesh(w, cr, mb, *mb.poly, ob);
        else concave_poly(w, cr, mb, *mb.poly, part(ob), ob, ref_scale);
        cr.swap = false;
    }
    // concave-concave: no algorithm, the result stays cleared
}

// SimpleNarrowPhase::calcContactResults (simple_narrow_phase.cpp:7-41)
template <class R> inline void stage_narrowphase(World<R>& w) {
    const size_t np = w.pairs.size();
    w.results.resize(np);  // new slots are freshly constructed, surviving slots keep stale distances
    for (size_t i = 0; i < np; i++) {
        ContactResult<R>& cr = w.results[i];
        clear_contact_points(cr);
        const int ia = w.pairs[i].first, ib = w.pairs[i].second;
        const Body<R>& A = w.bodies[ia];
        const Body<R>& B = w.bodies[ib];
        cr.a = ia; cr.b = ib;
        cr.mu = std::sqrt(A.mu * B.mu);
        cr.e = std::sqrt(A.e * B.e);
        const R ref_scale = (body_aabb_scale(A) + body_aabb_scale(B)) * R(0.05);
        if (A.type == ST_CONCAVE && B.type == ST_CONCAVE) continue;
        dispatch_pair(w, cr, A, B, ref_scale);
        sort_contact_points(cr);
    }
    count_neighbours(w);
    collision_events(w);
}

// The claim priority of pair p with npts contact points (the device library's rule): the point
// count first, more points lower (earlier), then a fixed bijection of the pair index onto [0, 2^28)
// (odd multiplications modulo 2^28 and right xor-shifts).
inline int color_prio(int p, int npts) {
    const uint32_t m = (1u << 28) - 1u;
    uint32_t x = (uint32_t)p & m;
    x = (x * 0x9E3779B1u) & m;
    x ^= x >> 15;
    x = (x * 0x2C1B3C6Du) & m;
    x ^= x >> 12;
    x = (x * 0x297A2D39u) & m;
    x ^= x >> 15;
    const int k = npts < 1 ? 1 : (npts > 8 ? 8 : npts);
    return (int)((uint32_t)(8 - k) << 28 | x);
}

// Constraint-graph colouring that fixes the solve order (the device library's rule; the
// reference itself sweeps in plain pair order).  Colour c is a maximal independent set built in
// sub-rounds: every remaining candidate claims each of its untaken dynamic bodies with its
// priority color_prio(pair index, point count) (lowest priority wins); a candidate holding all its claims
// joins colour c and takes its bodies; a candidate touching a body taken earlier in this colour
// waits for the next colour.  Kinematic bodies never change (rigidbody.cpp:167), so they impose
// no order.  Pairs of one colour share no dynamic body, so any order inside a colour gives the
// same result; this restatement lists each colour in pair-index order.
template <class R> inline void compute_color_order(World<R>& w) {
    const int np = (int)w.results.size();
    const int nb = (int)w.bodies.size();
    const int FREE = 0x7fffffff;
    w.color_of.assign(np, -1);
    w.solve_order.clear();
    std::vector<int> claim(nb, FREE);
    std::vector<char> taken(nb, 0);
    std::vector<int> active;
    for (int i = 0; i < np; i++) {
        const auto& cr = w.results[i];
        if (cr.npts > 0 && !(w.bodies[cr.a].kinematic && w.bodies[cr.b].kinematic)) active.push_back(i);
    }
    auto dyn = [&](int b) { return !w.bodies[b].kinematic; };
    int color = 0;
    while (!active.empty()) {
        std::vector<int> cand = active, deferred, members;
        while (!cand.empty()) {
            for (int p : cand) {
                const int a = w.results[p].a, b = w.results[p].b;
                if (dyn(a) && !taken[a]) claim[a] = std::min(claim[a], color_prio(p, w.results[p].npts));
                if (dyn(b) && !taken[b]) claim[b] = std::min(claim[b], color_prio(p, w.results[p].npts));
            }
            std::vector<int> next, winners;
            for (int p : cand) {
                const int a = w.results[p].a, b = w.results[p].b;
                const bool blk = (dyn(a) && taken[a]) || (dyn(b) && taken[b]);
                const bool win = !blk && (!dyn(a) || claim[a] == color_prio(p, w.results[p].npts)) && (!dyn(b) || claim[b] == color_prio(p, w.results[p].npts));
                if (blk) deferred.push_back(p);
                else if (win) winners.push_back(p);
                else next.push_back(p);
            }
            for (int p : cand) {
                const int a = w.results[p].a, b = w.results[p].b;
                if (dyn(a) && !taken[a]) claim[a] = FREE;
                if (dyn(b) && !taken[b]) claim[b] = FREE;
            }
            for (int p : winners) {
                const int a = w.results[p].a, b = w.results[p].b;
                if (dyn(a)) taken[a] = 1;
                if (dyn(b)) taken[b] = 1;
                members.push_back(p);
            }
            cand.swap(next);
        }
        std::sort(members.begin(), members.end());
        for (int p : members) {
            w.color_of[p] = color;
            w.solve_order.push_back(p);
            taken[w.results[p].a] = 0;
            taken[w.results[p].b] = 0;
        }
        std::sort(deferred.begin(), deferred.end());
        active.swap(deferred);
        color++;
    }
    w.ncolors = color;
}

// Solver rows (friction_contact_constraint.h:15-29)
template <class R> struct Row {
    V3<R> ra, rb, n, t0, t1;
    R rhs = 0, mn = 0, mt0 = 0, mt1 = 0, ln = 0, lt0 = 0, lt1 = 0, mu = 0;
};
template <class R> struct Solver {
    std::vector<std::vector<Row<R>>> rows;  // per result
};

// SequentialImpulseSolver::setupSolver + FrictionContactConstraint::initSequentialImpulse
// (sequential_impulse_solver.cpp:11-58, friction_contact_constraint.cpp:9-73)
template <class R> inline void stage_solver_setup(World<R>& w, Solver<R>& s) {
    for (auto& b : w.bodies) {
        if (b.kinematic) continue;  // clearTempVelocity is a no-op for kinematic bodies
        b.tv = b.v; b.tw = b.w;
    }
    const R kerp = R(0.2), pen_th = R(0.3), rest_th = R(0.1);
    s.rows.resize(w.results.size());
    for (size_t i = 0; i < w.results.size(); i++) {
        const ContactResult<R>& cr = w.results[i];
        const int np = cr.npts < 8 ? cr.npts : 8;
        auto& rows = s.rows[i];
        rows.resize(np);
        const Body<R>& A = w.bodies[cr.a];
        const Body<R>& B = w.bodies[cr.b];
        const R ims = (A.kinematic ? R(0) : A.inv_mass) + (B.kinematic ? R(0) : B.inv_mass);
        const M3<R> Ia = A.kinematic ? M3<R>() : A.world_inv_inertia;
        const M3<R> Ib = B.kinematic ? M3<R>() : B.world_inv_inertia;
        for (int k = 0; k < np; k++) {
            const ContactPoint<R>& cp = cr.pts[k];
            Row<R>& ci = rows[k];
            const V3<R> ra = mv(A.xf.basis, cp.la);
            const V3<R> rb = mv(B.xf.basis, cp.lb);
            ci.ra = ra; ci.rb = rb; ci.n = cp.normal; ci.t0 = cp.t0; ci.t1 = cp.t1;
            auto jmj = [&](const V3<R>& d) {
                const V3<R> xa = cross(ra, d), xb = cross(rb, d);
                return R(1.0) / (ims + dot(mv(Ia, xa), xa) + dot(mv(Ib, xb), xb));
            };
            ci.mn = jmj(ci.n);
            ci.mt0 = jmj(ci.t0);
            ci.mt1 = jmj(ci.t1);
            const V3<R> va = A.v + cross(A.w, ra);
            const V3<R> vb = B.v + cross(B.w, rb);
            R rev = -dot(ci.n, va - vb);
            if (std::abs(rev) < rest_th) rev = 0;
            rev *= cr.e;
            R pen = cp.dist;
            pen = pen > -pen_th ? pen : -pen_th;
            ci.rhs = (rev - kerp * pen / w.dt) * ci.mn;
            ci.ln = ci.lt0 = ci.lt1 = 0;
            ci.mu = cr.mu;
        }
    }
    for (auto& j : w.joints) joint_setup(w, j);
    compute_joint_order(w);
}
// RigidBody::applyTempImpulse (rigidbody.cpp:166-170)
template <class R> inline void apply_temp_impulse(Body<R>& b, const V3<R>& r, const V3<R>& J) {
    if (b.kinematic) return;
    b.tv = b.tv + J * b.inv_mass;
    b.tw = b.tw + mv(b.world_inv_inertia, cross(r, J));
}
// FrictionContactConstraint::iterateSequentialImpulse (friction_contact_constraint.cpp:75-108)
template <class R> inline void iterate_rows(World<R>& w, const ContactResult<R>& cr, std::vector<Row<R>>& rows) {
    Body<R>& A = w.bodies[cr.a];
    Body<R>& B = w.bodies[cr.b];
    for (auto& ci : rows) {
        const V3<R> vr = ((A.tv + cross(A.tw, ci.ra)) - B.tv) - cross(B.tw, ci.rb);
        R dn = ci.rhs - dot(ci.n, vr) * ci.mn;
        dn = dn > -ci.ln ? dn : -ci.ln;
        ci.ln += dn;
        R t0 = ci.lt0 - dot(ci.t0, vr) * ci.mt0;
        R t1 = ci.lt1 - dot(ci.t1, vr) * ci.mt1;
        const R scale = ci.mu * ci.ln / std::sqrt(t0 * t0 + t1 * t1);
        if (scale < 1) { t0 *= scale; t1 *= scale; }
        V3<R> J;
        for (int k = 0; k < 3; k++) J[k] = (dn * ci.n[k] + (t0 - ci.lt0) * ci.t0[k]) + (t1 - ci.lt1) * ci.t1[k];
        ci.lt0 = t0; ci.lt1 = t1;
        apply_temp_impulse(A, ci.ra, J);
        apply_temp_impulse(B, ci.rb, -J);
    }
}
// SequentialImpulseSolver::solve (sequential_impulse_solver.cpp:60-81).  The sweep runs over
// the colour order when use_color_order is set, otherwise over plain pair order (the reference).
template <class R> inline void stage_solve(World<R>& w, Solver<R>& s) {
    for (int it = 0; it < w.iterations; it++) {
        if (w.use_color_order) {
            for (int i : w.solve_order) iterate_rows(w, w.results[i], s.rows[i]);
        } else {
            for (size_t i = 0; i < w.results.size(); i++) iterate_rows(w, w.results[i], s.rows[i]);
        }
        for (int k : w.joint_order) joint_iterate(w, w.joints[k]);
    }
    for (auto& b : w.bodies) {
        if (b.kinematic) continue;
        b.v = b.tv; b.w = b.tw;
    }
}

// ---------------------------------------------------------------- joint constraints
// BallJointConstraint, HingeJointConstraint, SliderJointConstraint
// (constraint/{ball,hinge,slider}_joint_constraint.cpp): set up in setupSolver after the contacts,
// iterated after the contacts in every Gauss-Seidel sweep (sequential_impulse_solver.cpp:49-51, 66-68).
enum { JT_BALL = 0, JT_HINGE = 1, JT_SLIDER = 2 };
template <class R> struct Joint {
    int type = JT_BALL, a = 0, b = 0;
    V3<R> anchor_a, anchor_b, axis_a{1, 0, 0}, axis_b{1, 0, 0};
    bool use_limits = false, use_motor = false;
    R min_angle = 0, max_angle = 0, target_speed = 0;
    // initSequentialImpulse state
    V3<R> ra, rb, wax_a, wax_b, wt_a[2], wt_b[2];
    M3<R> jmj_ball, jmj_rot;
    V3<R> rhs_ball, rhs_rot;
    R rhs_hinge[2] = {0, 0}, jmj_hinge[2] = {0, 0}, rhs_pos[2] = {0, 0}, jmj_pos[2] = {0, 0};
    R rhs_motor = 0, rhs_limit = 0, jmj_ml = 0;
    bool limit_exceeded = false;
};

// Constraint::getSkewSymmetricMatrix (constraint.cpp:9-13)
template <class R> inline M3<R> skew(const V3<R>& v) {
    M3<R> m;
    m(0, 0) = 0; m(0, 1) = -v[2]; m(0, 2) = v[1];
    m(1, 0) = v[2]; m(1, 1) = 0; m(1, 2) = -v[0];
    m(2, 0) = -v[1]; m(2, 1) = v[0]; m(2, 2) = 0;
    return m;
}
// rx * I * rx^T with the second product summed as in (R * L) * R^T (body_update_world_inertia)
template <class R> inline M3<R> sandwich(const M3<R>& rx, const M3<R>& I) {
    const M3<R> T = mm(rx, I);
    M3<R> r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r(i, j) = T(i, 0) * rx(j, 0) + (T(i, 1) * rx(j, 1) + T(i, 2) * rx(j, 2));
    return r;
}
template <class R> inline M3<R> madd(const M3<R>& A, const M3<R>& B) {
    M3<R> r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r(i, j) = A(i, j) + B(i, j);
    return r;
}
// (I * (ima + imb) + rxA Ia rxA^T + rxB Ib rxB^T)^-1, the ball part's effective mass
template <class R> inline M3<R> ball_jmj(R ims, const V3<R>& ra, const M3<R>& Ia, const V3<R>& rb, const M3<R>& Ib) {
    M3<R> S;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) S(i, j) = (i == j ? R(1) : R(0)) * ims;
    return inverse3(madd(madd(S, sandwich(skew(ra), Ia)), sandwich(skew(rb), Ib)));
}
template <class R> inline R kin_inv_mass(const Body<R>& b) { return b.kinematic ? R(0) : b.inv_mass; }
template <class R> inline M3<R> kin_inv_inertia(const Body<R>& b) { return b.kinematic ? M3<R>() : b.world_inv_inertia; }
// RigidBody::applyTempAngularImpulse (rigidbody.cpp:178-181)
template <class R> inline void apply_temp_angular(Body<R>& b, const V3<R>& J) {
    if (b.kinematic) return;
    b.tw = b.tw + mv(b.world_inv_inertia, J);
}

template <class R> inline void joint_setup(World<R>& w, Joint<R>& j) {
    const Body<R>& A = w.bodies[j.a];
    const Body<R>& B = w.bodies[j.b];
    const R kerp = R(0.2), dt = w.dt;
    j.ra = mv(A.xf.basis, j.anchor_a);
    j.rb = mv(B.xf.basis, j.anchor_b);
    const R ims = kin_inv_mass(A) + kin_inv_mass(B);
    const M3<R> Ia = kin_inv_inertia(A), Ib = kin_inv_inertia(B);
    const V3<R> rel = A.xf.apply(j.anchor_a) - B.xf.apply(j.anchor_b);
    if (j.type == JT_BALL || j.type == JT_HINGE) {   // ball_joint_constraint.cpp:6-19, hinge_joint_constraint.cpp:25-34
        j.jmj_ball = ball_jmj(ims, j.ra, Ia, j.rb, Ib);
        j.rhs_ball = mv(j.jmj_ball, rel) * (-kerp / dt);
    }
    if (j.type == JT_HINGE || j.type == JT_SLIDER) {
        j.wax_a = mv(A.xf.basis, j.axis_a);
        j.wax_b = mv(B.xf.basis, j.axis_b);
        ortho_units(j.axis_a, j.wt_a[0], j.wt_a[1]);
        j.wt_a[0] = mv(A.xf.basis, j.wt_a[0]);
        j.wt_a[1] = mv(A.xf.basis, j.wt_a[1]);
        if (j.type == JT_SLIDER || j.use_limits) {
            ortho_units(j.axis_b, j.wt_b[0], j.wt_b[1]);
            j.wt_b[0] = mv(B.xf.basis, j.wt_b[0]);
            j.wt_b[1] = mv(B.xf.basis, j.wt_b[1]);
        }
    }
    if (j.type == JT_HINGE) {   // hinge_joint_constraint.cpp:36-70
        const M3<R> Is = madd(Ia, Ib);
        j.jmj_hinge[0] = R(1.0) / dot(j.wt_a[0], mv(Is, j.wt_a[0]));
        j.jmj_hinge[1] = R(1.0) / dot(j.wt_a[1], mv(Is, j.wt_a[1]));
        const V3<R> u = cross(j.wax_a, j.wax_b);
        j.rhs_hinge[0] = j.jmj_hinge[0] * (dot(u, j.wt_a[0]) * kerp / dt);
        j.rhs_hinge[1] = j.jmj_hinge[1] * (dot(u, j.wt_a[1]) * kerp / dt);
        if (j.use_motor || j.use_limits) j.jmj_ml = R(1.0) / dot(j.wax_a, mv(Is, j.wax_a));
        if (j.use_motor) j.rhs_motor = j.jmj_ml * -j.target_speed;
        if (j.use_limits) {
            const R ac = dot(j.wt_a[0], j.wt_b[0]);
            const R as = dot(cross(j.wt_a[0], j.wt_b[0]), j.wax_a);
            const R angle = std::atan2(as, ac);
            j.limit_exceeded = false;
            if (angle < j.min_angle) {
                j.limit_exceeded = true;
                j.rhs_limit = -j.jmj_ml * (j.min_angle - angle) * kerp / dt;
            } else if (angle > j.max_angle) {
                j.limit_exceeded = true;
                j.rhs_limit = -j.jmj_ml * (j.max_angle - angle) * kerp / dt;
            }
        }
    }
    if (j.type == JT_SLIDER) {   // slider_joint_constraint.cpp:24-53
        auto jmj_dir = [&](const V3<R>& d) {
            const V3<R> xa = cross(j.ra, d), xb = cross(j.rb, d);
            return R(1.0) / (ims + dot(mv(Ia, xa), xa) + dot(mv(Ib, xb), xb));
        };
        j.jmj_pos[0] = jmj_dir(j.wt_a[0]);
        j.jmj_pos[1] = jmj_dir(j.wt_a[1]);
        // the reference scales the second row by the first row's effective mass (:45)
        j.rhs_pos[0] = j.jmj_pos[0] * (dot(rel, j.wt_a[0]) * (-kerp / dt));
        j.rhs_pos[1] = j.jmj_pos[0] * (dot(rel, j.wt_a[1]) * (-kerp / dt));
        j.jmj_rot = inverse3(madd(Ia, Ib));
        j.rhs_rot = mv(j.jmj_rot, cross(j.wt_a[0], j.wt_b[0]) + cross(j.wt_a[1], j.wt_b[1])) * (kerp / dt);
    }
}

template <class R> inline void joint_iterate(World<R>& w, Joint<R>& j) {
    Body<R>& A = w.bodies[j.a];
    Body<R>& B = w.bodies[j.b];
    if (j.type == JT_BALL || j.type == JT_HINGE) {
        const V3<R> va = A.tv + cross(A.tw, j.ra);
        const V3<R> vb = B.tv + cross(B.tw, j.rb);
        const V3<R> imp = j.rhs_ball - mv(j.jmj_ball, va - vb);
        apply_temp_impulse(A, j.ra, imp);
        apply_temp_impulse(B, j.rb, -imp);
    }
    if (j.type == JT_HINGE) {
        // the angular velocities are read live: each impulse sees the ones applied before it
        const V3<R> dw = A.tw - B.tw;
        const R h0 = j.rhs_hinge[0] - j.jmj_hinge[0] * dot(j.wt_a[0], dw);
        const R h1 = j.rhs_hinge[1] - j.jmj_hinge[1] * dot(j.wt_a[1], dw);
        const V3<R> iv = j.wt_a[0] * h0 + j.wt_a[1] * h1;
        apply_temp_angular(A, iv);
        apply_temp_angular(B, -iv);
        if (j.use_motor) {
            const R m = j.rhs_motor - j.jmj_ml * dot(j.wax_a, A.tw - B.tw);
            apply_temp_angular(A, j.wax_a * m);
            apply_temp_angular(B, -j.wax_a * m);
        }
        if (j.use_limits && j.limit_exceeded) {
            const R delta = dot(j.wax_a, A.tw - B.tw);
            const R nd = delta * j.rhs_limit > 0 ? R(0) : delta;
            const R l = j.rhs_limit - j.jmj_ml * nd;
            apply_temp_angular(A, j.wax_a * l);
            apply_temp_angular(B, -j.wax_a * l);
        }
    }
    if (j.type == JT_SLIDER) {
        const V3<R> va = A.tv + cross(A.tw, j.ra);
        const V3<R> vb = B.tv + cross(B.tw, j.rb);
        const R p0 = j.rhs_pos[0] - j.jmj_pos[0] * dot(j.wt_a[0], va - vb);
        const R p1 = j.rhs_pos[1] - j.jmj_pos[1] * dot(j.wt_a[1], va - vb);
        const V3<R> pv = j.wt_a[0] * p0 + j.wt_a[1] * p1;
        apply_temp_impulse(A, j.ra, pv);
        apply_temp_impulse(B, j.rb, -pv);
        const V3<R> rot = -mv(j.jmj_rot, A.tw - B.tw);
        apply_temp_angular(A, rot);
        apply_temp_angular(B, -rot);
    }
}

// Joint sweep order: with use_color_order, joints grouped by a greedy colouring in joint order
// (joint k takes the lowest colour no earlier joint sharing one of its dynamic bodies holds); the
// device library colours its joints by the same rule.  Joints of one colour share no dynamic body.
template <class R> inline void compute_joint_order(World<R>& w) {
    const int nj = (int)w.joints.size();
    w.joint_order.resize(nj);
    if (!w.use_color_order) {
        for (int k = 0; k < nj; k++) w.joint_order[k] = k;
        return;
    }
    std::vector<int> col(nj, 0);
    int ncol = 0;
    for (int k = 0; k < nj; k++) {
        std::set<int> used;
        for (int q = 0; q < k; q++) {
            const auto dyn_shared = [&](int x) { return !w.bodies[x].kinematic && (x == w.joints[q].a || x == w.joints[q].b); };
            if (dyn_shared(w.joints[k].a) || dyn_shared(w.joints[k].b)) used.insert(col[q]);
        }
        int c = 0;
        while (used.count(c)) c++;
        col[k] = c;
        ncol = std::max(ncol, c + 1);
    }
    int o = 0;
    for (int c = 0; c < ncol; c++)
        for (int k = 0; k < nj; k++)
            if (col[k] == c) w.joint_order[o++] = k;
}

// World::step (world.cpp:267-300)
template <class R> inline void world_step(World<R>& w, Solver<R>& s) {
    stage_integrate(w);
    stage_broadphase(w);
    stage_narrowphase(w);
    stage_solver_setup(w, s);
    if (w.use_color_order) compute_color_order(w);
    stage_solve(w, s);
}

// ---------------------------------------------------------------- raycast (rigid/raycast)
// RaycastResult (raycast_solver_base.h:7-13) as DefaultRaycastSolver::performRaycast leaves it,
// with its return value.
template <class R> struct RayHit {
    bool hit = false;             // performRaycast's return value
    int body = -1;                // m_collisionObject
    R distance = Lim<R>::max();
    V3<R> point, normal;
};
// RaycastBox::rayHitBox (raycast_box.cpp:31-60): slab test with the direction's tiny components
// raised to PE_EPS; the normal is the entering slab's, x before y before z on ties.
template <class R> inline bool ray_hit_box(const V3<R>& start, const V3<R>& direction, const V3<R>& bmin,
                                           const V3<R>& bmax, R& distance, V3<R>& hp, V3<R>& hn) {
    const R eps = R(0.00001);
    V3<R> dir;
    for (int k = 0; k < 3; k++) dir[k] = std::abs(direction[k]) < eps ? eps : direction[k];
    V3<R> m, n, tmax, tmin;
    for (int k = 0; k < 3; k++) {
        m[k] = (bmax[k] - start[k]) / dir[k];
        n[k] = (bmin[k] - start[k]) / dir[k];
        tmax[k] = m[k] > n[k] ? m[k] : n[k];
        tmin[k] = m[k] < n[k] ? m[k] : n[k];
    }
    R t_enter = tmin[0] > tmin[1] ? tmin[0] : tmin[1];
    t_enter = t_enter > tmin[2] ? t_enter : tmin[2];
    R t_exit = tmax[0] < tmax[1] ? tmax[0] : tmax[1];
    t_exit = t_exit < tmax[2] ? t_exit : tmax[2];
    if (t_enter > t_exit || t_exit < 0) return false;
    distance = t_enter;
    hp = start + dir * t_enter;
    hn = V3<R>();
    if (t_enter == tmin[0]) hn[0] = dir[0] > 0 ? R(-1.0) : R(1.0);
    else if (t_enter == tmin[1]) hn[1] = dir[1] > 0 ? R(-1.0) : R(1.0);
    else hn[2] = dir[2] > 0 ? R(-1.0) : R(1.0);
    return true;
}
// RaycastBox::processRaycast (raycast_box.cpp:6-29)
template <class R> inline bool ray_box(const V3<R>& start, const V3<R>& dir, R maxd, const Body<R>& b, R& d,
                                       V3<R>& hp, V3<R>& hn) {
    const V3<R> sl = b.xf.inv_apply(start), dl = mtv(b.xf.basis, dir);
    const V3<R> sz(b.size[0], b.size[1], b.size[2]);
    if (ray_hit_box(sl, dl, (-sz) / R(2.0), sz / R(2.0), d, hp, hn)) {
        if (d <= maxd) {
            hp = b.xf.apply(hp);
            hn = mv(b.xf.basis, hn);
            return true;
        }
        return false;
    }
    hp = hn = V3<R>();
    d = Lim<R>::max();
    return false;
}
// RaycastSphere::processRaycast (raycast_sphere.cpp:6-29); a sphere behind the start, or around it,
// is hit at a negative distance.
template <class R> inline bool ray_sphere(const V3<R>& start, const V3<R>& dir, R maxd, const Body<R>& b, R& d,
                                          V3<R>& hp, V3<R>& hn) {
    const R r = b.radius;
    const V3<R> oc = b.xf.origin - start;
    const R c2s = norm(oc);
    const R s2p = dot(oc, dir);
    const R q = c2s * c2s - s2p * s2p;
    const R p2c = std::sqrt(q > R(0) ? q : R(0));
    if (p2c > r) {
        hp = hn = V3<R>();
        d = Lim<R>::max();
        return false;
    }
    const R p2h = std::sqrt(r * r - p2c * p2c);
    d = s2p - p2h;
    if (d <= maxd) {
        hp = start + dir * d;
        hn = normalized(hp - b.xf.origin);
        return true;
    }
    return false;
}
// RaycastCylinder::processRaycast (raycast_cylinder.cpp:6-77), its branches and its cap points
// (s + (h d - d * s) / d.y, with the top cap's y then set to h).
template <class R> inline bool ray_cylinder(const V3<R>& start, const V3<R>& dir, R maxd, const Body<R>& b, R& d,
                                            V3<R>& hp, V3<R>& hn) {
    const R radius = b.radius, height = b.height / R(2.0);
    const V3<R> sl = b.xf.inv_apply(start), dl = mtv(b.xf.basis, dir);
    const R eps = R(0.00001);
    auto cap = [&](R h) {
        V3<R> p;
        for (int k = 0; k < 3; k++) p[k] = sl[k] + (h * dl[k] - dl[k] * sl[k]) / dl[1];
        return p;
    };
    bool hit = false;
    if (std::abs(dl[0] - R(0)) < eps && std::abs(dl[2] - R(0)) < eps) {
        const R d2a = std::sqrt(sl[0] * sl[0] + sl[2] * sl[2]);
        if (d2a <= radius) {
            const R h = dl[1] < 0 ? height : -height;
            hp = cap(h);
            hn = V3<R>(0, 1, 0);
            if (dl[1] > 0) hn = -hn;
            d = norm(hp - sl);
            hit = true;
        }
    } else {
        const R d2a = std::abs(dot(sl, normalized(cross(dl, V3<R>(0, 1, 0)))));
        if (d2a <= radius) {
            const R den = dl[0] * dl[0] + dl[2] * dl[2];
            const R t = -(dl[0] * sl[0] + dl[2] * sl[2]) / den;
            const R dt = std::sqrt(radius * radius - d2a * d2a) / std::sqrt(den);
            const V3<R> in = sl + dl * (t - dt);
            const V3<R> out = sl + dl * (t + dt);
            if ((in[1] < -height && out[1] < -height) || (in[1] > height && out[1] > height)) {
                hit = false;
            } else if (in[1] >= -height && in[1] <= height) {
                hp = in;
                hn = normalized(V3<R>(in[0], R(0), in[2]));
                d = t - dt;
                hit = true;
            } else if (in[1] > height) {
                hp = cap(height);
                hp[1] = height;
                hn = V3<R>(0, 1, 0);
                d = norm(hp - sl);
                hit = true;
            } else {
                hp = cap(-height);
                hn = -V3<R>(0, 1, 0);
                d = norm(hp - sl);
                hit = true;
            }
        }
    }
    if (!hit) {
        hp = hn = V3<R>();
        d = Lim<R>::max();
        return false;
    }
    if (d <= maxd) {
        hp = b.xf.apply(hp);
        hn = mv(b.xf.basis, hn);
        return true;
    }
    return false;
}
// RaycastMesh::rayHitTriangle (raycast_mesh.cpp:52-77), Moller-Trumbore; distances behind the start
// are not rejected.
template <class R> inline bool ray_triangle(const V3<R>& s, const V3<R>& dir, const V3<R>& v0, const V3<R>& v1,
                                            const V3<R>& v2, R& d, V3<R>& hit) {
    const R eps = R(0.00001);
    const V3<R> e1 = v1 - v0, e2 = v2 - v0;
    const V3<R> p = cross(dir, e2);
    const R det = dot(e1, p);
    if (det > -eps && det < eps) return false;
    const R inv = R(1.0) / det;
    const V3<R> t = s - v0;
    const R u = dot(t, p) * inv;
    if (u < 0 || u > R(1.0)) return false;
    const V3<R> q = cross(t, e1);
    const R v = dot(dir, q) * inv;
    if (v < 0 || u + v > R(1.0)) return false;
    d = dot(e2, q) * inv;
    hit = s + dir * d;
    return true;
}
// RaycastMesh::processRaycast (raycast_mesh.cpp:9-50): the faces the BVH finds in the ray's box
// (start to start + length, grown by 0.1), each fanned from its first vertex, the first fan triangle
// hit counting for the face; the nearest face wins (strictly nearer replaces).
template <class R> inline bool ray_mesh(const V3<R>& start, const V3<R>& dir, R maxd, const Body<R>& b, R& d,
                                        V3<R>& hp, V3<R>& hn) {
    const Poly<R>& P = *b.poly;
    const V3<R> sl = b.xf.inv_apply(start), dl = mtv(b.xf.basis, dir);
    const V3<R> el = sl + dl * maxd;
    V3<R> bmin, bmax;
    for (int k = 0; k < 3; k++) {
        bmin[k] = (sl[k] < el[k] ? sl[k] : el[k]) - R(0.1);
        bmax[k] = (sl[k] > el[k] ? sl[k] : el[k]) + R(0.1);
    }
    std::vector<int> faces;
    bvh_search(P, bmin, bmax, faces);
    d = Lim<R>::max();
    hp = hn = V3<R>();
    for (int fi : faces) {
        const auto& f = P.f[fi];
        for (int i = 0; i < (int)f.size() - 2; i++) {
            R t;
            V3<R> hit;
            if (ray_triangle(sl, dl, P.v[f[0]], P.v[f[i + 1]], P.v[f[i + 2]], t, hit)) {
                if (t < d) {
                    d = t;
                    hp = hit;
                    hn = P.fn[fi];
                }
                break;
            }
        }
    }
    if (d < maxd) {
        hp = b.xf.apply(hp);
        hn = mv(b.xf.basis, hn);
        return true;
    }
    return false;
}
// DefaultRaycastSolver::performRaycast (default_raycast_solver.cpp:11-86) over the world's bodies in
// order: ignore-collision bodies and the call's ignore set are skipped; a mesh body whose AABB the
// (unbounded) ray misses ends the whole query with no hit, the result keeping what earlier bodies
// put there; otherwise the strictly nearest hit within length wins.
template <class R> inline RayHit<R> raycast(const World<R>& w, const V3<R>& start, const V3<R>& dir, R length,
                                            const std::set<int>& ignore) {
    RayHit<R> res;
    for (int i = 0; i < (int)w.bodies.size(); i++) {
        const Body<R>& b = w.bodies[i];
        if (b.removed || b.ignore || ignore.count(i)) continue;
        R d = 0;
        V3<R> hp, hn;
        bool ret = false;
        if (b.type == ST_BOX) ret = ray_box(start, dir, length, b, d, hp, hn);
        else if (b.type == ST_SPHERE) ret = ray_sphere(start, dir, length, b, d, hp, hn);
        else if (b.type == ST_CYLINDER) ret = ray_cylinder(start, dir, length, b, d, hp, hn);
        else {
            if (!ray_hit_box(start, dir, b.aabb_min, b.aabb_max, d, hp, hn)) {
                res.hit = false;
                return res;
            }
            if (b.type != ST_COMPOUND) {
                ret = ray_mesh(start, dir, length, b, d, hp, hn);
            } else {
                // each child in turn writes the same outputs, its hit or-ed into ret: the last child's
                // distance and point decide (default_raycast_solver.cpp:50-72)
                for (const Body<R>& kid : b.kids) {
                    Body<R> t = kid;
                    t.xf = xf_mul(b.xf, kid.xf);
                    if (kid.type == ST_BOX) ret |= ray_box(start, dir, length, t, d, hp, hn);
                    else if (kid.type == ST_SPHERE) ret |= ray_sphere(start, dir, length, t, d, hp, hn);
                    else if (kid.type == ST_CYLINDER) ret |= ray_cylinder(start, dir, length, t, d, hp, hn);
                    else if (kid.type == ST_CONVEX) ret |= ray_mesh(start, dir, length, t, d, hp, hn);
                }
            }
        }
        if (ret && d <= length && d < res.distance) {
            res.distance = d;
            res.point = hp;
            res.normal = hn;
            res.body = i;
        }
    }
    res.hit = res.body >= 0;
    return res;
}

// Parity table (what pins each function; the tests named here live in tests/):
//   quat / mat helpers ........ pinned: bit-exact vs compiled reference (test_oracle_vs_ref)
//   body_step / aabb / gravity  pinned: bit-exact vs compiled reference; closed-form KATs
//   stage_broadphase .......... pinned: pair set vs compiled reference and brute force
//   box_box / box_sphere / sphere_sphere + cache ... pinned: vs compiled reference, per pair
//   stage_solver_setup / solve  pinned: vs compiled reference in pair order (use_color_order=0)
//   compute_color_order ....... pinned: invariants (no dynamic body shared within a colour; every
//                               colour maximal) and brute-force check on small graphs (tests/)
//   pe_oracle_convex.hpp ...... pinned: box / cylinder meshes, convex and concave setMesh (volume,
//                               inertia, centring, BVH face order, unique edges) equal to the compiled
//                               reference's shapes; every shape pair type (box, sphere, cylinder,
//                               convex hull, heightfield) bit-exact per stage vs the compiled reference
//                               in FP32 and FP64 (test_oracle_vs_ref), and vs golden fixtures
//                               (test_golden)
//   joints / ignore pairs ..... pinned: bit-exact vs compiled reference (constraints demo, chains)
//   sleep / lifetime / removal / neighbour counts / collision reports ... pinned: bit-exact vs the
//                               compiled reference on one pool thread
//   raycast ................... pinned: bit-exact vs the compiled reference's DefaultRaycastSolver
//                               on every shape kind, and closed forms (sphere, box) on tiny cases

}  // namespace pe_oracle
