// Device-side building blocks of the rigid-body step (gfx950, FP32).
//
// Every function here restates a reference routine (cited per function) with the same
// operation order as the reference built in FP32 with Eigen 3.4: a 3-term sum -- dot
// product, matrix-vector row, trace -- is evaluated p0 + (p1 + p2).  The library is built
// with -ffp-contract=off and correctly rounded division and square root, so these
// functions round exactly like the reference's host build.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PE_HD __device__ __forceinline__

namespace pe {

constexpr float REAL_MAX = 1e30f;    // PE_REAL_MAX in FP32 (phys_general.h:36-42)
constexpr float REAL_MIN = -1e30f;
constexpr float EPS = 0.00001f;      // PE_EPS
constexpr float PI_R = 3.141592653589f;

enum : int { SHAPE_BOX = 0, SHAPE_SPHERE = 1, SHAPE_CYLINDER = 2, SHAPE_CONVEX = 3, SHAPE_CONCAVE = 4, SHAPE_COMPOUND = 5 };

struct f3 { float x, y, z; };
PE_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
PE_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PE_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PE_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PE_HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PE_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
PE_HD float sum3(float a, float b, float c) { return a + (b + c); }
PE_HD float dot(f3 a, f3 b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }
PE_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PE_HD float norm(f3 a) { return sqrtf(dot(a, a)); }
PE_HD f3 normalized(f3 a) {
    float n = dot(a, a);
    if (n > 0.0f) { float s = sqrtf(n); return f3{a.x / s, a.y / s, a.z / s}; }
    return a;
}
PE_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
PE_HD void setc(f3& a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }

// 3x3 matrix, row-major m[r*3+c]
struct m33 { float m[9]; };
PE_HD f3 col(const m33& A, int j) {   // by selects, so a run-time j keeps A in registers
    return j == 0 ? f3{A.m[0], A.m[3], A.m[6]} : (j == 1 ? f3{A.m[1], A.m[4], A.m[7]} : f3{A.m[2], A.m[5], A.m[8]});
}
PE_HD f3 mv(const m33& A, f3 v) {     // A * v
    return f3{sum3(A.m[0] * v.x, A.m[1] * v.y, A.m[2] * v.z),
              sum3(A.m[3] * v.x, A.m[4] * v.y, A.m[5] * v.z),
              sum3(A.m[6] * v.x, A.m[7] * v.y, A.m[8] * v.z)};
}
PE_HD f3 mtv(const m33& A, f3 v) {    // A^T * v
    return f3{sum3(A.m[0] * v.x, A.m[3] * v.y, A.m[6] * v.z),
              sum3(A.m[1] * v.x, A.m[4] * v.y, A.m[7] * v.z),
              sum3(A.m[2] * v.x, A.m[5] * v.y, A.m[8] * v.z)};
}
PE_HD m33 mm(const m33& A, const m33& B) {
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.m[i * 3 + j] = sum3(A.m[i * 3] * B.m[j], A.m[i * 3 + 1] * B.m[3 + j], A.m[i * 3 + 2] * B.m[6 + j]);
    return r;
}
PE_HD m33 mtm(const m33& A, const m33& B) {   // A^T * B
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.m[i * 3 + j] = sum3(A.m[i] * B.m[j], A.m[3 + i] * B.m[3 + j], A.m[6 + i] * B.m[6 + j]);
    return r;
}
PE_HD m33 mmt(const m33& A, const m33& B) {   // A * B^T
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.m[i * 3 + j] = sum3(A.m[i * 3] * B.m[j * 3], A.m[i * 3 + 1] * B.m[j * 3 + 1], A.m[i * 3 + 2] * B.m[j * 3 + 2]);
    return r;
}

// ------------------------------------------------------------------ body step
// Eigen quaternionbase_assign_impl<Other,3,3> (Shoemake), as used by
// RigidBody::step (rigidbody.cpp:222): quaternion (x, y, z, w) from a rotation matrix.
PE_HD void quat_from_matrix(const m33& m, float q[4]) {
    float t = m.m[0] + (m.m[4] + m.m[8]);
    if (t > 0.0f) {
        t = sqrtf(t + 1.0f);
        q[3] = 0.5f * t;
        t = 0.5f / t;
        q[0] = (m.m[7] - m.m[5]) * t;
        q[1] = (m.m[2] - m.m[6]) * t;
        q[2] = (m.m[3] - m.m[1]) * t;
    } else {
        int i = 0;
        if (m.m[4] > m.m[0]) i = 1;
        if (m.m[8] > m.m[i * 4]) i = 2;
        int j = (i + 1) % 3, k = (j + 1) % 3;
        t = sqrtf(m.m[i * 4] - m.m[j * 4] - m.m[k * 4] + 1.0f);
        q[i] = 0.5f * t;
        t = 0.5f / t;
        q[3] = (m.m[k * 3 + j] - m.m[j * 3 + k]) * t;
        q[j] = (m.m[j * 3 + i] + m.m[i * 3 + j]) * t;
        q[k] = (m.m[k * 3 + i] + m.m[i * 3 + k]) * t;
    }
}

// RigidBody::step (rigidbody.cpp:207-243) for a dynamic body: origin += v dt; orientation
// q += 1/2 dt (0, w) q in the FP32 quaternion-product order of Eigen's SSE path
// (Geometry_SIMD.h), normalise with sqrt((x^2+z^2)+(y^2+w^2)), back to a matrix
// (QuaternionBase::toRotationMatrix), then I_w^-1 = (R L) R^T (rigidbody.cpp:43-46).
PE_HD void integrate_pose(f3& pos, m33& R, f3 v, f3 w, float dt) {
    pos = add(pos, mul(v, dt));
    float q[4];
    quat_from_matrix(R, q);
    const f3 dr = mul(mul(w, dt), 0.5f);
    const float ax = dr.x, ay = dr.y, az = dr.z, aw = 0.0f;
    const float bx = q[0], by = q[1], bz = q[2], bw = q[3];
    const float dx = (ax * bw - az * by) + (ay * bz + aw * bx);
    const float dy = (ay * bw - ax * bz) + (az * bx + aw * by);
    const float dz = (az * bw - ay * bx) + (ax * by + aw * bz);
    const float dw = (aw * bw - ax * bx) - (az * bz + ay * by);
    float x = bx + dx, y = by + dy, z = bz + dz, ww = bw + dw;
    const float n = sqrtf((x * x + z * z) + (y * y + ww * ww));
    if (n > 0.0f) { x /= n; y /= n; z /= n; ww /= n; }
    const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    const float twx = tx * ww, twy = ty * ww, twz = tz * ww;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R.m[0] = 1.0f - (tyy + tzz); R.m[1] = txy - twz; R.m[2] = txz + twy;
    R.m[3] = txy + twz; R.m[4] = 1.0f - (txx + tzz); R.m[5] = tyz - twx;
    R.m[6] = txz - twy; R.m[7] = tyz + twx; R.m[8] = 1.0f - (txx + tyy);
}

// Shape::getAABB -- box (box_shape.cpp:27-43), sphere (sphere_shape.cpp:14-18),
// cylinder (cylinder_shape.cpp:34-43).  p = shape parameters (box size / radius / radius,height).
PE_HD void shape_aabb(int type, const float p[4], const m33& R, f3 pos, f3& mn, f3& mx) {
    if (type == SHAPE_BOX) {
        const float hx = p[0] / 2.0f, hy = p[1] / 2.0f, hz = p[2] / 2.0f;
        mn = f3{REAL_MAX, REAL_MAX, REAL_MAX};
        mx = f3{REAL_MIN, REAL_MIN, REAL_MIN};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const f3 c{(i & 1) ? hx : -hx, (i & 2) ? hy : -hy, (i & 4) ? hz : -hz};
            const f3 v = mv(R, c);
            mn.x = mn.x < v.x ? mn.x : v.x; mn.y = mn.y < v.y ? mn.y : v.y; mn.z = mn.z < v.z ? mn.z : v.z;
            mx.x = mx.x > v.x ? mx.x : v.x; mx.y = mx.y > v.y ? mx.y : v.y; mx.z = mx.z > v.z ? mx.z : v.z;
        }
        mn = add(mn, pos);
        mx = add(mx, pos);
    } else if (type == SHAPE_SPHERE) {
        const f3 r{p[0], p[0], p[0]};
        mn = sub(pos, r);
        mx = add(pos, r);
    } else {  // cylinder (axis = basis column 1)
        const f3 a = col(R, 1);
        const float hh = p[1] * 0.5f;
        f3 e{fabsf(a.x) * hh, fabsf(a.y) * hh, fabsf(a.z) * hh};
        e.y += p[0] * sqrtf(a.x * a.x + a.z * a.z);
        e.x += p[0] * sqrtf(a.y * a.y + a.z * a.z);
        e.z += p[0] * sqrtf(a.x * a.x + a.y * a.y);
        mn = sub(pos, e);
        mx = add(pos, e);
    }
}

// RigidBody::getAABBScale (rigidbody.cpp:119-123)
PE_HD float aabb_scale(f3 mn, f3 mx, f3 origin) {
    const float s = norm(sub(mx, mn));
    const float d = norm(sub(add(mx, mn), mul(origin, 2.0f)));
    return (s + d) * 0.5f;
}

// ------------------------------------------------------------------ contact cache
// One ContactResult (contact_result.h:76-101): 8 slots, each ContactPoint's fields.
struct Cache {
    float wx[8], wy[8], wz[8];      // world_pos (on B); wx >= REAL_MAX marks an invalid slot
    float nx[8], ny[8], nz[8];      // world_normal
    float ax[8], ay[8], az[8];      // local_pos_a
    float bx[8], by[8], bz[8];      // local_pos_b
    float d[8];                     // distance (kept on invalid slots, as in the reference)
    int trunc;                      // set when a clipped polygon outgrew its buffer (points lost)
};
static_assert((sizeof(Cache) / 4) % 2 == 1, "an odd row length: lanes of an LDS cache array hit distinct banks");

// Body pose needed by the narrowphase
struct Pose {
    m33 R;
    f3 p;
    float scale;       // aabb scale
};
PE_HD f3 xf_apply(const Pose& P, f3 v) { return add(mv(P.R, v), P.p); }
PE_HD f3 xf_inv(const Pose& P, f3 v) { return mtv(P.R, sub(v, P.p)); }

// The contact result's objects (their poses turn contact points into local positions) and the
// result's swap flag; the shapes' own poses are passed to each algorithm (a compound's children are
// posed apart from their body).
struct PairCtx {
    Pose A, B;
    bool swap;
    float same_th2;    // (min(scaleA, scaleB) * PE_DIST_TH)^2
};

// All-ones when i == k, else zero: masks for selecting among registers by a run-time index.
PE_HD int sel_mask(int i, int k) { return -(int)(i == k); }

// ContactResult::addContactPoint (contact_result.cpp:88-142)
PE_HD void add_point(Cache& c, const PairCtx& x, f3 world_normal, f3 world_pos, float depth) {
    f3 pa = world_pos, pb = world_pos, n = world_normal, point = world_pos;
    if (x.swap) {
        n = neg(n);
        pb = add(world_pos, mul(world_normal, depth));
        point = pb;
    } else {
        pa = add(world_pos, mul(world_normal, depth));
    }
    const f3 la = xf_inv(x.A, pa);
    const f3 lb = xf_inv(x.B, pb);
    float min_dist = x.same_th2;
    int near = -1;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (!(c.wx[i] < REAL_MAX + 0.1f)) continue;
        const f3 dd{c.bx[i] - lb.x, c.by[i] - lb.y, c.bz[i] - lb.z};
        const float dist = dot(dd, dd);
        if (dist < min_dist) { min_dist = dist; near = i; }
    }
    int slot = -1;
    if (near >= 0) {
        int dn = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) dn |= __float_as_int(c.d[i]) & sel_mask(near, i);
        if (depth < __int_as_float(dn)) slot = near;
    } else {
#pragma unroll
        for (int i = 7; i >= 0; i--) if (!(c.wx[i] < REAL_MAX + 0.1f)) slot = i;
        if (slot < 0) {
            float max_dist = REAL_MIN;
#pragma unroll
            for (int i = 0; i < 8; i++) if (c.d[i] > max_dist) { max_dist = c.d[i]; slot = i; }
        }
    }
    if (slot < 0) return;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (i != slot) continue;
        c.wx[i] = point.x; c.wy[i] = point.y; c.wz[i] = point.z;
        c.nx[i] = n.x; c.ny[i] = n.y; c.nz[i] = n.z;
        c.ax[i] = la.x; c.ay[i] = la.y; c.az[i] = la.z;
        c.bx[i] = lb.x; c.by[i] = lb.y; c.bz[i] = lb.z;
        c.d[i] = depth;
    }
}

// ContactResult::sortContactPoints (contact_result.cpp:144-156): libstdc++ std::sort of 8
// elements is an insertion sort, i.e. stable by distance, invalid slots included.  Returns
// point_size (the first invalid position after sorting) and, per slot k, its sorted position.
PE_HD int sort_rank(const Cache& c, int rank[8]) {
    int perm[8];
    float key[8];
#pragma unroll
    for (int i = 0; i < 8; i++) { perm[i] = i; key[i] = c.d[i]; }
#pragma unroll
    for (int i = 1; i < 8; i++) {
#pragma unroll
        for (int k = i; k > 0; k--) {
            if (key[k] < key[k - 1]) {
                const float tk = key[k]; key[k] = key[k - 1]; key[k - 1] = tk;
                const int tp = perm[k]; perm[k] = perm[k - 1]; perm[k - 1] = tp;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) if (perm[i] == k) rank[k] = i;
    }
    int np = 8;
#pragma unroll
    for (int k = 0; k < 8; k++) if (!(c.wx[k] < REAL_MAX + 0.1f) && rank[k] < np) np = rank[k];
    return np;
}

// ------------------------------------------------------------------ box-box
// BoxBoxCollisionAlgorithm (box_box_collision_algorithm.cpp:137-580), ODE dBoxBox.
// Small arrays indexed at run time, kept in registers: every access is unrolled over the slots
// and combines them with bit masks (a select chain would be folded back into an indexed
// private-memory access), so no scratch memory is used.
template <int N> struct RArr {
    float v[N];
    PE_HD float get(int i) const {
        int r = 0;
#pragma unroll
        for (int k = 0; k < N; k++) r |= __float_as_int(v[k]) & sel_mask(i, k);
        return __int_as_float(r);
    }
    PE_HD void set(int i, float x) {
        const int xb = __float_as_int(x);
#pragma unroll
        for (int k = 0; k < N; k++) {
            const int m = sel_mask(i, k);
            v[k] = __int_as_float((xb & m) | (__float_as_int(v[k]) & ~m));
        }
    }
};
template <int N> struct RArrI {
    int v[N];
    PE_HD int get(int i) const {
        int r = 0;
#pragma unroll
        for (int k = 0; k < N; k++) r |= v[k] & sel_mask(i, k);
        return r;
    }
    PE_HD void set(int i, int x) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const int m = sel_mask(i, k);
            v[k] = (x & m) | (v[k] & ~m);
        }
    }
};

// intersectRectQuad2 (:20-71): clips the quad (4 points) against the rectangle |x| <= h0,
// |y| <= h1, edge by edge (x then y, negative side first); stops as soon as a stage has
// produced 8 points.  Returns the point count; the points are (ox[k], oy[k]).
PE_HD int intersect_rect_quad2(float h0, float h1, const float quad[8], RArr<8>& ox, RArr<8>& oy) {
    RArr<8> qx, qy;
#pragma unroll
    for (int k = 0; k < 8; k++) { qx.v[k] = k < 4 ? quad[2 * k] : 0.0f; qy.v[k] = k < 4 ? quad[2 * k + 1] : 0.0f; }
    int nq = 4;
    bool done = false;
#pragma unroll
    for (int st = 0; st < 4; st++) {
        const int dir = st >> 1;
        const float sgn = (st & 1) ? 1.0f : -1.0f;
        const float hd = dir == 0 ? h0 : h1;
        RArr<8> rx, ry;
#pragma unroll
        for (int k = 0; k < 8; k++) { rx.v[k] = 0.0f; ry.v[k] = 0.0f; }
        int nr = 0;
        if (!done) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (i < nq && !done) {
                    const float cx = qx.v[i], cy = qy.v[i];
                    const bool wrap = !(i < 7 && i + 1 < nq);
                    const float nx = wrap ? qx.v[0] : qx.v[i < 7 ? i + 1 : 0];
                    const float ny = wrap ? qy.v[0] : qy.v[i < 7 ? i + 1 : 0];
                    const float cd = dir == 0 ? cx : cy, nd = dir == 0 ? nx : ny;
                    const float co = dir == 0 ? cy : cx, no = dir == 0 ? ny : nx;
                    const bool cin = sgn * cd < hd;
                    if (cin) {
                        rx.set(nr, cx); ry.set(nr, cy);
                        nr++;
                        if (nr & 8) done = true;
                    }
                    if (!done && (cin ^ (sgn * nd < hd))) {
                        const float other = co + (no - co) / (nd - cd) * (sgn * hd - cd);
                        rx.set(nr, dir == 0 ? sgn * hd : other);
                        ry.set(nr, dir == 0 ? other : sgn * hd);
                        nr++;
                        if (nr & 8) done = true;
                    }
                }
            }
            qx = rx; qy = ry;
            nq = nr;
        }
    }
    ox = qx; oy = qy;
    return nq;
}

// cullPoints2 (:73-135) for m = 4 of n (5..8) points: keeps point i0 and the three points
// closest in angle (around the centroid) to i0's angle + k 90 degrees.
PE_HD void cull_points2(int n, const RArr<8>& px, const RArr<8>& py, int i0, int iret[4]) {
    const int m = 4;
    float a = 0, cx = 0, cy = 0, q;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        if (i < n - 1) {
            const float p0 = px.v[i], p1 = py.v[i], p2 = px.v[i + 1], p3 = py.v[i + 1];
            q = p0 * p3 - p2 * p1;
            a += q;
            cx += q * (p0 + p2);
            cy += q * (p1 + p3);
        }
    }
    {
        const float p0 = px.v[0], p1 = py.v[0], pm1 = py.get(n - 1), pm2 = px.get(n - 1);
        q = pm2 * p1 - p0 * pm1;
        if (fabsf(a + q) > EPS) a = 1.0f / (3.0f * (a + q));
        else a = REAL_MAX;
        cx = a * (cx + q * (pm2 + p0));
        cy = a * (cy + q * (pm1 + p1));
    }
    RArr<8> A;
    RArrI<8> avail;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        A.v[i] = i < n ? atan2f(py.v[i] - cy, px.v[i] - cx) : 0.0f;
        avail.v[i] = i < n ? 1 : 0;
    }
    avail.set(i0, 0);
    iret[0] = i0;
    const float Ai0 = A.get(i0);
#pragma unroll
    for (int j = 1; j < m; j++) {
        a = (float)j * (2.0f * PI_R / (float)m) + Ai0;
        if (a > PI_R) a -= 2.0f * PI_R;
        float max_diff = REAL_MAX;
        int pick = i0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (avail.v[i]) {
                float diff = fabsf(A.v[i] - a);
                if (diff > PI_R) diff = 2.0f * PI_R - diff;
                if (diff < max_diff) { max_diff = diff; pick = i; }
            }
        }
        iret[j] = pick;
        avail.set(pick, 0);
    }
}

// PA, PB: the two boxes' poses (the bodies', or compound children's); x carries the result's objects.
PE_HD void box_box(Cache& c, PairCtx& x, const Pose& PA, const Pose& PB, const float sa[3], const float sb[3]) {
    const float fudge = 1.05f;
    const f3 p1 = PA.p, p2 = PB.p;
    const m33& R1 = PA.R;
    const m33& R2 = PB.R;
    const f3 p = sub(p2, p1);
    const f3 pp = mtv(R1, p);
    const float A[3] = {sa[0] * 0.5f, sa[1] * 0.5f, sa[2] * 0.5f};
    const float B[3] = {sb[0] * 0.5f, sb[1] * 0.5f, sb[2] * 0.5f};
    const m33 Rij = mtm(R1, R2);
#define RIJ(i, j) Rij.m[(i) * 3 + (j)]
    float Q11 = fabsf(RIJ(0, 0)), Q12 = fabsf(RIJ(0, 1)), Q13 = fabsf(RIJ(0, 2));
    float Q21 = fabsf(RIJ(1, 0)), Q22 = fabsf(RIJ(1, 1)), Q23 = fabsf(RIJ(1, 2));
    float Q31 = fabsf(RIJ(2, 0)), Q32 = fabsf(RIJ(2, 1)), Q33 = fabsf(RIJ(2, 2));
    float s = REAL_MIN, s2, l;
    int invert_normal = 0, code = 0;
    bool use_normalR = false;
    int normalR_col = 0;          // normalR is column normalR_col of R1 (1-3) or R2 (4-6)
    f3 normalC{0, 0, 0};
#define PE_TST(e1, e2, nrcol, cc)                                       \
    {                                                                   \
        const float ex1 = (e1);                                         \
        s2 = fabsf(ex1) - (e2);                                         \
        if (s2 > 0) return;                                             \
        if (s2 > s) { s = s2; normalR_col = nrcol; use_normalR = true; invert_normal = ex1 < 0; code = cc; } \
    }
    PE_TST(pp.x, A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13, 0, 1)
    PE_TST(pp.y, A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23, 1, 2)
    PE_TST(pp.z, A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33, 2, 3)
    PE_TST(dot(col(R2, 0), p), A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0], 3, 4)
    PE_TST(dot(col(R2, 1), p), A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1], 4, 5)
    PE_TST(dot(col(R2, 2), p), A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2], 5, 6)
#undef PE_TST
#define PE_TST(e1, e2, n1, n2, n3, cc)                                  \
    {                                                                   \
        const float ex1 = (e1);                                         \
        s2 = fabsf(ex1) - (e2);                                         \
        if (s2 > EPS) return;                                           \
        const float m1 = (n1), m2 = (n2), m3 = (n3);                    \
        l = sqrtf(m1 * m1 + m2 * m2 + m3 * m3);                         \
        if (l > EPS) {                                                  \
            s2 /= l;                                                    \
            if (s2 * fudge > s) {                                       \
                s = s2; use_normalR = false;                            \
                normalC = f3{m1 / l, m2 / l, m3 / l};                   \
                invert_normal = ex1 < 0; code = cc;                     \
            }                                                           \
        }                                                               \
    }
    Q11 += EPS; Q12 += EPS; Q13 += EPS;
    Q21 += EPS; Q22 += EPS; Q23 += EPS;
    Q31 += EPS; Q32 += EPS; Q33 += EPS;
    PE_TST(pp.z * RIJ(1, 0) - pp.y * RIJ(2, 0), (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0.0f, -RIJ(2, 0), RIJ(1, 0), 7)
    PE_TST(pp.z * RIJ(1, 1) - pp.y * RIJ(2, 1), (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0.0f, -RIJ(2, 1), RIJ(1, 1), 8)
    PE_TST(pp.z * RIJ(1, 2) - pp.y * RIJ(2, 2), (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0.0f, -RIJ(2, 2), RIJ(1, 2), 9)
    PE_TST(pp.x * RIJ(2, 0) - pp.z * RIJ(0, 0), (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), RIJ(2, 0), 0.0f, -RIJ(0, 0), 10)
    PE_TST(pp.x * RIJ(2, 1) - pp.z * RIJ(0, 1), (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), RIJ(2, 1), 0.0f, -RIJ(0, 1), 11)
    PE_TST(pp.x * RIJ(2, 2) - pp.z * RIJ(0, 2), (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), RIJ(2, 2), 0.0f, -RIJ(0, 2), 12)
    PE_TST(pp.y * RIJ(0, 0) - pp.x * RIJ(1, 0), (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -RIJ(1, 0), RIJ(0, 0), 0.0f, 13)
    PE_TST(pp.y * RIJ(0, 1) - pp.x * RIJ(1, 1), (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -RIJ(1, 1), RIJ(0, 1), 0.0f, 14)
    PE_TST(pp.y * RIJ(0, 2) - pp.x * RIJ(1, 2), (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -RIJ(1, 2), RIJ(0, 2), 0.0f, 15)
#undef PE_TST
#undef RIJ
    if (!code) return;
    f3 normal;
    if (use_normalR) normal = normalR_col < 3 ? col(R1, normalR_col) : col(R2, normalR_col - 3);
    else normal = mv(R1, normalC);
    if (invert_normal) normal = neg(normal);
    const float depth = -s;

    if (code > 6) {  // edge-edge (:281-348)
        f3 pa = p1;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const f3 cj = col(R1, j);
            const float sign = dot(normal, cj) > 0 ? 1.0f : -1.0f;
            pa.x += sign * A[j] * cj.x; pa.y += sign * A[j] * cj.y; pa.z += sign * A[j] * cj.z;
        }
        f3 pb = p2;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const f3 cj = col(R2, j);
            const float sign = dot(normal, cj) > 0 ? -1.0f : 1.0f;
            pb.x += sign * B[j] * cj.x; pb.y += sign * B[j] * cj.y; pb.z += sign * B[j] * cj.z;
        }
        const int quotient = (code - 7) / 3;
        const f3 ua = col(R1, quotient), ub = col(R2, quotient);
        const f3 b_a = sub(pb, pa);
        const float ua_ub = dot(ua, ub);
        const float q1 = dot(ua, b_a);
        const float q2 = -dot(ub, b_a);
        float d = 1 - ua_ub * ua_ub;
        float beta;
        if (d <= EPS) { beta = 0; }
        else { d = 1.0f / d; beta = (ua_ub * q1 + q2) * d; }
        pb.x += ub.x * beta; pb.y += ub.y * beta; pb.z += ub.z * beta;
        add_point(c, x, neg(normal), add(pb, mul(normal, 0.0f)), -depth + 0.0f * 2);
        return;
    }

    // face-something (:350-579)
    const f3 SA3{A[0], A[1], A[2]}, SB3{B[0], B[1], B[2]};
    f3 Sa, Sb;
    m33 Ra, Rb;
    f3 pa, pb;
    if (code <= 3) { x.swap = false; Ra = R1; Rb = R2; pa = p1; pb = p2; Sa = SA3; Sb = SB3; }
    else { x.swap = true; Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = SB3; Sb = SA3; }
    const f3 normal2 = code <= 3 ? normal : neg(normal);
    const f3 nr = mtv(Rb, normal2);
    const f3 anr{fabsf(nr.x), fabsf(nr.y), fabsf(nr.z)};
    int lan_r, a1, a2;
    if (anr.y > anr.x) {
        if (anr.y > anr.z) { a1 = 0; lan_r = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lan_r = 2; }
    } else {
        if (anr.x > anr.z) { lan_r = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lan_r = 2; }
    }
    f3 center;
    {
        const float sbl = comp(Sb, lan_r);
        const f3 rl = col(Rb, lan_r);
        if (comp(nr, lan_r) < 0) center = f3{pb.x - pa.x + sbl * rl.x, pb.y - pa.y + sbl * rl.y, pb.z - pa.z + sbl * rl.z};
        else center = f3{pb.x - pa.x - sbl * rl.x, pb.y - pa.y - sbl * rl.y, pb.z - pa.z - sbl * rl.z};
    }
    const int codeN = code <= 3 ? code - 1 : code - 4;
    int code1, code2;
    if (codeN == 0) { code1 = 1; code2 = 2; }
    else if (codeN == 1) { code1 = 0; code2 = 2; }
    else { code1 = 0; code2 = 1; }
    const f3 rc1 = col(Ra, code1), rc2 = col(Ra, code2), rba1 = col(Rb, a1), rba2 = col(Rb, a2);
    const float c1 = dot(center, rc1), c2 = dot(center, rc2);
    float m11 = dot(rc1, rba1), m12 = dot(rc1, rba2), m21 = dot(rc2, rba1), m22 = dot(rc2, rba2);
    float quad[8];
    {
        const float sba1 = comp(Sb, a1), sba2 = comp(Sb, a2);
        const float k1 = m11 * sba1, k2 = m21 * sba1, k3 = m12 * sba2, k4 = m22 * sba2;
        const float c1mk1 = c1 - k1, c2mk2 = c2 - k2;
        quad[0] = c1mk1 - k3; quad[1] = c2mk2 - k4; quad[2] = c1mk1 + k3; quad[3] = c2mk2 + k4;
        const float c1pk1 = c1 + k1, c2pk2 = c2 + k2;
        quad[4] = c1pk1 + k3; quad[5] = c2pk2 + k4; quad[6] = c1pk1 - k3; quad[7] = c2pk2 - k4;
    }
    RArr<8> rx, ry;
    const int n = intersect_rect_quad2(comp(Sa, code1), comp(Sa, code2), quad, rx, ry);
    if (n < 1) return;
    RArr<8> ptx, pty, ptz, dep;
    const float det1 = 1.0f / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    const float san = comp(Sa, codeN);
    int c_num = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (j < n) {
            const float rxj = rx.v[j], ryj = ry.v[j];
            const float k1 = m22 * (rxj - c1) - m12 * (ryj - c2);
            const float k2 = -m21 * (rxj - c1) + m11 * (ryj - c2);
            const f3 pt{center.x + k1 * rba1.x + k2 * rba2.x, center.y + k1 * rba1.y + k2 * rba2.y,
                        center.z + k1 * rba1.z + k2 * rba2.z};
            ptx.set(c_num, pt.x); pty.set(c_num, pt.y); ptz.set(c_num, pt.z);
            const float dj = san - dot(normal2, pt);
            dep.set(c_num, dj);
            if (dj >= 0) {
                rx.set(c_num, rxj);
                ry.set(c_num, ryj);
                c_num++;
            }
        }
    }
    if (c_num < 1) return;
    int max_c = 4 < c_num ? 4 : c_num;
    max_c = max_c > 1 ? max_c : 1;
    const f3 nv = normal;
    const f3 marg = mul(nv, 0.0f);
    if (c_num <= max_c) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < c_num) {
                const f3 pw{ptx.v[j] + pa.x, pty.v[j] + pa.y, ptz.v[j] + pa.z};
                if (code < 4) add_point(c, x, neg(nv), add(pw, marg), -dep.v[j] + 0.0f * 2);
                else add_point(c, x, nv, sub(pw, marg), -dep.v[j] + 0.0f * 2);
            }
        }
    } else {
        int i1 = 0;
        float maxd = dep.v[0];
#pragma unroll
        for (int i = 1; i < 8; i++) if (i < c_num && dep.v[i] > maxd) { maxd = dep.v[i]; i1 = i; }
        int iret[4];
        cull_points2(c_num, rx, ry, i1, iret);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = iret[j];
            const f3 pw{ptx.get(k) + pa.x, pty.get(k) + pa.y, ptz.get(k) + pa.z};
            const float dk = dep.get(k);
            if (code < 4) add_point(c, x, neg(nv), add(pw, marg), -dk + 0.0f * 2);
            else add_point(c, x, nv, sub(pw, marg), -dk + 0.0f * 2);
        }
    }
}

// ------------------------------------------------------------------ sphere pairs
// BoxSphereCollisionAlgorithm (box_sphere_collision_algorithm.cpp:6-140).  The +y face of the
// inside case uses half_extent.x (:51), as in the reference.
PE_HD void box_sphere(Cache& c, PairCtx& x, bool a_is_box, const Pose& PA, const Pose& PB, const float box_size[3],
                      float radius) {
    const Pose& box = a_is_box ? PA : PB;
    const Pose& sph = a_is_box ? PB : PA;
    x.swap = a_is_box;
    const f3 he{box_size[0] / 2.0f, box_size[1] / 2.0f, box_size[2] / 2.0f};
    const f3 ps = xf_inv(box, sph.p);
    f3 cp = ps;
    cp.x = he.x < cp.x ? he.x : cp.x; cp.x = -he.x > cp.x ? -he.x : cp.x;
    cp.y = he.y < cp.y ? he.y : cp.y; cp.y = -he.y > cp.y ? -he.y : cp.y;
    cp.z = he.z < cp.z ? he.z : cp.z; cp.z = -he.z > cp.z ? -he.z : cp.z;
    f3 normal = sub(ps, cp);
    const float dist2 = dot(normal, normal);
    if (!(dist2 > radius * radius)) {
        float distance;
        if (dist2 <= EPS * EPS) {
            float face = he.x - ps.x, md = face;
            cp.x = he.x; normal = f3{1, 0, 0};
            face = he.x + ps.x;
            if (face < md) { md = face; cp = ps; cp.x = -he.x; normal = f3{-1, 0, 0}; }
            face = he.y - ps.y;
            if (face < md) { md = face; cp = ps; cp.y = he.x; normal = f3{0, 1, 0}; }
            face = he.y + ps.y;
            if (face < md) { md = face; cp = ps; cp.y = -he.y; normal = f3{0, -1, 0}; }
            face = he.z - ps.z;
            if (face < md) { md = face; cp = ps; cp.z = he.z; normal = f3{0, 0, 1}; }
            face = he.z + ps.z;
            if (face < md) { md = face; cp = ps; cp.z = -he.z; normal = f3{0, 0, -1}; }
            distance = -md;
        } else {
            distance = sqrtf(dist2);
            normal = divs(normal, distance);
        }
        normal = mv(box.R, normal);
        const f3 pt = xf_apply(box, cp);
        const float dist = distance - radius;
        add_point(c, x, normal, sub(pt, mul(normal, 0.0f)), dist + 2 * 0.0f);
    }
    x.swap = false;
}

// SphereSphereCollisionAlgorithm (sphere_sphere_collision_algorithm.cpp:23-37)
PE_HD void sphere_sphere(Cache& c, PairCtx& x, const Pose& PA, const Pose& PB, float ra, float rb) {
    const f3 rel = sub(PA.p, PB.p);
    const float dist = norm(rel);
    if (dist > ra + rb || fabsf(dist - 0.0f) < EPS) return;
    const f3 normal = divs(rel, dist);
    const f3 pt = add(PB.p, mul(normal, rb));
    const float depth = dist - ra - rb;
    add_point(c, x, normal, sub(pt, mul(normal, 0.0f)), depth + 2 * 0.0f);
}

// ------------------------------------------------------------------ solver rows
// ContactPoint::getOrthoUnits (contact_result.cpp:8-31)
PE_HD void ortho_units(f3 n, f3& t1, f3& t2) {
    n = normalized(n);
    if (fabsf(n.z) > 0.7071f) {
        const float a = n.y * n.y + n.z * n.z;
        const float k = 1.0f / sqrtf(a);
        t1 = f3{0.0f, -n.z * k, n.y * k};
        t2 = f3{a * k, -n.x * t1.z, n.x * t1.y};
    } else {
        const float a = n.x * n.x + n.y * n.y;
        const float k = 1.0f / sqrtf(a);
        t1 = f3{-n.y * k, n.x * k, 0.0f};
        t2 = f3{-n.z * t1.y, n.z * t1.x, a * k};
    }
}

// FrictionContactConstraint::ConstraintInfo (friction_contact_constraint.h:15-29)
struct Row {
    f3 ra, rb, n, t0, t1;
    float rhs, mn, mt0, mt1, mu;
};

// One body's dynamic quantities as the solver sees them (kinematic: zero inverse mass/inertia).
struct SolverBody {
    m33 Iw;
    float im;
    bool kin;
};

// FrictionContactConstraint::initSequentialImpulse (friction_contact_constraint.cpp:9-73)
PE_HD Row make_row(const m33& RA, const m33& RB, f3 la, f3 lb, f3 n, float dist,
                   float ims, const m33& Ia, const m33& Ib, f3 va_lin, f3 wa, f3 vb_lin, f3 wb,
                   float mu, float e, float dt) {
    Row r;
    r.ra = mv(RA, la);
    r.rb = mv(RB, lb);
    r.n = n;
    ortho_units(n, r.t0, r.t1);
    {
        const f3 xa = cross(r.ra, r.n), xb = cross(r.rb, r.n);
        r.mn = 1.0f / (ims + dot(mv(Ia, xa), xa) + dot(mv(Ib, xb), xb));
    }
    {
        const f3 xa = cross(r.ra, r.t0), xb = cross(r.rb, r.t0);
        r.mt0 = 1.0f / (ims + dot(mv(Ia, xa), xa) + dot(mv(Ib, xb), xb));
    }
    {
        const f3 xa = cross(r.ra, r.t1), xb = cross(r.rb, r.t1);
        r.mt1 = 1.0f / (ims + dot(mv(Ia, xa), xa) + dot(mv(Ib, xb), xb));
    }
    const f3 va = add(va_lin, cross(wa, r.ra));
    const f3 vb = add(vb_lin, cross(wb, r.rb));
    float rev = -dot(r.n, sub(va, vb));
    if (fabsf(rev) < 0.1f) rev = 0;
    rev *= e;
    float pen = dist;
    pen = pen > -0.3f ? pen : -0.3f;
    r.rhs = (rev - 0.2f * pen / dt) * r.mn;
    r.mu = mu;
    return r;
}

// FrictionContactConstraint::iterateSequentialImpulse (friction_contact_constraint.cpp:75-108)
// for one row; updates the accumulated impulses and returns the impulse vector J
// (applied +J to A, -J to B).
PE_HD f3 iterate_row(const Row& r, float& ln, float& lt0, float& lt1,
                     f3 tva, f3 twa, f3 tvb, f3 twb) {
    const f3 vr = sub(sub(add(tva, cross(twa, r.ra)), tvb), cross(twb, r.rb));
    float dn = r.rhs - dot(r.n, vr) * r.mn;
    dn = dn > -ln ? dn : -ln;
    ln += dn;
    float t0 = lt0 - dot(r.t0, vr) * r.mt0;
    float t1 = lt1 - dot(r.t1, vr) * r.mt1;
    const float scale = r.mu * ln / sqrtf(t0 * t0 + t1 * t1);
    if (scale < 1) { t0 *= scale; t1 *= scale; }
    f3 J;
    J.x = (dn * r.n.x + (t0 - lt0) * r.t0.x) + (t1 - lt1) * r.t1.x;
    J.y = (dn * r.n.y + (t0 - lt0) * r.t0.y) + (t1 - lt1) * r.t1.y;
    J.z = (dn * r.n.z + (t0 - lt0) * r.t0.z) + (t1 - lt1) * r.t1.z;
    lt0 = t0;
    lt1 = t1;
    return J;
}

}  // namespace pe
