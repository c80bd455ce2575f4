// Raycasts against one world's bodies (rigid/raycast: RaycastBox, RaycastSphere, RaycastCylinder,
// RaycastMesh, DefaultRaycastSolver), FP32 in the reference's evaluation order.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"
#include "pe_shapes.hpp"

namespace pe {

// RaycastBox::rayHitBox (raycast_box.cpp:31-60): slab test with the direction's tiny components raised
// to PE_EPS; the normal is the entering slab's, x before y before z on ties.
PE_HD bool ray_hit_box(f3 start, f3 direction, f3 bmin, f3 bmax, float& distance, f3& hp, f3& hn) {
    const float eps = 0.00001f;
    const f3 dir{fabsf(direction.x) < eps ? eps : direction.x, fabsf(direction.y) < eps ? eps : direction.y,
                 fabsf(direction.z) < eps ? eps : direction.z};
    const f3 m{(bmax.x - start.x) / dir.x, (bmax.y - start.y) / dir.y, (bmax.z - start.z) / dir.z};
    const f3 n{(bmin.x - start.x) / dir.x, (bmin.y - start.y) / dir.y, (bmin.z - start.z) / dir.z};
    const f3 tmax{m.x > n.x ? m.x : n.x, m.y > n.y ? m.y : n.y, m.z > n.z ? m.z : n.z};
    const f3 tmin{m.x < n.x ? m.x : n.x, m.y < n.y ? m.y : n.y, m.z < n.z ? m.z : n.z};
    float t_enter = tmin.x > tmin.y ? tmin.x : tmin.y;
    t_enter = t_enter > tmin.z ? t_enter : tmin.z;
    float t_exit = tmax.x < tmax.y ? tmax.x : tmax.y;
    t_exit = t_exit < tmax.z ? t_exit : tmax.z;
    if (t_enter > t_exit || t_exit < 0) return false;
    distance = t_enter;
    hp = add(start, mul(dir, t_enter));
    hn = f3{0.0f, 0.0f, 0.0f};
    if (t_enter == tmin.x) hn.x = dir.x > 0 ? -1.0f : 1.0f;
    else if (t_enter == tmin.y) hn.y = dir.y > 0 ? -1.0f : 1.0f;
    else hn.z = dir.z > 0 ? -1.0f : 1.0f;
    return true;
}

// RaycastBox::processRaycast (raycast_box.cpp:6-29)
PE_HD bool ray_box(f3 start, f3 dir, float maxd, const float size[3], const m33& R, f3 o, float& d, f3& hp, f3& hn) {
    const f3 sl = mtv(R, sub(start, o)), dl = mtv(R, dir);
    const f3 sz{size[0], size[1], size[2]};
    if (ray_hit_box(sl, dl, divs(neg(sz), 2.0f), divs(sz, 2.0f), d, hp, hn)) {
        if (d <= maxd) {
            hp = add(mv(R, hp), o);
            hn = mv(R, hn);
            return true;
        }
        return false;
    }
    hp = hn = f3{0.0f, 0.0f, 0.0f};
    d = REAL_MAX;
    return false;
}

// RaycastSphere::processRaycast (raycast_sphere.cpp:6-29); a sphere behind the start, or around it,
// is hit at a negative distance.
PE_HD bool ray_sphere(f3 start, f3 dir, float maxd, float r, f3 o, float& d, f3& hp, f3& hn) {
    const f3 oc = sub(o, start);
    const float c2s = norm(oc);
    const float s2p = dot(oc, dir);
    const float q = c2s * c2s - s2p * s2p;
    const float p2c = sqrtf(q > 0.0f ? q : 0.0f);
    if (p2c > r) {
        hp = hn = f3{0.0f, 0.0f, 0.0f};
        d = REAL_MAX;
        return false;
    }
    const float p2h = sqrtf(r * r - p2c * p2c);
    d = s2p - p2h;
    if (d <= maxd) {
        hp = add(start, mul(dir, d));
        hn = normalized(sub(hp, o));
        return true;
    }
    return false;
}

// RaycastCylinder::processRaycast (raycast_cylinder.cpp:6-77), its branches and its cap points
// (s + (h d - d * s) / d.y, with the top cap's y then set to h).
PE_HD f3 ray_cap(f3 sl, f3 dl, float h) {
    return f3{sl.x + (h * dl.x - dl.x * sl.x) / dl.y, sl.y + (h * dl.y - dl.y * sl.y) / dl.y,
              sl.z + (h * dl.z - dl.z * sl.z) / dl.y};
}
PE_HD bool ray_cylinder(f3 start, f3 dir, float maxd, float radius, float hgt, const m33& R, f3 o, float& d, f3& hp,
                        f3& hn) {
    const float height = hgt / 2.0f;
    const f3 sl = mtv(R, sub(start, o)), dl = mtv(R, dir);
    const float eps = 0.00001f;
    bool hit = false;
    if (fabsf(dl.x - 0.0f) < eps && fabsf(dl.z - 0.0f) < eps) {
        const float d2a = sqrtf(sl.x * sl.x + sl.z * sl.z);
        if (d2a <= radius) {
            const float h = dl.y < 0 ? height : -height;
            hp = ray_cap(sl, dl, h);
            hn = f3{0.0f, 1.0f, 0.0f};
            if (dl.y > 0) hn = neg(hn);
            d = norm(sub(hp, sl));
            hit = true;
        }
    } else {
        const float d2a = fabsf(dot(sl, normalized(cross(dl, f3{0.0f, 1.0f, 0.0f}))));
        if (d2a <= radius) {
            const float den = dl.x * dl.x + dl.z * dl.z;
            const float t = -(dl.x * sl.x + dl.z * sl.z) / den;
            const float dt = sqrtf(radius * radius - d2a * d2a) / sqrtf(den);
            const f3 in = add(sl, mul(dl, t - dt));
            const f3 out = add(sl, mul(dl, t + dt));
            if ((in.y < -height && out.y < -height) || (in.y > height && out.y > height)) {
                hit = false;
            } else if (in.y >= -height && in.y <= height) {
                hp = in;
                hn = normalized(f3{in.x, 0.0f, in.z});
                d = t - dt;
                hit = true;
            } else if (in.y > height) {
                hp = ray_cap(sl, dl, height);
                hp.y = height;
                hn = f3{0.0f, 1.0f, 0.0f};
                d = norm(sub(hp, sl));
                hit = true;
            } else {
                hp = ray_cap(sl, dl, -height);
                hn = neg(f3{0.0f, 1.0f, 0.0f});
                d = norm(sub(hp, sl));
                hit = true;
            }
        }
    }
    if (!hit) {
        hp = hn = f3{0.0f, 0.0f, 0.0f};
        d = REAL_MAX;
        return false;
    }
    if (d <= maxd) {
        hp = add(mv(R, hp), o);
        hn = mv(R, hn);
        return true;
    }
    return false;
}

// RaycastMesh::rayHitTriangle (raycast_mesh.cpp:52-77), Moller-Trumbore; distances behind the start
// are not rejected.
PE_HD bool ray_triangle(f3 s, f3 dir, f3 v0, f3 v1, f3 v2, float& d, f3& hit) {
    const float eps = 0.00001f;
    const f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    const f3 p = cross(dir, e2);
    const float det = dot(e1, p);
    if (det > -eps && det < eps) return false;
    const float inv = 1.0f / det;
    const f3 t = sub(s, v0);
    const float u = dot(t, p) * inv;
    if (u < 0 || u > 1.0f) return false;
    const f3 q = cross(t, e1);
    const float v = dot(dir, q) * inv;
    if (v < 0 || u + v > 1.0f) return false;
    d = dot(e2, q) * inv;
    hit = add(s, mul(dir, d));
    return true;
}

// RaycastMesh::processRaycast (raycast_mesh.cpp:9-50): the faces the BVH finds in the ray's box (start
// to start + length, grown by 0.1), each fanned from its first vertex, the first fan triangle hit
// counting for the face; the nearest face wins (strictly nearer replaces).  The BVH walk's stack is
// the LDS row of this lane (kernels of at most 64 threads per workgroup).
__device__ inline bool ray_mesh(const MeshSet& ms, int mesh, f3 start, f3 dir, float maxd, const m33& R, f3 o,
                                float& d, f3& hp, f3& hn) {
    const Poly P = poly_mesh(&ms, mesh);
    const f3 sl = mtv(R, sub(start, o)), dl = mtv(R, dir);
    const f3 el = add(sl, mul(dl, maxd));
    const f3 bmin{(sl.x < el.x ? sl.x : el.x) - 0.1f, (sl.y < el.y ? sl.y : el.y) - 0.1f, (sl.z < el.z ? sl.z : el.z) - 0.1f};
    const f3 bmax{(sl.x > el.x ? sl.x : el.x) + 0.1f, (sl.y > el.y ? sl.y : el.y) + 0.1f, (sl.z > el.z ? sl.z : el.z) + 0.1f};
    d = REAL_MAX;
    hp = hn = f3{0.0f, 0.0f, 0.0f};
    bvh_visit<1>(ms, mesh, bmin, bmax, [&](int f) {
        const int nf = face_len(P, f);
        const f3 v0 = face_vertex(P, f, 0);
        for (int i = 0; i < nf - 2; i++) {
            float t;
            f3 hit;
            if (ray_triangle(sl, dl, v0, face_vertex(P, f, i + 1), face_vertex(P, f, i + 2), t, hit)) {
                if (t < d) {
                    d = t;
                    hp = hit;
                    hn = face_normal(P, f);
                }
                break;
            }
        }
    });
    if (d < maxd) {
        hp = add(mv(R, hp), o);
        hn = mv(R, hn);
        return true;
    }
    return false;
}

}  // namespace pe
