// Device narrowphase for cylinders, convex meshes and concave (triangle) meshes (gfx950, FP32).
//
// Each routine restates the reference algorithm cited beside it, in the reference's FP32
// operation order (a 3-term sum is p0 + (p1 + p2)), so a contact computed here is the one the
// reference computes.  Shapes are read through Poly, a view that serves vertices, faces and
// unique edges of a box or cylinder (generated from the unit meshes and the body's size) or of
// a prepared mesh in the MeshSet arena, or of one triangle of a concave mesh.
#pragma once
#include "pe_common.hpp"
#include "pe_device.hpp"

namespace pe {

enum PolyKind : int { PK_BOX = 0, PK_CYL = 1, PK_MESH = 2, PK_TRI = 3 };

// Unit box mesh (shape/default_mesh.h:7-57) as sign patterns: vertex k of face f is
// BOX_V[4 f + k]; faces +y, -y, +z, -z, -x, +x.
__device__ __constant__ const signed char BOX_V[24][3] = {
    {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {1, -1, -1}, {1, -1, 1},
    {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {-1, 1, -1},
    {-1, 1, -1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}, {1, 1, 1}, {1, -1, 1}, {1, -1, -1}, {1, 1, -1}};
__device__ __constant__ const signed char BOX_N[6][3] = {{0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {-1, 0, 0}, {1, 0, 0}};
__device__ __constant__ const signed char BOX_E[12][2][3] = {
    {{-1, 1, 1}, {1, 1, 1}}, {{-1, -1, 1}, {1, -1, 1}}, {{-1, -1, -1}, {1, -1, -1}}, {{-1, 1, -1}, {1, 1, -1}},
    {{1, -1, 1}, {1, 1, 1}}, {{-1, -1, 1}, {-1, 1, 1}}, {{-1, -1, -1}, {-1, 1, -1}}, {{1, -1, -1}, {1, 1, -1}},
    {{1, 1, 1}, {1, 1, -1}}, {{-1, 1, 1}, {-1, 1, -1}}, {{-1, -1, 1}, {-1, -1, -1}}, {{1, -1, 1}, {1, -1, -1}}};

PE_HD f3 ld4(const float4& a) { return f3{a.x, a.y, a.z}; }

struct Poly {
    int kind;
    f3 h;               // box: half extents; cylinder: (radius, half height, 0)
    float r2, hgt;      // cylinder: 2 radius, height (the mesh scale)
    const MeshSet* ms;  // cylinder unit tables / mesh arena
    const float4* v;    // mesh: vertices, face normals, faces, face indices, edges
    const float4* fn;
    const int2* face;
    const int32_t* fi;
    const float4* e;
    int nv, nf, ne;
    f3 t[3], tn;        // triangle: vertices and normal (mesh frame)
};

PE_HD f3 cyl_ring(const Poly& P, bool top, int i) {
    const float* u = P.ms->cyl + (top ? 0 : 40);
    return f3{u[i] * P.r2, top ? 0.5f * P.hgt : -0.5f * P.hgt, u[20 + i] * P.r2};
}
PE_HD int n_faces(const Poly& P) { return P.kind == PK_BOX ? 6 : P.kind == PK_CYL ? 42 : P.kind == PK_MESH ? P.nf : 1; }
PE_HD int n_edges(const Poly& P) { return P.kind == PK_BOX ? 12 : P.kind == PK_CYL ? 80 : P.kind == PK_MESH ? P.ne : 0; }
PE_HD f3 face_normal(const Poly& P, int f) {
    if (P.kind == PK_BOX) return f3{(float)BOX_N[f][0], (float)BOX_N[f][1], (float)BOX_N[f][2]};
    if (P.kind == PK_CYL) {
        if (f < 2) return f3{0.0f, f == 0 ? 1.0f : -1.0f, 0.0f};
        const float* n = P.ms->cyl + 80 + 3 * (f - 2);
        return f3{n[0], n[1], n[2]};
    }
    if (P.kind == PK_MESH) return ld4(P.fn[f]);
    return P.tn;
}
PE_HD int face_len(const Poly& P, int f) {
    if (P.kind == PK_BOX) return 4;
    if (P.kind == PK_CYL) return f < 2 ? 20 : 3;
    if (P.kind == PK_MESH) return P.face[f].y;
    return 3;
}
// k-th vertex of face f
PE_HD f3 face_vertex(const Poly& P, int f, int k) {
    if (P.kind == PK_BOX) {
        const signed char* s = BOX_V[4 * f + k];
        return f3{s[0] * P.h.x, s[1] * P.h.y, s[2] * P.h.z};
    }
    if (P.kind == PK_CYL) {
        if (f == 0) return cyl_ring(P, true, k);
        if (f == 1) return cyl_ring(P, false, 19 - k);
        if (f < 22) {  // up faces {60 + j, 40 + (j + 1) % 20, 40 + j}
            const int j = f - 2;
            return k == 0 ? cyl_ring(P, false, j) : cyl_ring(P, true, k == 1 ? (j + 1) % 20 : j);
        }
        const int j = f - 22;  // down faces {40 + j, 60 + (j + 19) % 20, 60 + j}
        return k == 0 ? cyl_ring(P, true, j) : cyl_ring(P, false, k == 1 ? (j + 19) % 20 : j);
    }
    if (P.kind == PK_MESH) return ld4(P.v[P.fi[P.face[f].x + k]]);
    return P.t[k];
}
PE_HD void edge_of(const Poly& P, int e, f3& a, f3& b) {
    if (P.kind == PK_BOX) {
        const signed char* s = BOX_E[e][0];
        const signed char* t = BOX_E[e][1];
        a = f3{s[0] * P.h.x, s[1] * P.h.y, s[2] * P.h.z};
        b = f3{t[0] * P.h.x, t[1] * P.h.y, t[2] * P.h.z};
    } else if (P.kind == PK_CYL) {
        if (e < 20) { a = cyl_ring(P, true, e); b = cyl_ring(P, true, (e + 1) % 20); }
        else if (e < 40) { const int k = e - 20; a = cyl_ring(P, false, 19 - k); b = cyl_ring(P, false, (38 - k) % 20); }
        else {
            const int k = (e - 40) >> 1;
            if (((e - 40) & 1) == 0) { a = cyl_ring(P, true, k); b = cyl_ring(P, false, k); }
            else { a = cyl_ring(P, false, k); b = cyl_ring(P, true, (k + 1) % 20); }
        }
    } else {
        a = ld4(P.e[2 * e]);
        b = ld4(P.e[2 * e + 1]);
    }
}

PE_HD Poly poly_box(const float size[3]) {
    Poly P;
    P.kind = PK_BOX;
    P.h = f3{size[0] / 2.0f, size[1] / 2.0f, size[2] / 2.0f};
    P.ms = nullptr;
    return P;
}
PE_HD Poly poly_cyl(const MeshSet* ms, float r, float h) {
    Poly P;
    P.kind = PK_CYL;
    P.h = f3{r, h * 0.5f, 0.0f};
    P.r2 = r * 2.0f;
    P.hgt = h;
    P.ms = ms;
    return P;
}
PE_HD Poly poly_mesh(const MeshSet* ms, int mesh) {
    Poly P;
    P.kind = PK_MESH;
    P.ms = ms;
    const int32_t* r = ms->rec + MRW * mesh;
    const float4* A = ms->arena;
    P.v = A + r[MR_V]; P.fn = A + r[MR_FN];
    P.face = reinterpret_cast<const int2*>(A + r[MR_FACE]);
    P.fi = reinterpret_cast<const int32_t*>(A + r[MR_FI]);
    P.e = A + r[MR_E];
    P.nv = r[MR_NV]; P.nf = r[MR_NF]; P.ne = r[MR_NE];
    return P;
}

// Pose helpers: Transform::inverse and Transform::operator* (transform.cpp:7-9, 66-70)
PE_HD Pose pose_inv(const Pose& t) {
    Pose r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.R.m[i * 3 + j] = t.R.m[j * 3 + i];
    r.p = neg(mtv(t.R, t.p));
    r.scale = t.scale;
    return r;
}
PE_HD Pose pose_mul(const Pose& a, const Pose& b) {
    Pose r;
    r.R = mm(a.R, b.R);
    r.p = add(mv(a.R, b.p), a.p);
    r.scale = b.scale;
    return r;
}

// ------------------------------------------------------------------ shape queries
// ConvexMeshShape::getAABB (convex_mesh_shape.cpp:138-150)
__device__ inline void mesh_aabb(const MeshSet& ms, int mesh, const m33& R, f3 pos, f3& mn, f3& mx);

// Compound child k of the table, posed by its body's basis R and origin o (Transform::operator*,
// transform.cpp:7-9): type, shape parameters, world basis and origin.
__device__ __forceinline__ int child_pose(const MeshSet& ms, int k, const m33& R, f3 o, float sp[3], m33& Rc, f3& pc) {
    const float4* r = ms.kids + (size_t)k * (CHW / 4);
    float c[CHW];
#pragma unroll
    for (int q = 0; q < CHW / 4; q++) { const float4 v = r[q]; c[4 * q] = v.x; c[4 * q + 1] = v.y; c[4 * q + 2] = v.z; c[4 * q + 3] = v.w; }
    m33 L;
#pragma unroll
    for (int q = 0; q < 9; q++) L.m[q] = c[CH_R + q];
    Rc = mm(R, L);
    pc = add(mv(R, f3{c[CH_P], c[CH_P + 1], c[CH_P + 2]}), o);
    sp[0] = c[CH_SP]; sp[1] = c[CH_SP + 1]; sp[2] = c[CH_SP + 2];
    return __float_as_int(c[CH_TYPE]);
}
// CompoundShape::getAABB (compound_shape.cpp:49-58): the union of the children's AABBs.
__device__ inline void compound_aabb(const MeshSet& ms, int cid, const m33& R, f3 pos, f3& mn, f3& mx) {
    const int2 kr = ms.krange[cid];
    mn = f3{REAL_MAX, REAL_MAX, REAL_MAX};
    mx = f3{REAL_MIN, REAL_MIN, REAL_MIN};
    for (int k = 0; k < kr.y; k++) {
        float sp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        m33 Rc;
        f3 pc, a, b;
        const int t = child_pose(ms, kr.x + k, R, pos, sp, Rc, pc);
        if (t <= SHAPE_CYLINDER) shape_aabb(t, sp, Rc, pc, a, b);
        else mesh_aabb(ms, __float_as_int(sp[0]), Rc, pc, a, b);
        mn = f3{mn.x < a.x ? mn.x : a.x, mn.y < a.y ? mn.y : a.y, mn.z < a.z ? mn.z : a.z};
        mx = f3{mx.x > b.x ? mx.x : b.x, mx.y > b.y ? mx.y : b.y, mx.z > b.z ? mx.z : b.z};
    }
}

__device__ inline void mesh_aabb(const MeshSet& ms, int mesh, const m33& R, f3 pos, f3& mn, f3& mx) {
    const int32_t* r = ms.rec + MRW * mesh;
    const float4* vv = ms.arena + r[MR_V];
    mn = f3{REAL_MAX, REAL_MAX, REAL_MAX};
    mx = f3{REAL_MIN, REAL_MIN, REAL_MIN};
    for (int i = 0; i < r[MR_NV]; i++) {
        const f3 v = mv(R, ld4(vv[i]));
        mn.x = mn.x < v.x ? mn.x : v.x; mn.y = mn.y < v.y ? mn.y : v.y; mn.z = mn.z < v.z ? mn.z : v.z;
        mx.x = mx.x > v.x ? mx.x : v.x; mx.y = mx.y > v.y ? mx.y : v.y; mx.z = mx.z > v.z ? mx.z : v.z;
    }
    mn = add(mn, pos);
    mx = add(mx, pos);
}

// Shape::project -- box (box_shape.cpp:51-68), cylinder (cylinder_shape.cpp:50-79), mesh
// (convex_mesh_shape.cpp:160-184: the max test is the else-branch of the min test).
PE_HD void project(const Poly& S, const Pose& t, f3 axis, float& mn, float& mx, f3& pmn, f3& pmx) {
    const f3 la = mtv(t.R, axis);
    const float off = dot(t.p, axis);
    if (S.kind == PK_BOX) {
        const f3 ext{la.x > 0 ? S.h.x : -S.h.x, la.y > 0 ? S.h.y : -S.h.y, la.z > 0 ? S.h.z : -S.h.z};
        const float he = dot(ext, la);
        mn = off - he; mx = off + he;
        pmn = xf_apply(t, neg(ext)); pmx = xf_apply(t, ext);
    } else if (S.kind == PK_CYL) {
        const float rad = S.h.x, hh = S.h.y;
        if (fabsf(la.x - 0.0f) < EPS && fabsf(la.z - 0.0f) < EPS) {
            const f3 up{0.0f * hh, 1.0f * hh, 0.0f * hh};
            pmn = xf_apply(t, neg(up)); pmx = xf_apply(t, up);
            mn = off - hh; mx = off + hh;
        } else if (fabsf(la.y - 0.0f) < EPS) {
            const f3 r = mul(la, rad);
            pmn = xf_apply(t, neg(r)); pmx = xf_apply(t, r);
            mn = off - rad; mx = off + rad;
        } else {
            const float r = sqrtf(la.x * la.x + la.z * la.z);
            const f3 ext{rad * la.x / r, hh * (la.y > 0 ? 1.0f : -1.0f), rad * la.z / r};
            pmn = xf_apply(t, neg(ext)); pmx = xf_apply(t, ext);
            const float el = dot(ext, la);
            mn = off - el; mx = off + el;
        }
    } else {
        mn = REAL_MAX; mx = REAL_MIN;
        f3 lmn{0, 0, 0}, lmx{0, 0, 0};
        const int nv = S.kind == PK_MESH ? S.nv : 3;
        for (int i = 0; i < nv; i++) {
            const f3 q = S.kind == PK_MESH ? ld4(S.v[i]) : S.t[i];
            const float x = dot(q, la);
            if (x < mn) { mn = x; lmn = q; }
            else if (x > mx) { mx = x; lmx = q; }
        }
        mn += off; mx += off;
        pmn = xf_apply(t, lmn); pmx = xf_apply(t, lmx);
    }
}
// The mesh branch of project over nv vertices at v (a wave's copy of a mesh's vertices in LDS).
PE_HD void project_verts(const float4* v, int nv, const Pose& t, f3 axis, float& mn, float& mx, f3& pmn, f3& pmx) {
    const f3 la = mtv(t.R, axis);
    const float off = dot(t.p, axis);
    mn = REAL_MAX; mx = REAL_MIN;
    f3 lmn{0, 0, 0}, lmx{0, 0, 0};
    for (int i = 0; i < nv; i++) {
        const f3 q = ld4(v[i]);
        const float x = dot(q, la);
        if (x < mn) { mn = x; lmn = q; }
        else if (x > mx) { mx = x; lmx = q; }
    }
    mn += off; mx += off;
    pmn = xf_apply(t, lmn); pmx = xf_apply(t, lmx);
}
// Shape::localIsInside -- box_shape.cpp:45-49, cylinder_shape.cpp:45-48, convex_mesh_shape.cpp:152-158
PE_HD bool inside(const Poly& S, f3 q) {
    if (S.kind == PK_BOX) return fabsf(q.x) <= S.h.x && fabsf(q.y) <= S.h.y && fabsf(q.z) <= S.h.z;
    if (S.kind == PK_CYL)
        return q.y >= -S.hgt * 0.5f && q.y <= S.hgt * 0.5f && q.x * q.x + q.z * q.z <= S.h.x * S.h.x;
    const int nf = n_faces(S);
    for (int f = 0; f < nf; f++)
        if (dot(face_normal(S, f), sub(q, face_vertex(S, f, 0))) >= 0) return false;
    return true;
}

// ------------------------------------------------------------------ convex SAT + clipping
// convex_convex_collision_algorithm.cpp
constexpr int CLIP_CAP = 2 * MAX_FACE_VERTS;   // clipped polygon capacity (a 20-gon cap clipped by 20 planes needs 40)

// clipFace (:29-61).  Clipping a convex polygon by a plane adds at most one vertex, and every
// mesh face has at most MAX_FACE_VERTS vertices (checked when a mesh is prepared), so an incident
// face clipped by a reference face's side planes stays within CLIP_CAP; a polygon that rounding
// made non-convex could still grow beyond it, and then trunc is set (reported as a capacity error)
// instead of points being dropped silently.
PE_HD int clip_face(const f3* in, int nin, f3* out, f3 pn, float pe, int& trunc) {
    if (nin < 2) return 0;
    int no = 0;
    f3 first = in[nin - 1];
    float ds = dot(pn, first) + pe;
    auto put = [&](f3 q) {
        if (no < CLIP_CAP) out[no++] = q;
        else trunc = 1;
    };
    for (int ve = 0; ve < nin; ve++) {
        const f3 end = in[ve];
        const float de = dot(pn, end) + pe;
        if (ds < 0) {
            if (de < 0) put(end);
            else put(add(first, mul(sub(end, first), ds * 1.0f / (ds - de))));
        } else if (de < 0) {
            put(add(first, mul(sub(end, first), ds * 1.0f / (ds - de))));
            put(end);
        }
        first = end;
        ds = de;
    }
    return no;
}
// Lanes that work on one pair together: G = 1 is a thread alone; G = 64 is a whole wave, all
// of whose lanes hold the same pair, run the same control flow on the same values, and split
// the long axis loops of the separating-axis test (lane r takes items r, r + G, ...).
template <int G>
struct Grp {
    static_assert(G == 1 || G == WAVE, "a pair is handled by one lane or by a whole wave");
    static __device__ __forceinline__ int rank() { return G == 1 ? 0 : (int)(threadIdx.x & (G - 1)); }
    static __device__ __forceinline__ bool any(bool v) {
        if (G == 1) return v;
        return __ballot(v) != 0ull;
    }
    // The lexicographically least (d, i) of the group in every lane; i = INT_MAX marks a lane
    // without a candidate.  Each lane offers the first index of its own least d, so the result
    // is the first index (in item order) reaching the least d, as a sequential scan with a
    // strict comparison finds it.
    static __device__ __forceinline__ void argmin(float& d, int& i) {
        if (G == 1) return;
#pragma unroll
        for (int m = 1; m < G; m <<= 1) {
            const float od = __shfl_xor(d, m);
            const int oi = __shfl_xor(i, m);
            if (oi != 0x7fffffff && (i == 0x7fffffff || od < d || (od == d && oi < i))) { d = od; i = oi; }
        }
    }
    static __device__ __forceinline__ float bcast(float v, int src) { return G == 1 ? v : __shfl(v, src); }
    static __device__ __forceinline__ f3 bcast(f3 v, int src) { return G == 1 ? v : f3{__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)}; }
};

// clipHullAgainstHull + clipFaceAgainstHull (:63-170)
template <int G>
__device__ __noinline__ void clip_hull_hull(Cache& c, const PairCtx& x, f3 sep, const Poly& A, const Poly& B,
                                            const Pose& ta, const Pose& tb, float min_dist, float max_dist) {
    using L = Grp<G>;
    // B's face most along sep (first of the largest), A's face most against it (first of the
    // least): as argmin of (-d, f) and (d, f)
    float bd = REAL_MAX;
    int cb = 0x7fffffff;
    const int nfb = n_faces(B);
    for (int f = L::rank(); f < nfb; f += G) {
        const float d = dot(mv(tb.R, face_normal(B, f)), sep);
        if (d > REAL_MIN && -d < bd) { bd = -d; cb = f; }
    }
    L::argmin(bd, cb);
    if (cb == 0x7fffffff) return;
    float ad = REAL_MAX;
    int ca = 0x7fffffff;
    const int nfa = n_faces(A);
    for (int f = L::rank(); f < nfa; f += G) {
        const float d = dot(mv(ta.R, face_normal(A, f)), sep);
        if (d < ad) { ad = d; ca = f; }
    }
    L::argmin(ad, ca);
    if (ca == 0x7fffffff) return;
    // polygon buffers: a lane's own (scratch) when it works alone; one copy per wave in LDS when the
    // wave's lanes work together (every lane writes the same value to the same word)
    f3 pbuf0[G == 1 ? CLIP_CAP : 1], pbuf1[G == 1 ? CLIP_CAP : 1];
    __shared__ f3 s_clip[G == 1 ? 1 : 4][2][CLIP_CAP];
    f3* buf0 = G == 1 ? pbuf0 : s_clip[(threadIdx.x / WAVE) % 4][0];
    f3* buf1 = G == 1 ? pbuf1 : s_clip[(threadIdx.x / WAVE) % 4][1];
    int n = face_len(B, cb);
    for (int k = 0; k < n; k++) buf0[k] = xf_apply(tb, face_vertex(B, cb, k));
    const int nva = face_len(A, ca);
    const f3 wn = mv(ta.R, face_normal(A, ca));
    f3* pin = buf0;
    f3* pout = buf1;
    for (int e0 = 0; e0 < nva; e0++) {
        const f3 a = face_vertex(A, ca, e0), b = face_vertex(A, ca, (e0 + 1) % nva);
        const f3 ew = mv(ta.R, sub(a, b));
        const f3 pn = neg(cross(ew, wn));
        const f3 wa = xf_apply(ta, a);
        const float pe = -dot(wa, pn);
        n = clip_face(pin, n, pout, pn, pe, c.trunc);
        f3* t = pin; pin = pout; pout = t;
    }
    const f3 pv = face_vertex(A, ca, 0);
    const f3 ln = face_normal(A, ca);
    for (int i = 0; i < n; i++) {
        const f3 vtx = pin[i];
        const float depth = dot(ln, sub(xf_inv(ta, vtx), pv));
        if (depth <= max_dist && depth >= min_dist) add_point(c, x, sep, sub(vtx, mul(sep, 0.0f)), depth + 0.0f * 2);
    }
}
// testSepAxis (:216-243)
// projA(ax, mn, mx, pmn, pmx) and projB: Shape::project of the two shapes.
template <class PA, class PB>
PE_HD bool test_axis_with(PA&& projA, PB&& projB, f3 ax, float& depth, f3& wa, f3& wb) {
    float min0, max0, min1, max1;
    f3 pmna, pmxa, pmnb, pmxb;
    projA(ax, min0, max0, pmna, pmxa);
    projB(ax, min1, max1, pmnb, pmxb);
    if (max0 < min1 || max1 < min0) return false;
    const float d0 = max0 - min1, d1 = max1 - min0;
    if (d0 < d1) { depth = d0; wa = pmxa; wb = pmnb; }
    else { depth = d1; wa = pmna; wb = pmxb; }
    return true;
}
// Orders a wave's LDS stores before its later LDS loads of other lanes' words (a wave's LDS
// instructions execute in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// A wave's copies of two convex meshes' vertices in LDS (up to MESH_LDS_VERTS each), for the
// projections of the separating-axis test: every axis projects every vertex of both shapes, and
// LDS serves a wave's same-address loads faster than the vector memory path.
constexpr int MESH_LDS_VERTS = 64;
// The same for the world directions of both shapes' edges (up to EDGE_LDS_DIRS each), the operands
// every edge-pair axis reads: computed once per pair instead of once per axis.
constexpr int EDGE_LDS_DIRS = 128;
// isOnLine (:245-249)
PE_HD bool on_line(f3 v0, f3 v1, f3 v2) {
    const f3 n = cross(sub(v1, v0), sub(v2, v0));
    return dot(n, n) < EPS;
}
// findSeparatingAxis (:251-402) + ConvexConvexCollisionAlgorithm::getClosestPoints (:408-432).
// The deep-point contact is added before the edge axes are tested and stays even when an edge
// axis then separates the shapes.  With G lanes, the face axes and the edge-pair axes are split
// over the lanes: the shapes are separated if any axis separates them (the sequential scan
// returns at the first), and the axis kept is the first in scan order of the least depth among
// those the scan would accept (an edge axis only below the face depth and with both witness
// points on their edges), which is the one the sequential scan ends with.
template <int G>
// keep_a (a wave only): the wave's LDS already holds A's vertex copy, staged by its previous call
// with the same A (the triangles of one concave pair), so it is not staged again.
__device__ __noinline__ void convex_convex(Cache& c, const PairCtx& x, const Poly& A, const Poly& B, const Pose& ta,
                                           const Pose& tb, float ref_scale, bool keep_a = false) {
    using L = Grp<G>;
    const f3 dc = sub(ta.p, tb.p);
    float dmin = REAL_MAX;
    f3 dmin_pt{0, 0, 0}, sep{0, 0, 0};
    bool pt_on_a = false;
    const int nfa = n_faces(A), nfb = n_faces(B);
    // with a wave: the vertices of a convex mesh A or B staged in LDS (same values, same order)
    __shared__ float4 s_vtx[G == 1 ? 1 : 4][2][G == 1 ? 1 : MESH_LDS_VERTS];
    float4* va = s_vtx[G == 1 ? 0 : (threadIdx.x / WAVE) % 4][0];
    float4* vb = s_vtx[G == 1 ? 0 : (threadIdx.x / WAVE) % 4][1];
    const bool sa = G > 1 && A.kind == PK_MESH && A.nv <= MESH_LDS_VERTS;
    const bool sb = G > 1 && B.kind == PK_MESH && B.nv <= MESH_LDS_VERTS;
    if (G > 1 && ((sa && !keep_a) || sb)) {
        wave_lds_sync();
        if (sa && !keep_a) for (int i = L::rank(); i < A.nv; i += G) va[i] = A.v[i];
        if (sb) for (int i = L::rank(); i < B.nv; i += G) vb[i] = B.v[i];
        wave_lds_sync();
    }
    // testSepAxis, with a mesh's vertices from the wave's LDS copy when there is one
    auto test_axis = [&](const Poly& A, const Poly& B, const Pose& ta, const Pose& tb, f3 ax, float& depth, f3& wa,
                         f3& wb) {
        return test_axis_with(
            [&](f3 a, float& mn, float& mx, f3& pn, f3& px) {
                if (sa) project_verts(va, A.nv, ta, a, mn, mx, pn, px);
                else project(A, ta, a, mn, mx, pn, px);
            },
            [&](f3 a, float& mn, float& mx, f3& pn, f3& px) {
                if (sb) project_verts(vb, B.nv, tb, a, mn, mx, pn, px);
                else project(B, tb, a, mn, mx, pn, px);
            },
            ax, depth, wa, wb);
    };
    {
        float bd = REAL_MAX;
        int bi = 0x7fffffff;
        f3 bn{0, 0, 0}, bpt{0, 0, 0};
        bool separated = false;
        for (int i = L::rank(); i < nfa + nfb; i += G) {
            const bool on_b = i >= nfa;
            f3 n = on_b ? mv(tb.R, face_normal(B, i - nfa)) : mv(ta.R, face_normal(A, i));
            if (dot(dc, n) < 0) n = mul(n, -1.0f);
            float d;
            f3 wa, wb;
            if (!test_axis(A, B, ta, tb, n, d, wa, wb)) { separated = true; break; }
            if (d < bd) { bd = d; bi = i; bn = n; bpt = on_b ? add(wa, mul(n, d)) : wb; }
        }
        if (L::any(separated)) return;
        L::argmin(bd, bi);
        if (bi != 0x7fffffff) {
            const int src = bi & (G - 1);
            dmin = bd;
            sep = L::bcast(bn, src);
            dmin_pt = L::bcast(bpt, src);
            pt_on_a = bi >= nfa;
        }
    }
    if ((pt_on_a && inside(B, xf_inv(tb, dmin_pt))) || (!pt_on_a && inside(A, xf_inv(ta, dmin_pt))))
        add_point(c, x, sep, sub(dmin_pt, mul(sep, 0.0f)), -dmin + 0.0f * 2);
    int ea = -1, eb = -1;
    f3 wea{0, 0, 0}, web{0, 0, 0}, wpa{0, 0, 0}, wpb{0, 0, 0};
    const int nea = n_edges(A), neb = n_edges(B);
    if (nea > 0 && neb > 0) {
        float bd = dmin;
        int bi = 0x7fffffff;
        f3 bsep{0, 0, 0}, bwea{0, 0, 0}, bweb{0, 0, 0}, bwpa{0, 0, 0}, bwpb{0, 0, 0};
        bool separated = false;
        const int ne = nea * neb;
        // with a wave: B's world edge directions (the inner loop's operand), computed once into LDS
        __shared__ float4 s_edir[G == 1 ? 1 : 4][2][G == 1 ? 1 : EDGE_LDS_DIRS];
        float4* adir = s_edir[G == 1 ? 0 : (threadIdx.x / WAVE) % 4][0];
        float4* edir = s_edir[G == 1 ? 0 : (threadIdx.x / WAVE) % 4][1];
        const bool sea = G > 1 && nea <= EDGE_LDS_DIRS;
        const bool se = G > 1 && neb <= EDGE_LDS_DIRS;
        if (G > 1 && (sea || se)) {
            wave_lds_sync();
            if (sea) {
                for (int e = L::rank(); e < nea; e += G) {
                    f3 p0, p1;
                    edge_of(A, e, p0, p1);
                    const f3 d = mv(ta.R, normalized(sub(p0, p1)));
                    adir[e] = make_float4(d.x, d.y, d.z, 0.0f);
                }
            }
            if (se) {
                for (int e = L::rank(); e < neb; e += G) {
                    f3 b0, b1;
                    edge_of(B, e, b0, b1);
                    const f3 d = mv(tb.R, normalized(sub(b0, b1)));
                    edir[e] = make_float4(d.x, d.y, d.z, 0.0f);
                }
            }
            wave_lds_sync();
        }
        int t = L::rank();
        int e0 = t / neb, e1 = t - e0 * neb;
        int cur0 = -1;
        f3 a0, a1, we0;
        for (; t < ne; t += G) {
            if (e0 != cur0) {
                if (sea) {
                    we0 = ld4(adir[e0]);
                } else {
                    edge_of(A, e0, a0, a1);
                    we0 = mv(ta.R, normalized(sub(a0, a1)));
                }
                cur0 = e0;
            }
            f3 we1;
            if (se) {
                we1 = ld4(edir[e1]);
            } else {
                f3 b0, b1;
                edge_of(B, e1, b0, b1);
                we1 = mv(tb.R, normalized(sub(b0, b1)));
            }
            f3 cr = cross(we0, we1);
            if (!(fabsf(norm(cr) - 0.0f) < EPS)) {
                cr = normalized(cr);
                if (dot(dc, cr) < 0) cr = mul(cr, -1.0f);
                float dist;
                f3 wa, wb;
                if (!test_axis(A, B, ta, tb, cr, dist, wa, wb)) { separated = true; break; }
                if (dist < bd) {
                    if (sea) edge_of(A, e0, a0, a1);
                }
                if (dist < bd && on_line(a0, a1, xf_inv(ta, wa))) {
                    f3 b0, b1;
                    edge_of(B, e1, b0, b1);
                    if (on_line(b0, b1, xf_inv(tb, wb))) {
                        bd = dist; bi = t; bsep = cr; bwea = we0; bweb = we1; bwpa = wa; bwpb = wb;
                    }
                }
            }
            e1 += G;
            while (e1 >= neb) { e1 -= neb; e0++; }
        }
        if (L::any(separated)) return;
        L::argmin(bd, bi);
        if (bi != 0x7fffffff) {
            const int src = bi & (G - 1);
            dmin = bd;
            ea = bi / neb;
            eb = bi - ea * neb;
            sep = L::bcast(bsep, src);
            wea = L::bcast(bwea, src);
            web = L::bcast(bweb, src);
            wpa = L::bcast(bwpa, src);
            wpb = L::bcast(bwpb, src);
        }
    }
    if (ea >= 0 && eb >= 0) {
        // segmentsClosestPoints (:172-214) with half lengths PE_REAL_MAX
        const f3 tr = sub(wpb, wpa);
        const float ha = REAL_MAX, hb = REAL_MAX;
        const float dadb = dot(wea, web), dat = dot(wea, tr), dbt = dot(web, tr);
        const float d = 1.0f - dadb * dadb;
        float t_a, t_b;
        if (d == 0) t_a = 0;
        else {
            t_a = (dat - dbt * dadb) / d;
            if (t_a < -ha) t_a = -ha;
            else if (t_a > ha) t_a = ha;
        }
        t_b = t_a * dadb - dbt;
        if (t_b < -hb) {
            t_b = -hb;
            t_a = t_b * dadb + dat;
            if (t_a < -ha) t_a = -ha;
            else if (t_a > ha) t_a = ha;
        } else if (t_b > hb) {
            t_b = hb;
            t_a = t_b * dadb + dat;
            if (t_a < -ha) t_a = -ha;
            else if (t_a > ha) t_a = ha;
        }
        const f3 oa = mul(wea, t_a), ob = mul(web, t_b);
        f3 pv = add(sub(tr, oa), ob);
        const float nl2 = dot(pv, pv);
        if (nl2 > EPS) {
            const float nl = sqrtf(nl2);
            pv = mul(pv, 1.0f / nl);
            if (dot(pv, dc) < 0) pv = mul(pv, -1.0f);
            const f3 pob = add(wpb, ob);
            add_point(c, x, pv, sub(pob, mul(pv, 0.0f)), -nl + 0.0f * 2);
        }
    }
    if (dot(dc, sep) < 0) sep = neg(sep);
    clip_hull_hull<G>(c, x, sep, A, B, ta, tb, -ref_scale, 0.0f);
}

// ------------------------------------------------------------------ sphere-cylinder
// SphereCylinderCollisionAlgorithm::getClosestPoints (sphere_cylinder_collision_algorithm.cpp:32-77)
__device__ __noinline__ void sphere_cylinder(Cache& c, const PairCtx& x, const Pose& ps, float s_r, const Pose& pc,
                                             float c_r, float height) {
    const float c_h = height * 0.5f;
    const f3 sp = xf_inv(pc, ps.p);
    const float r = sqrtf(sp.x * sp.x + sp.z * sp.z);
    const float d = fabsf(sp.y) - c_h;
    if (sp.y < -c_h - s_r || sp.y > c_h + s_r || r * r > (c_r + s_r) * (c_r + s_r) ||
        (d > 0 && r > c_r && (r - c_r) * (r - c_r) + d * d > s_r * s_r))
        return;
    f3 n, pos;
    float depth;
    if (sp.y >= -c_h && sp.y <= c_h && r > 0) {
        n = normalized(f3{-sp.x, 0.0f, -sp.z});
        pos = add(mul(n, s_r), sp);
        depth = c_r + s_r - r;
    } else if (r <= c_r) {
        n = f3{0.0f, 1.0f, 0.0f};
        if (sp.y > 0) n = neg(n);
        pos = add(sp, mul(n, s_r));
        depth = s_r + c_h - fabsf(sp.y);
    } else {
        f3 q = mul(normalized(f3{sp.x, 0.0f, sp.z}), c_r);
        q.y = sp.y > 0 ? c_h : -c_h;
        n = normalized(sub(q, sp));
        pos = add(mul(n, s_r), sp);
        depth = s_r - norm(sub(sp, q));
    }
    n = mv(pc.R, n);
    pos = xf_apply(pc, pos);
    add_point(c, x, n, sub(pos, mul(n, 0.0f)), -depth + 2 * 0.0f);
}

// ------------------------------------------------------------------ box-cylinder
// box_cylinder_collision_algorithm.cpp
// addContactPointOnBox (:79-114)
PE_HD void point_on_box(Cache& c, const PairCtx& x, f3 pc, int face, f3 he, const Pose& to) {
    if (!(face >= -3 && face <= 3 && face != 0)) return;
    for (int i = 0; i < 3; i++)
        if (comp(pc, i) > comp(he, i) || comp(pc, i) < -comp(he, i)) return;
    f3 p = pc, n{0, 0, 0};
    float pen;
    if (face > 0) { const int i = face - 1; setc(p, i, comp(he, i)); setc(n, i, 1.0f); pen = comp(pc, i) - comp(he, i); }
    else { const int i = -face - 1; setc(p, i, -comp(he, i)); setc(n, i, -1.0f); pen = -comp(pc, i) - comp(he, i); }
    const f3 nw = mv(to.R, n);
    const f3 pw = xf_apply(to, p);
    add_point(c, x, nw, sub(pw, mul(nw, 0.0f)), pen + 2 * 0.0f);
}
// addContactPointOnCylinder (:116-139)
PE_HD void point_on_cyl(Cache& c, const PairCtx& x, f3 p, f3 pc, f3 ax, float rad, float h, const Pose& to) {
    const f3 r = sub(p, pc);
    const float dy = dot(r, ax);
    const float dxz = sqrtf(dot(r, r) - dy * dy);
    const float d_side = rad - dxz, d_top = h - fabsf(dy);
    f3 n;
    float dist;
    if (d_side < d_top) { dist = d_side; n = normalized(sub(mul(ax, dy), r)); }
    else { dist = d_top; n = dy < 0 ? ax : neg(ax); }
    n = mv(to.R, n);
    const f3 pw = xf_apply(to, p);
    add_point(c, x, n, sub(pw, mul(n, 0.0f)), -dist + 2 * 0.0f);
}
// findClosestBoxFace (:141-157)
PE_HD int closest_box_face(f3 he, f3 q) {
    int code = 0;
    float dist = REAL_MAX;
    for (int i = 0; i < 3; i++) {
        const float pd = fabsf(comp(q, i) - comp(he, i)), nd = fabsf(comp(q, i) + comp(he, i));
        if (pd < dist) { code = i + 1; dist = pd; }
        if (nd < dist) { code = -(i + 1); dist = nd; }
    }
    return code;
}
PE_HD float clampf(float v, float lo, float hi) { const float m = v < hi ? v : hi; return m > lo ? m : lo; }
// intersectSegmentBox (:171-231)
PE_HD bool seg_box(f3 he, f3 pc, f3 ax, float h, float tol, float& tmin, float& tmax) {
    tmin = REAL_MIN; tmax = REAL_MAX;
    for (int k = 0; k < 3; k++) {
        const float a = comp(ax, k), p = comp(pc, k), e = comp(he, k);
        if (fabsf(a) < tol) {
            if (fabsf(p) > e) return false;
        } else {
            const float t1 = (-e - p) / a, t2 = (+e - p) / a;
            const float lo = t1 < t2 ? t1 : t2, hi = t1 > t2 ? t1 : t2;
            tmin = tmin > lo ? tmin : lo;
            tmax = tmax < hi ? tmax : hi;
            if (tmin > tmax) return false;
        }
    }
    if ((tmin < -h && tmax < -h) || (tmin > +h && tmax > +h)) return false;
    tmin = clampf(tmin, -h, +h);
    tmax = clampf(tmax, -h, +h);
    return true;
}
// intersectLinePlane (:159-169)
PE_HD bool line_plane(f3 s, f3 d, f3 pp, f3 pn, float tol, float& t) {
    const float nd = dot(pn, d);
    if (fabsf(nd) < tol) return false;
    t = dot(pn, sub(pp, s)) / nd;
    return true;
}
// intersectSegmentCylinder (:233-304)
PE_HD bool seg_cyl(f3 s, f3 sd, float len, f3 pc, f3 ax, float h, float rad, float tol, float& tmin, float& tmax) {
    tmin = REAL_MIN; tmax = REAL_MAX;
    const f3 v = sub(s, pc);
    const float cd_sd = dot(ax, sd), v_cd = dot(v, ax), v_sd = dot(v, sd), v_v = dot(v, v);
    const float a = 1.0f - cd_sd * cd_sd;
    const float b = v_sd - v_cd * cd_sd;
    const float cc = v_v - v_cd * v_cd - rad * rad;
    if (fabsf(a) < tol) {
        const f3 q = sub(v, mul(ax, v_cd));
        if (dot(q, q) > rad * rad) return false;
        tmin = -len; tmax = +len;
    } else {
        float delta = b * b - a * cc;
        if (delta < 0) return false;
        delta = sqrtf(delta);
        tmin = (-b - delta) / a;
        tmax = (-b + delta) / a;
    }
    float t1, t2;
    const bool c1 = line_plane(s, sd, add(pc, mul(ax, h)), ax, tol, t1);
    const bool c2 = line_plane(s, sd, sub(pc, mul(ax, h)), ax, tol, t2);
    if (c1 && c2) {
        if (t1 < t2) { tmin = tmin > t1 ? tmin : t1; tmax = tmax < t2 ? tmax : t2; }
        else { tmin = tmin > t2 ? tmin : t2; tmax = tmax < t1 ? tmax : t1; }
        if (tmax < tmin) return false;
    } else {
        const float d1 = fabsf(dot(ax, sub(add(pc, mul(ax, h)), s)));
        const float d2 = fabsf(dot(ax, sub(sub(pc, mul(ax, h)), s)));
        if (d1 > 2 * h || d2 > 2 * h) return false;
    }
    if ((tmin < -len && tmax < -len) || (tmin > +len && tmax > +len)) return false;
    tmin = clampf(tmin, -len, +len);
    tmax = clampf(tmax, -len, +len);
    return true;
}
// BoxCylinderCollisionAlgorithm::getClosestPoints (:306-388)
__device__ __noinline__ void box_cylinder(Cache& c, const PairCtx& x, const Pose& pb, const float size[3],
                                          const Pose& pcy, float rad, float height) {
    const float h = height / 2.0f;
    const f3 he{size[0] / 2.0f, size[1] / 2.0f, size[2] / 2.0f};
    const Pose rel = pose_mul(pose_inv(pb), pcy);
    const f3 ax = col(rel.R, 1);
    const f3 pc = rel.p;
    for (int id = 0; id < 3; id++) {
        f3 nd{0, 0, 0};
        setc(nd, id, 1.0f);
        if (fabsf(comp(ax, id) - 1) < EPS || fabsf(comp(ax, id) + 1) < EPS) continue;
        const f3 r = mul(normalized(cross(cross(nd, ax), ax)), rad);
        for (int j = 0; j < 2; j++) {
            const float jd = j == 0 ? -1.0f : 1.0f;
            const f3 cs = add(pc, f3{jd * r.x, jd * r.y, jd * r.z});
            float tmin, tmax;
            if (seg_box(he, cs, ax, h, EPS, tmin, tmax)) {
                const f3 pmin = add(cs, mul(ax, tmin));
                const f3 pmax = add(cs, mul(ax, tmax));
                const f3 pmid = add(cs, mul(ax, (tmin + tmax) / 2));
                const int face = closest_box_face(he, pmid);
                point_on_box(c, x, pmin, face, he, pb);
                point_on_box(c, x, pmax, face, he, pb);
            }
        }
    }
    for (int id = 0; id < 3; id++) {
        f3 ed{0, 0, 0};
        setc(ed, id, 1.0f);
        const float eh = comp(he, id);
        const int jd = (id + 1) % 3, kd = (id + 2) % 3;
        for (int j = -1; j <= 1; j += 2) {
            for (int k = -1; k <= 1; k += 2) {
                f3 ec{0, 0, 0};
                setc(ec, id, 0.0f);
                setc(ec, jd, (float)j * comp(he, jd));
                setc(ec, kd, (float)k * comp(he, kd));
                float tmin, tmax;
                if (seg_cyl(ec, ed, eh, pc, ax, h, rad, EPS, tmin, tmax)) {
                    const f3 pmid = add(ec, mul(ed, (tmin + tmax) / 2));
                    point_on_cyl(c, x, pmid, pc, ax, rad, h, pb);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ BVH search
// BVH::search (bvh.cpp:105-125): depth first, right child popped before left; calls
// visit(face) for every face of every overlapped leaf, in the reference's output order.
// The stack lives in LDS (a private array would live in scratch memory, whose every push and
// pop is a round trip beyond L1): one row per lane when each lane walks its own pair (G = 1,
// kernels of at most BVH_ROWS threads per workgroup), one row per wave when the wave's lanes walk
// one pair together (G = WAVE, bvh_visit_wave: every lane writes the same value to the same word;
// kernels of at most 4 waves per workgroup).
constexpr int BVH_ROWS = 64;
__device__ __forceinline__ int* bvh_stack_row() {
    __shared__ int s_stack[BVH_ROWS * (BVH_STACK + 1)];   // rows padded to an odd length
    return s_stack + (int)(threadIdx.x % BVH_ROWS) * (BVH_STACK + 1);
}
// The walk of a wave (G = WAVE): a stack entry carries what its node's record and leaf entry hold
// (a leaf's face range; an inner node's children), read when the node was tested as its parent's
// child.  A node then costs one round of dependent loads -- its children's records, when it is an
// inner node -- rather than one for its own record and one more for its leaf entry, and a child
// that misses the box is never pushed.  Children are tested and pushed left then right, so the
// nodes that pass are popped, and the faces visited, in the same order as by the walk below.
__device__ __forceinline__ int4* bvh_wave_stack_row() {
    __shared__ int4 s_st[4][BVH_STACK];   // a row per wave (every lane writes the same value)
    return s_st[(threadIdx.x / WAVE) % 4];
}
template <class F>
__device__ inline void bvh_visit_wave(const float4* node, const int2* leaf, f3 aa, f3 bb, F&& visit) {
    int4* st = bvh_wave_stack_row();
    auto hit = [&](float4 lo, float4 hi) {
        return !(lo.x > bb.x || lo.y > bb.y || lo.z > bb.z) && !(hi.x < aa.x || hi.y < aa.y || hi.z < aa.z);
    };
    int sp = 0;
    {
        const float4 lo = node[0], hi = node[1];
        const int2 lf = leaf[0];
        if (hit(lo, hi)) st[sp++] = make_int4(lf.x, lf.y, __float_as_int(lo.w), __float_as_int(hi.w));
    }
    while (sp > 0) {
        const int4 e = st[--sp];
        if (e.x > 0) {
            for (int i = 0; i < e.x; i++) visit(e.y + i);
            continue;
        }
        const int cl = e.z, cr = e.w;
        float4 llo{}, lhi{}, rlo{}, rhi{};
        int2 llf{}, rlf{};
        if (cl >= 0) { llo = node[2 * cl]; lhi = node[2 * cl + 1]; llf = leaf[cl]; }
        if (cr >= 0) { rlo = node[2 * cr]; rhi = node[2 * cr + 1]; rlf = leaf[cr]; }
        if (cl >= 0 && hit(llo, lhi) && sp < BVH_STACK)
            st[sp++] = make_int4(llf.x, llf.y, __float_as_int(llo.w), __float_as_int(lhi.w));
        if (cr >= 0 && hit(rlo, rhi) && sp < BVH_STACK)
            st[sp++] = make_int4(rlf.x, rlf.y, __float_as_int(rlo.w), __float_as_int(rhi.w));
    }
}
template <int G = 1, class F>
__device__ inline void bvh_visit(const MeshSet& ms, int mesh, f3 aa, f3 bb, F&& visit) {
    const int32_t* r = ms.rec + MRW * mesh;
    if (r[MR_NB] == 0) return;
    const float4* node = ms.arena + r[MR_NODE];
    const int2* leaf = reinterpret_cast<const int2*>(ms.arena + r[MR_LEAF]);
    if (G > 1) {
        bvh_visit_wave(node, leaf, aa, bb, visit);
        return;
    }
    int* st = bvh_stack_row();
    int sp = 0;
    st[sp++] = 0;
    while (sp > 0) {
        const int id = st[--sp];
        if (id < 0) continue;
        const float4 lo = node[2 * id], hi = node[2 * id + 1];
        if (lo.x > bb.x || lo.y > bb.y || lo.z > bb.z) continue;
        if (hi.x < aa.x || hi.y < aa.y || hi.z < aa.z) continue;
        const int2 lf = leaf[id];
        if (lf.x > 0) {
            for (int i = 0; i < lf.x; i++) visit(lf.y + i);
        } else if (sp + 2 <= BVH_STACK) {
            st[sp++] = __float_as_int(lo.w);
            st[sp++] = __float_as_int(hi.w);
        }
    }
}

// The faces bvh_visit<WAVE> finds, tested by a wave's lanes in batches of WAVE and added in the
// walk's order.  test(f), on the lane that holds face f, returns the bit mask of f's items (the
// routine's own units: triangles of a fan, segments, edges) that give contact points; commit(f, i)
// then runs on every lane with the same values for each such item, face by face in walk order and
// item by item in ascending order -- the order in which a sequential walk adds the points, so the
// contact cache ends as it would.  Most faces yield no point, so the tests (most of the work) run
// in parallel and only the hits are computed again, serially, to be added.
// With rec, the walk's faces are also kept in rec (a wave's LDS row: every lane writes the same
// value) while there are at most FACE_REC of them, and *nrec is their count: a second walk over the
// same box then runs over the list (faces_list_wave) instead of the tree.
constexpr int FACE_REC = 128;
__device__ __forceinline__ int* face_rec_row() {
    __shared__ int s_rec[4][FACE_REC];
    return s_rec[(threadIdx.x / WAVE) % 4];
}
template <class Test, class Commit>
__device__ inline void bvh_faces_wave(const MeshSet& ms, int mesh, f3 aa, f3 bb, Test&& test, Commit&& commit,
                                      int* rec = nullptr, int* nrec = nullptr) {
    const int r = (int)(threadIdx.x & (WAVE - 1));
    int mine = -1, cnt = 0, all = 0;
    auto flush = [&]() {
        const uint32_t m = r < cnt ? test(mine) : 0u;
        uint64_t hit = __ballot(m != 0u);
        while (hit) {
            const int src = __ffsll((unsigned long long)hit) - 1;
            hit &= hit - 1;
            const int f = __shfl(mine, src);
            uint32_t mk = (uint32_t)__shfl((int)m, src);
            while (mk) {
                const int i = __ffs(mk) - 1;
                mk &= mk - 1;
                commit(f, i);
            }
        }
        cnt = 0;
    };
    bvh_visit<WAVE>(ms, mesh, aa, bb, [&](int f) {
        if (rec && all < FACE_REC) rec[all] = f;
        all++;
        if (r == cnt) mine = f;
        if (++cnt == WAVE) flush();
    });
    if (cnt > 0) flush();
    if (nrec) *nrec = all;
}
// The same tests and commits over faces already found: rec[0 .. n), in order.
template <class Test, class Commit>
__device__ inline void faces_list_wave(const int* rec, int n, Test&& test, Commit&& commit) {
    const int r = (int)(threadIdx.x & (WAVE - 1));
    for (int base = 0; base < n; base += WAVE) {
        const int cnt = n - base < WAVE ? n - base : WAVE;
        const int mine = r < cnt ? rec[base + r] : -1;
        const uint32_t m = r < cnt ? test(mine) : 0u;
        uint64_t hit = __ballot(m != 0u);
        while (hit) {
            const int src = __ffsll((unsigned long long)hit) - 1;
            hit &= hit - 1;
            const int f = __shfl(mine, src);
            uint32_t mk = (uint32_t)__shfl((int)m, src);
            while (mk) {
                const int i = __ffs(mk) - 1;
                mk &= mk - 1;
                commit(f, i);
            }
        }
    }
}

// ------------------------------------------------------------------ cylinder-mesh
// intersectSegmentFace (cylinder_convex_collision_algorithm.cpp:53-107)
PE_HD bool seg_face(const Poly& M, int f, f3 ps, f3 sd, float l, float& t1, float& t2, float& d1, float& d2) {
    const f3 fn = face_normal(M, f);
    const f3 org = face_vertex(M, f, 0);
    const f3 end = add(ps, mul(sd, l));
    const f3 r1 = sub(ps, org), r2 = sub(end, org);
    const float dist1 = dot(r1, fn), dist2 = dot(r2, fn);
    if ((dist1 > 0 && dist2 > 0) || (dist1 <= 0 && dist2 <= 0)) return false;
    const float t = dist1 / (dist1 - dist2) * l;
    const f3 v = add(ps, mul(sd, t));
    float tmin = 0, tmax = l, mind = REAL_MAX;
    const int nf = face_len(M, f);
    for (int i = 0; i < nf; i++) {
        const f3 v0 = face_vertex(M, f, i);
        const f3 v1 = face_vertex(M, f, (i + 1) % nf);
        const f3 tc = normalized(cross(fn, sub(v1, v0)));
        const float ddc = dot(sd, tc);
        const float ed = fabsf(dot(sub(v0, v), tc));
        mind = mind < ed ? mind : ed;
        if (!(fabsf(ddc - 0.0f) < EPS)) {
            const float q = dot(sub(v0, ps), tc) / ddc;
            if (ddc > 0) tmin = tmin > q ? tmin : q;
            else tmax = tmax < q ? tmax : q;
        } else if (dot(cross(sub(v0, v), sub(v1, v)), fn) < 0) {
            return false;
        }
    }
    if (t < tmin || t > tmax) return false;
    t1 = dist1 <= 0 ? tmin : t;
    t2 = dist2 <= 0 ? tmax : t;
    d1 = dist1 <= 0 ? dist1 * (t - tmin) / t : 0;
    d2 = dist2 <= 0 ? dist2 * (tmax - t) / (l - t) : 0;
    if (-d1 > mind + 0.0f || -d2 > mind + 0.0f) return false;
    return true;
}
// CylinderConvexCollisionAlgorithm::getClosestPoints (:109-203), split into its items.  The
// cylinder in the mesh's frame: axis ax, centre pc, half height h.
struct CylInMesh {
    f3 ax, pc, aa, bb;
    float rad, h;
};
__device__ __forceinline__ CylInMesh cyl_in_mesh(const Pose& pm, const Pose& pcy, float rad, float height) {
    CylInMesh q;
    q.h = height / 2.0f;
    q.rad = rad;
    const Pose rel = pose_mul(pose_inv(pm), pcy);
    q.ax = col(rel.R, 1);
    q.pc = rel.p;
    const float sp[4] = {rad, height, 0.0f, 0.0f};
    shape_aabb(SHAPE_CYLINDER, sp, rel.R, rel.p, q.aa, q.bb);
    return q;
}
// First walk, face f: the cylinder's section in the plane of the axis and the face normal (ux
// across, the axis up; uz the second section when the axis is along the normal).  Item i < 4 is
// side i of the section, items 4 and 5 the sides of the second section along the axis.
struct CylFace {
    f3 ux, uz;
    bool vert;
};
PE_HD CylFace cyl_face_frame(const Poly& M, int f, f3 ax) {
    const f3 fn = face_normal(M, f);
    const float adc = dot(ax, fn);
    CylFace F;
    F.uz = f3{0, 0, 0};
    F.vert = false;
    if (fabsf(fabsf(adc) - 1) < EPS) {
        F.uz = normalized(cross(ax, f3{1.0f, 0.0f, 0.0f}));
        F.ux = normalized(cross(ax, F.uz));
        F.vert = true;
    } else {
        F.ux = normalized(cross(cross(fn, ax), ax));
    }
    return F;
}
PE_HD int cyl_face_items(const CylFace& F) { return F.vert ? 6 : 4; }
PE_HD void cyl_face_seg(const CylFace& F, const CylInMesh& q, int i, f3& ps, f3& sd, float& l) {
    constexpr int dx[4] = {1, -1, -1, 1};
    constexpr int dy[4] = {1, 1, -1, -1};
    const f3 uy = q.ax;
    if (i < 4) {
        ps = add(add(q.pc, mul(mul(F.ux, (float)dx[i]), q.rad)), mul(mul(uy, (float)dy[i]), q.h));
        sd = mul(dx[i] == dy[i] ? uy : F.ux, (float)-dx[i]);
        l = (dx[i] == dy[i] ? q.h : q.rad) * 2;
    } else {
        const int j = (i - 4) * 2;
        ps = add(add(q.pc, mul(mul(F.uz, (float)dx[j]), q.rad)), mul(mul(uy, (float)dy[j]), q.h));
        sd = mul(uy, (float)-dx[j]);
        l = q.h * 2;
    }
}
PE_HD bool cyl_face_hit(const Poly& M, int f, const CylFace& F, const CylInMesh& q, int i) {
    f3 ps, sd;
    float l, t1, t2, d1, d2;
    cyl_face_seg(F, q, i, ps, sd, l);
    return seg_face(M, f, ps, sd, l, t1, t2, d1, d2);
}
PE_HD void cyl_face_add(Cache& c, const PairCtx& x, const Poly& M, int f, const CylFace& F, const CylInMesh& q,
                        int i, const Pose& pm) {
    f3 ps, sd;
    float l, t1, t2, d1, d2;
    cyl_face_seg(F, q, i, ps, sd, l);
    if (seg_face(M, f, ps, sd, l, t1, t2, d1, d2)) {
        const f3 fn = face_normal(M, f);
        const f3 p1 = xf_apply(pm, sub(add(ps, mul(sd, t1)), mul(fn, d1)));
        const f3 p2 = xf_apply(pm, sub(add(ps, mul(sd, t2)), mul(fn, d2)));
        const f3 n = mv(pm.R, fn);
        add_point(c, x, n, sub(p1, mul(n, 0.0f)), d1 + 0.0f * 2);
        add_point(c, x, n, sub(p2, mul(n, 0.0f)), d2 + 0.0f * 2);
    }
}
// Second walk, face f: item i is the face's edge from vertex i to vertex i + 1.
PE_HD bool cyl_edge_seg(const Poly& M, int f, int i, const CylInMesh& q, f3& ss, f3& sd, float& t1, float& t2) {
    const int nf = face_len(M, f);
    const f3 v1 = face_vertex(M, f, i);
    const f3 v2 = face_vertex(M, f, (i + 1) % nf);
    ss = divs(add(v1, v2), 2.0f);
    const float l = norm(sub(v2, v1)) / 2.0f;
    sd = divs(sub(v2, v1), l * 2);
    return seg_cyl(ss, sd, l, q.pc, q.ax, q.h, q.rad, 0.0f, t1, t2);
}
PE_HD void cyl_edge_add(Cache& c, const PairCtx& x, const Poly& M, int f, int i, const CylInMesh& q, const Pose& pm) {
    f3 ss, sd;
    float t1, t2;
    if (cyl_edge_seg(M, f, i, q, ss, sd, t1, t2)) {
        const f3 p1 = add(ss, mul(sd, t1));
        const f3 p2 = add(ss, mul(sd, t2));
        const f3 pmid = divs(add(p1, p2), 2.0f);
        point_on_cyl(c, x, pmid, q.pc, q.ax, q.rad, q.h, pm);
    }
}
// The whole routine for the mesh of a convex or concave body (pm: mesh body pose): G = 1 on one
// lane; G = WAVE on a wave's lanes, the faces of each walk split over them (bvh_faces_wave).
template <int G = 1>
__device__ __noinline__ void cylinder_mesh(Cache& c, const PairCtx& x, const MeshSet& ms, int mesh, const Pose& pm,
                                           const Pose& pcy, float rad, float height) {
    const Poly M = poly_mesh(&ms, mesh);
    const CylInMesh q = cyl_in_mesh(pm, pcy, rad, height);
    if (G == 1) {
        bvh_visit(ms, mesh, q.aa, q.bb, [&](int f) {
            const CylFace F = cyl_face_frame(M, f, q.ax);
            for (int i = 0; i < cyl_face_items(F); i++) cyl_face_add(c, x, M, f, F, q, i, pm);
        });
        bvh_visit(ms, mesh, q.aa, q.bb, [&](int f) {
            const int nf = face_len(M, f);
            for (int i = 0; i < nf; i++) cyl_edge_add(c, x, M, f, i, q, pm);
        });
    } else {
        // the second walk is over the same box: it visits the faces the first one found, in the same
        // order, so it runs over their list when the first walk could keep it
        int* rec = face_rec_row();
        int nrec = 0;
        bvh_faces_wave(ms, mesh, q.aa, q.bb,
            [&](int f) {
                const CylFace F = cyl_face_frame(M, f, q.ax);
                uint32_t m = 0;
                for (int i = 0; i < cyl_face_items(F); i++) m |= (cyl_face_hit(M, f, F, q, i) ? 1u : 0u) << i;
                return m;
            },
            [&](int f, int i) { cyl_face_add(c, x, M, f, cyl_face_frame(M, f, q.ax), q, i, pm); }, rec, &nrec);
        auto edge_test = [&](int f) {
            const int nf = face_len(M, f);
            uint32_t m = 0;
            for (int i = 0; i < nf; i++) {
                f3 ss, sd;
                float t1, t2;
                m |= (cyl_edge_seg(M, f, i, q, ss, sd, t1, t2) ? 1u : 0u) << i;
            }
            return m;
        };
        auto edge_commit = [&](int f, int i) { cyl_edge_add(c, x, M, f, i, q, pm); };
        if (nrec <= FACE_REC) faces_list_wave(rec, nrec, edge_test, edge_commit);
        else bvh_faces_wave(ms, mesh, q.aa, q.bb, edge_test, edge_commit);
    }
}

// ------------------------------------------------------------------ sphere-mesh
// sphere_convex_collision_algorithm.cpp
// segmentSqrDistance (:36-57)
PE_HD float seg_sqdist(f3 from, f3 to, f3 p, f3& nearest) {
    f3 diff = sub(p, from);
    const f3 v = sub(to, from);
    float t = dot(v, diff);
    if (t > 0) {
        const float vv = dot(v, v);
        if (t < vv) { t /= vv; diff = sub(diff, f3{t * v.x, t * v.y, t * v.z}); }
        else { t = 1; diff = sub(diff, v); }
    } else {
        t = 0;
    }
    nearest = add(from, f3{t * v.x, t * v.y, t * v.z});
    return dot(diff, diff);
}
// pointInTriangle (:59-83)
PE_HD bool point_in_tri(const f3* vt, f3 n, f3 p) {
    const f3 e1 = sub(vt[1], vt[0]), e2 = sub(vt[2], vt[1]), e3 = sub(vt[0], vt[2]);
    const float r1 = dot(cross(e1, n), sub(p, vt[0]));
    const float r2 = dot(cross(e2, n), sub(p, vt[1]));
    const float r3 = dot(cross(e3, n), sub(p, vt[2]));
    return (r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0);
}
// collideSphereTriangle (:91-163)
PE_HD bool sphere_tri(f3 cc, float rad, const f3* vt, f3& point, f3& rn, float& depth) {
    f3 n = cross(sub(vt[1], vt[0]), sub(vt[2], vt[0]));
    const float l2 = dot(n, n);
    bool has = false;
    f3 cp{0, 0, 0};
    if (l2 >= EPS * EPS) {
        n = divs(n, sqrtf(l2));
        float dist = dot(sub(cc, vt[0]), n);
        if (dist < 0) { dist *= -1.0f; n = mul(n, -1.0f); }
        if (dist < rad) {
            if (point_in_tri(vt, n, cc)) {
                has = true;
                cp = sub(cc, mul(n, dist));
            } else {
                float mind = rad * rad;
                for (int i = 0; i < 3; i++) {
                    f3 ne;
                    const float d2 = seg_sqdist(vt[i], vt[(i + 1) % 3], cc, ne);
                    if (d2 < mind) { mind = d2; has = true; cp = ne; }
                }
            }
        }
    }
    if (has) {
        const f3 q = sub(cc, cp);
        const float d2 = dot(q, q);
        if (d2 < rad * rad) {
            if (d2 > EPS * EPS) {
                const float d = sqrtf(d2);
                rn = normalized(q);
                point = cp;
                depth = -(rad - d);
            } else {
                rn = n;
                point = cp;
                depth = -rad;
            }
            return true;
        }
    }
    return false;
}
// SphereConvexCollisionAlgorithm::getClosestPoints (:165-211) for the mesh of a convex or
// concave body: every face of every BVH leaf overlapping the sphere's box, fan-triangulated; item
// i of face f is the fan's triangle (0, i + 1, i + 2).  G as for cylinder_mesh.
PE_HD bool sphere_fan_tri(const Poly& M, int f, int i, f3 cc, float rad, f3& p, f3& n, float& depth) {
    const f3 vt[3] = {face_vertex(M, f, 0), face_vertex(M, f, i + 1), face_vertex(M, f, i + 2)};
    depth = 0;
    return sphere_tri(cc, rad, vt, p, n, depth);
}
PE_HD void sphere_fan_add(Cache& c, const PairCtx& x, const Poly& M, int f, int i, f3 cc, float rad, const Pose& pm) {
    f3 p, n;
    float depth;
    if (sphere_fan_tri(M, f, i, cc, rad, p, n, depth)) {
        n = mv(pm.R, n);
        p = xf_apply(pm, p);
        add_point(c, x, n, sub(p, mul(n, 0.0f)), depth + 2 * 0.0f);
    }
}
template <int G = 1>
__device__ __noinline__ void sphere_mesh(Cache& c, const PairCtx& x, const MeshSet& ms, int mesh, const Pose& pm,
                                         const Pose& psph, float rad) {
    const Poly M = poly_mesh(&ms, mesh);
    const f3 cc = xf_inv(pm, psph.p);
    const f3 aa = sub(cc, f3{rad, rad, rad}), bb = add(cc, f3{rad, rad, rad});
    if (G == 1) {
        bvh_visit(ms, mesh, aa, bb, [&](int f) {
            const int nf = face_len(M, f);
            for (int i = 0; i < nf - 2; i++) sphere_fan_add(c, x, M, f, i, cc, rad, pm);
        });
    } else {
        bvh_faces_wave(ms, mesh, aa, bb,
            [&](int f) {
                const int nf = face_len(M, f);
                uint32_t m = 0;
                for (int i = 0; i < nf - 2; i++) {
                    f3 p, n;
                    float depth;
                    m |= (sphere_fan_tri(M, f, i, cc, rad, p, n, depth) ? 1u : 0u) << i;
                }
                return m;
            },
            [&](int f, int i) { sphere_fan_add(c, x, M, f, i, cc, rad, pm); });
    }
}

// ------------------------------------------------------------------ concave mesh + polyhedron
// ConcaveConvexCollisionAlgorithm::getClosestPoints (concave_convex_collision_algorithm.cpp:38-68):
// each triangle of a BVH leaf overlapping the polyhedron's box (in the mesh frame) is a one-face
// shape without unique edges, run through the convex routine with the polyhedron as A.
template <int G>
__device__ __noinline__ void concave_poly(Cache& c, const PairCtx& x, const MeshSet& ms, int mesh, const Pose& pm,
                                          const Poly& A, const Pose& pa, float ref_scale) {
    const Pose rel = pose_mul(pose_inv(pm), pa);
    f3 aa, bb;
    if (A.kind == PK_BOX) {
        const float sp[4] = {A.h.x * 2.0f, A.h.y * 2.0f, A.h.z * 2.0f, 0.0f};
        shape_aabb(SHAPE_BOX, sp, rel.R, rel.p, aa, bb);
    } else {
        aa = f3{REAL_MAX, REAL_MAX, REAL_MAX};
        bb = f3{REAL_MIN, REAL_MIN, REAL_MIN};
        for (int i = 0; i < A.nv; i++) {
            const f3 v = mv(rel.R, ld4(A.v[i]));
            aa.x = aa.x < v.x ? aa.x : v.x; aa.y = aa.y < v.y ? aa.y : v.y; aa.z = aa.z < v.z ? aa.z : v.z;
            bb.x = bb.x > v.x ? bb.x : v.x; bb.y = bb.y > v.y ? bb.y : v.y; bb.z = bb.z > v.z ? bb.z : v.z;
        }
        aa = add(aa, rel.p);
        bb = add(bb, rel.p);
    }
    const Poly M = poly_mesh(&ms, mesh);
    bool staged = false;   // a wave copies A's vertices to LDS once per pair (convex_convex's keep_a)
    bvh_visit<G>(ms, mesh, aa, bb, [&](int f) {
        Poly T;
        T.kind = PK_TRI;
        T.ms = &ms;
        T.t[0] = face_vertex(M, f, 0);
        T.t[1] = face_vertex(M, f, 1);
        T.t[2] = face_vertex(M, f, 2);
        T.tn = face_normal(M, f);
        convex_convex<G>(c, x, A, T, pa, pm, ref_scale, staged);
        staged = G > 1;
    });
}

}  // namespace pe
