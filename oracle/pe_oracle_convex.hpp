// CPU ORACLE -- test infrastructure only (see pe_oracle.hpp).
//
// Restatement of the reference's polyhedral shapes and of the narrowphase algorithms that use
// them: box / cylinder / convex-mesh / triangle projections, the SAT + clipping convex routine,
// sphere-cylinder, box-cylinder, cylinder-mesh, sphere-mesh and the concave (triangle-mesh)
// wrappers, the mesh set-up (ConvexMeshShape::setMesh, ConcaveMeshShape::setMesh) and the BVH.
// Included by pe_oracle.hpp after the sphere pairs; every function cites the reference code
// it follows.
#pragma once

namespace pe_oracle {

// ---------------------------------------------------------------- polyhedral data
// A polyhedral shape as the reference stores it: pe::Mesh (vertices, faces with normals),
// the unique edge list and the BVH over faces (utils/bvh.h).
template <class R> struct Poly {
    std::vector<V3<R>> v;                        // vertex positions
    std::vector<std::vector<int>> f;             // face vertex indices
    std::vector<V3<R>> fn;                       // face normals
    std::vector<std::pair<V3<R>, V3<R>>> e;      // unique edges
    struct Node { V3<R> aa, bb; int left = -1, right = -1, n = 0, index = 0; };
    std::vector<Node> bvh;                       // root = node 0 (empty: no BVH)
    R volume = 0;
    M3<R> inertia = M3<R>::identity();           // unit-mass local inertia
};

// The unit box mesh and its unique edges (shape/default_mesh.h:7-57): 24 vertices, 4 per face
// in the order +y, -y, +z, -z, -x, +x, each face listing its corners counter-clockwise.
namespace unit_box {
static const signed char V[24][3] = {
    {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1},       // +y
    {-1, -1, 1}, {-1, -1, -1}, {1, -1, -1}, {1, -1, 1},   // -y
    {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},       // +z
    {1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {-1, 1, -1},   // -z
    {-1, 1, -1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1},   // -x
    {1, 1, 1}, {1, -1, 1}, {1, -1, -1}, {1, 1, -1}};      // +x
static const signed char N[6][3] = {{0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {-1, 0, 0}, {1, 0, 0}};
static const signed char E[12][2][3] = {
    {{-1, 1, 1}, {1, 1, 1}}, {{-1, -1, 1}, {1, -1, 1}}, {{-1, -1, -1}, {1, -1, -1}}, {{-1, 1, -1}, {1, 1, -1}},
    {{1, -1, 1}, {1, 1, 1}}, {{-1, -1, 1}, {-1, 1, 1}}, {{-1, -1, -1}, {-1, 1, -1}}, {{1, -1, -1}, {1, 1, -1}},
    {{1, 1, 1}, {1, 1, -1}}, {{-1, 1, 1}, {-1, 1, -1}}, {{-1, -1, 1}, {-1, -1, -1}}, {{1, -1, 1}, {1, -1, -1}}};
}  // namespace unit_box

// BoxShape::BoxShape (box_shape.cpp:7-24): unit mesh and edges scaled by the full size.
template <class R> inline Poly<R> box_poly(const Body<R>& b) {
    Poly<R> p;
    const R s[3] = {b.size[0], b.size[1], b.size[2]};
    auto sc = [&](const signed char* u) { return V3<R>(R(u[0] * 0.5) * s[0], R(u[1] * 0.5) * s[1], R(u[2] * 0.5) * s[2]); };
    for (int i = 0; i < 24; i++) p.v.push_back(sc(unit_box::V[i]));
    for (int f = 0; f < 6; f++) {
        p.f.push_back({4 * f, 4 * f + 1, 4 * f + 2, 4 * f + 3});
        p.fn.push_back(V3<R>(R(unit_box::N[f][0]), R(unit_box::N[f][1]), R(unit_box::N[f][2])));
    }
    for (int k = 0; k < 12; k++) p.e.push_back({sc(unit_box::E[k][0]), sc(unit_box::E[k][1])});
    return p;
}

// The unit cylinder mesh (shape/default_mesh.h:59-273): a 20-gon prism of radius 0.5 and
// height 1 whose bottom ring is turned by 9 degrees.  Vertex coordinates are 0.5 cos / sin of
// the ring angles rounded to 6 decimals (ring angle of top vertex i: -18 i degrees, bottom
// vertex i: -(9 + 18 i) degrees).  Side-face normals are the reference's 6-decimal values,
// tabulated by the face's ring angle; face 8 of the table carries x = -0.493982 where symmetry
// gives -0.453982, and it is kept.
namespace unit_cyl {
inline double r6(double x) { return std::round(x * 1e6) / 1e6 + 0.0; }
inline void ring(int i, bool top, double& x, double& z) {
    const double a = (top ? 18.0 * i : 9.0 + 18.0 * i) * 3.14159265358979323846 / 180.0;
    x = r6(0.5 * std::cos(a));
    z = r6(-0.5 * std::sin(a));
}
// |cos|, |sin| of the side-face normals: up faces (ring angle 9 + 18k), down faces (18k)
static const double UPC[5] = {0.98767, 0.89099, 0.707093, 0.453982, 0.156432};
static const double DNC[6] = {0.999981, 0.951038, 0.809002, 0.587774, 0.309011, 0.0};
inline void side_normal(bool up, int k, double n[3]) {
    // quadrant-reduced magnitudes, signs from the angle
    const double ang = up ? 9.0 + 18.0 * k : 18.0 * k;
    const double c = std::cos(ang * 3.14159265358979323846 / 180.0), s = std::sin(ang * 3.14159265358979323846 / 180.0);
    double ra = std::fmod(ang, 180.0);
    if (ra > 90.0) ra = 180.0 - ra;
    double mc, ms;
    if (up) { const int q = (int)std::lround((ra - 9.0) / 18.0); mc = UPC[q]; ms = UPC[4 - q]; }
    else { const int q = (int)std::lround(ra / 18.0); mc = DNC[q]; ms = DNC[5 - q]; }
    n[0] = std::fabs(c) < 1e-9 ? 0.0 : (c > 0 ? mc : -mc);
    n[1] = up ? 0.006156 : -0.006156;
    n[2] = std::fabs(s) < 1e-9 ? 0.0 : (s > 0 ? -ms : ms);
    if (up && k == 6) n[0] = -0.493982;
}
}  // namespace unit_cyl

// CylinderShape::CylinderShape (cylinder_shape.cpp:6-32): unit mesh and edges scaled by
// (2r, h, 2r); face normals are not rescaled.
template <class R> inline Poly<R> cylinder_poly(const Body<R>& b) {
    Poly<R> p;
    const R r2 = b.radius * 2, h = b.height;
    V3<R> top[20], bot[20];
    for (int i = 0; i < 20; i++) {
        double x, z;
        unit_cyl::ring(i, true, x, z);
        top[i] = V3<R>(R(x) * r2, R(0.5) * h, R(z) * r2);
        unit_cyl::ring(i, false, x, z);
        bot[i] = V3<R>(R(x) * r2, R(-0.5) * h, R(z) * r2);
    }
    for (int i = 0; i < 20; i++) p.v.push_back(top[i]);   // 0..19  top cap
    for (int i = 0; i < 20; i++) p.v.push_back(bot[i]);   // 20..39 bottom cap
    for (int i = 0; i < 20; i++) p.v.push_back(top[i]);   // 40..59 side, top ring
    for (int i = 0; i < 20; i++) p.v.push_back(bot[i]);   // 60..79 side, bottom ring
    std::vector<int> cap;
    for (int i = 0; i < 20; i++) cap.push_back(i);
    p.f.push_back(cap); p.fn.push_back(V3<R>(0, 1, 0));
    cap.clear();
    for (int i = 39; i >= 20; i--) cap.push_back(i);
    p.f.push_back(cap); p.fn.push_back(V3<R>(0, -1, 0));
    for (int k = 0; k < 20; k++) {
        p.f.push_back({60 + k, 40 + (k + 1) % 20, 40 + k});
        double n[3];
        unit_cyl::side_normal(true, k, n);
        p.fn.push_back(V3<R>(R(n[0]), R(n[1]), R(n[2])));
    }
    for (int k = 0; k < 20; k++) {
        p.f.push_back({40 + k, 60 + (k + 19) % 20, 60 + k});
        double n[3];
        unit_cyl::side_normal(false, k, n);
        p.fn.push_back(V3<R>(R(n[0]), R(n[1]), R(n[2])));
    }
    for (int k = 0; k < 20; k++) p.e.push_back({top[k], top[(k + 1) % 20]});
    for (int k = 0; k < 20; k++) p.e.push_back({bot[19 - k], bot[(38 - k) % 20]});
    for (int k = 0; k < 20; k++) {
        p.e.push_back({top[k], bot[k]});
        p.e.push_back({bot[k], top[(k + 1) % 20]});
    }
    return p;
}

// ---------------------------------------------------------------- BVH (utils/bvh.cpp)
template <class R> inline void face_min_max(const Poly<R>& p, const std::vector<int>& f, V3<R>& mn, V3<R>& mx) {  // :11-31
    mn = V3<R>(Lim<R>::max(), Lim<R>::max(), Lim<R>::max());
    mx = V3<R>(Lim<R>::min(), Lim<R>::min(), Lim<R>::min());
    for (int i : f)
        for (int k = 0; k < 3; k++) {
            if (p.v[i][k] < mn[k]) mn[k] = p.v[i][k];
            if (p.v[i][k] > mx[k]) mx[k] = p.v[i][k];
        }
}
// buildBVH (:45-83).  Faces are reordered with std::sort over the face centroid along the
// node's longest axis; the permutation depends only on the comparison results, so the same
// standard library gives the reference's order.
template <class R> inline int bvh_build(Poly<R>& p, int l, int r, int nmax) {
    if (l > r) return -1;
    const int id = (int)p.bvh.size();
    p.bvh.emplace_back();
    V3<R> aa(Lim<R>::max(), Lim<R>::max(), Lim<R>::max()), bb(Lim<R>::min(), Lim<R>::min(), Lim<R>::min());
    for (int i = l; i <= r; i++) {
        V3<R> mn, mx;
        face_min_max(p, p.f[i], mn, mx);
        for (int k = 0; k < 3; k++) {
            if (mn[k] < aa[k]) aa[k] = mn[k];
            if (mx[k] > bb[k]) bb[k] = mx[k];
        }
    }
    p.bvh[id].aa = aa; p.bvh[id].bb = bb;
    if (r - l + 1 <= nmax) { p.bvh[id].n = r - l + 1; p.bvh[id].index = l; return id; }
    const R lx = bb[0] - aa[0], ly = bb[1] - aa[1], lz = bb[2] - aa[2];
    const int axis = (lx >= ly && lx >= lz) ? 0 : (ly >= lz ? 1 : 2);
    struct F { std::vector<int> idx; V3<R> n; };
    std::vector<F> fs;
    for (int i = l; i <= r; i++) fs.push_back({p.f[i], p.fn[i]});
    auto centre = [&](const F& f) {
        R s = 0;
        for (int i : f.idx) s += p.v[i][axis];
        s /= (R)f.idx.size();
        return s;
    };
    std::sort(fs.begin(), fs.end(), [&](const F& a, const F& b) { return centre(a) < centre(b); });
    for (int i = l; i <= r; i++) { p.f[i] = fs[i - l].idx; p.fn[i] = fs[i - l].n; }
    const int mid = (l + r) / 2;
    const int left = bvh_build(p, l, mid, nmax);
    const int right = bvh_build(p, mid + 1, r, nmax);
    p.bvh[id].left = left; p.bvh[id].right = right;
    return id;
}
// BVH::search (:105-125): depth first, right child popped before left.
template <class R> inline void bvh_search(const Poly<R>& p, const V3<R>& aa, const V3<R>& bb, std::vector<int>& out) {
    if (p.bvh.empty()) return;
    std::vector<int> st{0};
    while (!st.empty()) {
        const int id = st.back();
        st.pop_back();
        if (id < 0) continue;
        const auto& nd = p.bvh[id];
        if (nd.aa[0] > bb[0] || nd.aa[1] > bb[1] || nd.aa[2] > bb[2]) continue;
        if (nd.bb[0] < aa[0] || nd.bb[1] < aa[1] || nd.bb[2] < aa[2]) continue;
        if (nd.n > 0) {
            for (int i = 0; i < nd.n; i++) out.push_back(nd.index + i);
        } else {
            st.push_back(nd.left);
            st.push_back(nd.right);
        }
    }
}

// ---------------------------------------------------------------- mesh set-up
// pe::Vector3Hash (phys_general.h:133-143) on the value the reference hashes.
template <class R> inline uint32_t vec_hash(const V3<R>& v) {
    auto ui = [](R c) { return (uint32_t)(int64_t)((double)c * 1e5); };
    uint32_t h = 0x995af;
    h ^= ui(v[0]) + 0x9e3779b9u + (h << 6) + (h >> 2);
    h ^= ui(v[1]) + 0x9e3779b9u + (h << 6) + (h >> 2);
    h ^= ui(v[2]) + 0x9e3779b9u + (h << 6) + (h >> 2);
    return h;
}
template <class R> inline bool approx_eq(const V3<R>& a, const V3<R>& b) {
    const R eps = R(0.00001);
    return std::abs(a[0] - b[0]) < eps && std::abs(a[1] - b[1]) < eps && std::abs(a[2] - b[2]) < eps;
}
// unique edges (convex_mesh_shape.cpp:110-128): an edge is kept at its first appearance in
// face order; vertices are identified through a hash list with approximate equality.
template <class R> inline void unique_edges(Poly<R>& p) {
    p.e.clear();
    std::vector<V3<R>> vm;
    std::vector<uint32_t> vh;
    auto find_or_add = [&](const V3<R>& x) {
        const uint32_t h = vec_hash(x);
        for (size_t i = 0; i < vm.size(); i++)
            if (vh[i] == h && approx_eq(vm[i], x)) return (int)i;
        vm.push_back(x); vh.push_back(h);
        return (int)vm.size() - 1;
    };
    std::vector<std::pair<int, int>> seen;
    for (auto& f : p.f) {
        const int nf = (int)f.size();
        for (int i = 0; i < nf; i++) {
            const int v0 = f[i], v1 = f[(i + 1) % nf];
            int id0 = find_or_add(p.v[v0]);
            int id1 = find_or_add(p.v[v1]);
            if (id0 > id1) std::swap(id0, id1);
            if (std::find(seen.begin(), seen.end(), std::make_pair(id0, id1)) == seen.end()) {
                seen.push_back({id0, id1});
                p.e.push_back({p.v[v0], p.v[v1]});
            }
        }
    }
}
// ConvexMeshShape::setMesh (convex_mesh_shape.cpp:12-132): volume, centre of mass and unit-mass
// inertia by the polyhedral mass-property integrals, vertices moved to the centre of mass,
// BVH with one face per leaf (faces reordered), unique edges.  Returns the relocation.
template <class R> inline V3<R> convex_set_mesh(Poly<R>& p) {
    R mult[10] = {R(1.0) / 6, R(1.0) / 24, R(1.0) / 24, R(1.0) / 24, R(1.0) / 60,
                  R(1.0) / 60, R(1.0) / 60, R(1.0) / 120, R(1.0) / 120, R(1.0) / 120};
    R intg[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto sub = [](R w0, R w1, R w2, R& f1, R& f2, R& f3, R& g0, R& g1, R& g2) {  // :29-40
        const R tmp0 = w0 + w1;
        f1 = tmp0 + w2;
        const R tmp1 = w0 * w0;
        const R tmp2 = tmp1 + w1 * tmp0;
        f2 = tmp2 + w2 * f1;
        f3 = w0 * tmp1 + w1 * tmp2 + w2 * f2;
        g0 = f2 + w0 * (f1 + w0);
        g1 = f2 + w1 * (f1 + w1);
        g2 = f2 + w2 * (f1 + w2);
    };
    for (auto& f : p.f) {
        const V3<R>& v0 = p.v[f[0]];
        for (int i = 0; i < (int)f.size() - 2; i++) {
            const V3<R>& v1 = p.v[f[i + 1]];
            const V3<R>& v2 = p.v[f[i + 2]];
            const R x0 = v0[0], y0 = v0[1], z0 = v0[2];
            const R x1 = v1[0], y1 = v1[1], z1 = v1[2];
            const R x2 = v2[0], y2 = v2[1], z2 = v2[2];
            const R a1 = x1 - x0, b1 = y1 - y0, c1 = z1 - z0;
            const R a2 = x2 - x0, b2 = y2 - y0, c2 = z2 - z0;
            const R d0 = b1 * c2 - b2 * c1;
            const R d1 = a2 * c1 - a1 * c2;
            const R d2 = a1 * b2 - a2 * b1;
            R f1x, f2x, f3x, g0x, g1x, g2x, f1y, f2y, f3y, g0y, g1y, g2y, f1z, f2z, f3z, g0z, g1z, g2z;
            sub(x0, x1, x2, f1x, f2x, f3x, g0x, g1x, g2x);
            sub(y0, y1, y2, f1y, f2y, f3y, g0y, g1y, g2y);
            sub(z0, z1, z2, f1z, f2z, f3z, g0z, g1z, g2z);
            intg[0] += d0 * f1x;
            intg[1] += d0 * f2x;
            intg[2] += d1 * f2y;
            intg[3] += d2 * f2z;
            intg[4] += d0 * f3x;
            intg[5] += d1 * f3y;
            intg[6] += d2 * f3z;
            intg[7] += d0 * (y0 * g0x + y1 * g1x + y2 * g2x);
            intg[8] += d1 * (z0 * g0y + z1 * g1y + z2 * g2y);
            intg[9] += d2 * (x0 * g0z + x1 * g1z + x2 * g2z);
        }
    }
    for (int i = 0; i < 10; i++) intg[i] *= mult[i];
    const R V = intg[0];
    p.volume = V;
    const V3<R> cm(intg[1] / V, intg[2] / V, intg[3] / V);
    M3<R> I;
    I(0, 0) = intg[5] + intg[6] - V * (cm[1] * cm[1] + cm[2] * cm[2]);
    I(1, 1) = intg[4] + intg[6] - V * (cm[2] * cm[2] + cm[0] * cm[0]);
    I(2, 2) = intg[4] + intg[5] - V * (cm[0] * cm[0] + cm[1] * cm[1]);
    I(0, 1) = I(1, 0) = -intg[7] + V * cm[0] * cm[1];
    I(1, 2) = I(2, 1) = -intg[8] + V * cm[1] * cm[2];
    I(0, 2) = I(2, 0) = -intg[9] + V * cm[2] * cm[0];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) I(r, c) = I(r, c) / V;
    p.inertia = I;
    for (auto& x : p.v) x = x - cm;
    const int node = std::max((int)p.f.size() / 8191, 1);
    p.bvh.clear();
    bvh_build(p, 0, (int)p.f.size() - 1, node);
    unique_edges(p);
    return cm;
}
// ConcaveMeshShape::setMesh (concave_mesh_shape.cpp:6-17): identity inertia, zero volume, BVH
// with max(F / 2047, 1) faces per leaf; no unique edges.
template <class R> inline void concave_set_mesh(Poly<R>& p) {
    p.inertia = M3<R>::identity();
    p.volume = 0;
    const int node = std::max((int)p.f.size() / 2047, 1);
    p.bvh.clear();
    bvh_build(p, 0, (int)p.f.size() - 1, node);
}

// ---------------------------------------------------------------- shape queries
// Xf helpers: Transform::inverse and Transform::operator* (transform.cpp:7-9, 66-70)
template <class R> inline Xf<R> xf_inverse(const Xf<R>& t) {
    Xf<R> r;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.basis(i, j) = t.basis(j, i);
    r.origin = -mtv(t.basis, t.origin);
    return r;
}
template <class R> inline Xf<R> xf_mul(const Xf<R>& a, const Xf<R>& b) {
    return Xf<R>{mm(a.basis, b.basis), mv(a.basis, b.origin) + a.origin};
}

// The participant of a convex routine: box, cylinder (analytic projection over a polyhedral
// mesh), or a mesh shape (hull, or the one-face shape a concave wrapper builds per triangle).
template <class R> struct Part {
    int kind;                 // ST_BOX, ST_CYLINDER or ST_CONVEX
    const Body<R>* b;         // box / cylinder parameters
    const Poly<R>* p;         // faces, edges, vertices
};

// Shape::project: box_shape.cpp:51-68, cylinder_shape.cpp:50-79, convex_mesh_shape.cpp:160-184
template <class R> inline void shape_project(const Part<R>& s, const Xf<R>& t, const V3<R>& axis, R& mn, R& mx,
                                             V3<R>& pmn, V3<R>& pmx) {
    const V3<R> la = mtv(t.basis, axis);
    const R off = dot(t.origin, axis);
    if (s.kind == ST_BOX) {
        const R* h = s.b->half;
        const V3<R> ext(la[0] > 0 ? h[0] : -h[0], la[1] > 0 ? h[1] : -h[1], la[2] > 0 ? h[2] : -h[2]);
        const R he = dot(ext, la);
        mn = off - he; mx = off + he;
        pmn = t.apply(-ext); pmx = t.apply(ext);
    } else if (s.kind == ST_CYLINDER) {
        const R hh = s.b->height * R(0.5), rad = s.b->radius;
        const R eps = R(0.00001);
        if (std::abs(la[0] - 0) < eps && std::abs(la[2] - 0) < eps) {
            const V3<R> up = V3<R>(0, 1, 0) * hh;
            pmn = t.apply(-up); pmx = t.apply(up);
            mn = off - hh; mx = off + hh;
        } else if (std::abs(la[1] - 0) < eps) {
            const V3<R> r = la * rad;
            pmn = t.apply(-r); pmx = t.apply(r);
            mn = off - rad; mx = off + rad;
        } else {
            const R r = std::sqrt(la[0] * la[0] + la[2] * la[2]);
            const R x = rad * la[0] / r, z = rad * la[2] / r;
            const R y = hh * R(la[1] > 0 ? 1 : -1);
            const V3<R> ext(x, y, z);
            pmn = t.apply(-ext); pmx = t.apply(ext);
            const R el = dot(ext, la);
            mn = off - el; mx = off + el;
        }
    } else {
        // ConvexMeshShape::project: the max test is an else-branch of the min test, so a vertex
        // that lowers the minimum never raises the maximum (and the first vertex never does).
        mn = Lim<R>::max(); mx = Lim<R>::min();
        V3<R> lmn, lmx;
        for (const auto& q : s.p->v) {
            const R x = dot(q, la);
            if (x < mn) { mn = x; lmn = q; }
            else if (x > mx) { mx = x; lmx = q; }
        }
        mn += off; mx += off;
        pmn = t.apply(lmn); pmx = t.apply(lmx);
    }
}
// Shape::localIsInside: box_shape.cpp:45-49, cylinder_shape.cpp:45-48, convex_mesh_shape.cpp:152-158
template <class R> inline bool shape_inside(const Part<R>& s, const V3<R>& q) {
    if (s.kind == ST_BOX) {
        const R* h = s.b->half;
        return std::abs(q[0]) <= h[0] && std::abs(q[1]) <= h[1] && std::abs(q[2]) <= h[2];
    }
    if (s.kind == ST_CYLINDER) {
        const R hh = s.b->height;
        return q[1] >= -hh * R(0.5) && q[1] <= hh * R(0.5) && q[0] * q[0] + q[2] * q[2] <= s.b->radius * s.b->radius;
    }
    for (size_t f = 0; f < s.p->f.size(); f++)
        if (dot(s.p->fn[f], q - s.p->v[s.p->f[f][0]]) >= 0) return false;
    return true;
}

// ---------------------------------------------------------------- convex SAT + clipping
// convex_convex_collision_algorithm.cpp
template <class R> inline void clip_face(const std::vector<V3<R>>& in, std::vector<V3<R>>& out,
                                         const V3<R>& pn, R pe) {  // :29-61
    const int nv = (int)in.size();
    if (nv < 2) return;
    V3<R> first = in[nv - 1];
    R ds = dot(pn, first) + pe;
    for (int ve = 0; ve < nv; ve++) {
        const V3<R> end = in[ve];
        const R de = dot(pn, end) + pe;
        if (ds < 0) {
            if (de < 0) out.push_back(end);
            else out.push_back(first + (end - first) * R(ds * R(1.f) / (ds - de)));
        } else if (de < 0) {
            out.push_back(first + (end - first) * R(ds * R(1.f) / (ds - de)));
            out.push_back(end);
        }
        first = end;
        ds = de;
    }
}
template <class R> inline void clip_hull_hull(const World<R>& w, ContactResult<R>& res, const V3<R>& sep,
                                              const Part<R>& A, const Part<R>& B, const Xf<R>& ta, const Xf<R>& tb,
                                              R min_dist, R max_dist) {  // :63-170
    const Poly<R>& pa = *A.p;
    const Poly<R>& pb = *B.p;
    int cb = -1;
    R dmax = Lim<R>::min();
    for (size_t f = 0; f < pb.f.size(); f++) {
        const R d = dot(mv(tb.basis, pb.fn[f]), sep);
        if (d > dmax) { dmax = d; cb = (int)f; }
    }
    if (cb < 0) return;
    std::vector<V3<R>> b1, b2;
    for (int i : pb.f[cb]) b1.push_back(tb.apply(pb.v[i]));
    int ca = -1;
    R dmin = Lim<R>::max();
    for (size_t f = 0; f < pa.f.size(); f++) {
        const R d = dot(mv(ta.basis, pa.fn[f]), sep);
        if (d < dmin) { dmin = d; ca = (int)f; }
    }
    if (ca < 0) return;
    const auto& poly = pa.f[ca];
    const int nva = (int)poly.size();
    const V3<R> wn = mv(ta.basis, pa.fn[ca]);
    std::vector<V3<R>>* pin = &b1;
    std::vector<V3<R>>* pout = &b2;
    for (int e0 = 0; e0 < nva; e0++) {
        const V3<R> a = pa.v[poly[e0]], b = pa.v[poly[(e0 + 1) % nva]];
        const V3<R> ew = mv(ta.basis, a - b);
        const V3<R> pn = -cross(ew, wn);
        const V3<R> wa = ta.apply(a);
        const R pe = -dot(wa, pn);
        clip_face(*pin, *pout, pn, pe);
        std::swap(pin, pout);
        pout->clear();
    }
    const V3<R> pv = pa.v[poly[0]];
    const V3<R> ln = pa.fn[ca];
    for (const V3<R>& vtx : *pin) {
        const R depth = dot(ln, ta.inv_apply(vtx) - pv);
        if (depth <= max_dist && depth >= min_dist) add_contact_point(w, res, sep, vtx - sep * R(0), depth + R(0) * 2);
    }
}
template <class R> inline bool test_sep_axis(const Part<R>& A, const Part<R>& B, const Xf<R>& ta, const Xf<R>& tb,
                                             const V3<R>& ax, R& depth, V3<R>& wa, V3<R>& wb) {  // :216-243
    R min0, max0, min1, max1;
    V3<R> pmna, pmxa, pmnb, pmxb;
    shape_project(A, ta, ax, min0, max0, pmna, pmxa);
    shape_project(B, tb, ax, min1, max1, pmnb, pmxb);
    if (max0 < min1 || max1 < min0) return false;
    const R d0 = max0 - min1, d1 = max1 - min0;
    if (d0 < d1) { depth = d0; wa = pmxa; wb = pmnb; }
    else { depth = d1; wa = pmna; wb = pmxb; }
    return true;
}
template <class R> inline bool on_line(const V3<R>& v0, const V3<R>& v1, const V3<R>& v2) {  // :245-249
    return sqnorm(cross(v1 - v0, v2 - v0)) < R(0.00001);
}
template <class R> inline void segments_closest(V3<R>& pv, V3<R>& oa, V3<R>& ob, R& ta_, R& tb_, const V3<R>& tr,
                                                const V3<R>& da, R ha, const V3<R>& db, R hb) {  // :172-214
    const R dadb = dot(da, db), dat = dot(da, tr), dbt = dot(db, tr);
    const R d = R(1.0) - dadb * dadb;
    if (d == 0) ta_ = 0;
    else {
        ta_ = (dat - dbt * dadb) / d;
        if (ta_ < -ha) ta_ = -ha;
        else if (ta_ > ha) ta_ = ha;
    }
    tb_ = ta_ * dadb - dbt;
    if (tb_ < -hb) {
        tb_ = -hb;
        ta_ = tb_ * dadb + dat;
        if (ta_ < -ha) ta_ = -ha;
        else if (ta_ > ha) ta_ = ha;
    } else if (tb_ > hb) {
        tb_ = hb;
        ta_ = tb_ * dadb + dat;
        if (ta_ < -ha) ta_ = -ha;
        else if (ta_ > ha) ta_ = ha;
    }
    oa = da * ta_;
    ob = db * tb_;
    pv = tr - oa + ob;
}
// findSeparatingAxis (:251-402).  The deep-point contact is added before the edge axes are
// tested, so it stays in the result even when an edge axis then separates the shapes.
template <class R> inline bool find_sep_axis(const World<R>& w, ContactResult<R>& res, const Part<R>& A,
                                             const Part<R>& B, const Xf<R>& ta, const Xf<R>& tb, V3<R>& sep) {
    const V3<R> dc = ta.origin - tb.origin;
    R dmin = Lim<R>::max();
    V3<R> dmin_pt;
    bool pt_on_a = false;
    for (size_t f = 0; f < A.p->f.size(); f++) {
        V3<R> n = mv(ta.basis, A.p->fn[f]);
        if (dot(dc, n) < 0) n = n * R(-1.0);
        R d;
        V3<R> wa, wb;
        if (!test_sep_axis(A, B, ta, tb, n, d, wa, wb)) return false;
        if (d < dmin) { dmin = d; dmin_pt = wb; pt_on_a = false; sep = n; }
    }
    for (size_t f = 0; f < B.p->f.size(); f++) {
        V3<R> n = mv(tb.basis, B.p->fn[f]);
        if (dot(dc, n) < 0) n = n * R(-1.0);
        R d;
        V3<R> wa, wb;
        if (!test_sep_axis(A, B, ta, tb, n, d, wa, wb)) return false;
        if (d < dmin) { dmin = d; dmin_pt = wa + n * d; pt_on_a = true; sep = n; }
    }
    if ((pt_on_a && shape_inside(B, tb.inv_apply(dmin_pt))) || (!pt_on_a && shape_inside(A, ta.inv_apply(dmin_pt))))
        add_contact_point(w, res, sep, dmin_pt - sep * R(0), -dmin + R(0) * 2);
    int ea = -1, eb = -1;
    V3<R> wea, web, wpa(0, 0, 0), wpb(0, 0, 0);
    const auto& EA = A.p->e;
    const auto& EB = B.p->e;
    for (int e0 = 0; e0 < (int)EA.size(); e0++) {
        const V3<R> we0 = mv(ta.basis, normalized(EA[e0].first - EA[e0].second));
        for (int e1 = 0; e1 < (int)EB.size(); e1++) {
            const V3<R> we1 = mv(tb.basis, normalized(EB[e1].first - EB[e1].second));
            V3<R> c = cross(we0, we1);
            if (!(std::abs(norm(c) - 0) < R(0.00001))) {
                c = normalized(c);
                if (dot(dc, c) < 0) c = c * R(-1.0);
                R dist;
                V3<R> wa, wb;
                if (!test_sep_axis(A, B, ta, tb, c, dist, wa, wb)) return false;
                if (dist < dmin) {
                    if (!on_line(EA[e0].first, EA[e0].second, ta.inv_apply(wa)) ||
                        !on_line(EB[e1].first, EB[e1].second, tb.inv_apply(wb)))
                        continue;
                    dmin = dist; sep = c; ea = e0; eb = e1; wea = we0; web = we1; wpa = wa; wpb = wb;
                }
            }
        }
    }
    if (ea >= 0 && eb >= 0) {
        V3<R> pv, oa, ob;
        R t_a, t_b;
        const V3<R> tr = wpb - wpa;
        segments_closest(pv, oa, ob, t_a, t_b, tr, wea, Lim<R>::max(), web, Lim<R>::max());
        const R nl2 = sqnorm(pv);
        if (nl2 > R(0.00001)) {
            const R nl = std::sqrt(nl2);
            pv = pv * (R(1.0) / nl);
            if (dot(pv, dc) < 0) pv = pv * R(-1.0);
            const V3<R> pob = wpb + ob;
            add_contact_point(w, res, pv, pob - pv * R(0), -nl + R(0) * 2);
        }
    }
    if (dot(dc, sep) < 0) sep = -sep;
    return true;
}
// ConvexConvexCollisionAlgorithm::getClosestPoints (:408-432)
template <class R> inline bool convex_convex(const World<R>& w, ContactResult<R>& res, const Part<R>& A,
                                             const Part<R>& B, const Xf<R>& ta, const Xf<R>& tb, R ref_scale) {
    V3<R> sep;
    if (!find_sep_axis(w, res, A, B, ta, tb, sep)) return false;
    clip_hull_hull(w, res, sep, A, B, ta, tb, -ref_scale, R(0));
    return true;
}

// ---------------------------------------------------------------- sphere-cylinder
// sphere_cylinder_collision_algorithm.cpp:32-77 (A side = cylinder, B side = sphere)
template <class R> inline bool sphere_cylinder(const World<R>& w, ContactResult<R>& res, const Body<R>& sph,
                                               const Body<R>& cyl) {
    const R s_r = sph.radius, c_r = cyl.radius, c_h = cyl.height * R(0.5);
    const V3<R> sp = cyl.xf.inv_apply(sph.xf.origin);
    const R r = std::sqrt(sp[0] * sp[0] + sp[2] * sp[2]);
    const R d = std::abs(sp[1]) - c_h;
    if (sp[1] < -c_h - s_r || sp[1] > c_h + s_r || r * r > (c_r + s_r) * (c_r + s_r) ||
        (d > 0 && r > c_r && (r - c_r) * (r - c_r) + d * d > s_r * s_r))
        return false;
    V3<R> n, pos;
    R depth;
    if (sp[1] >= -c_h && sp[1] <= c_h && r > 0) {
        n = normalized(V3<R>(-sp[0], 0, -sp[2]));
        pos = n * s_r + sp;
        depth = c_r + s_r - r;
    } else if (r <= c_r) {
        n = V3<R>(0, 1, 0);
        if (sp[1] > 0) n = -n;
        pos = sp + n * s_r;
        depth = s_r + c_h - std::abs(sp[1]);
    } else {
        V3<R> pc = normalized(V3<R>(sp[0], 0, sp[2])) * c_r;
        pc[1] = sp[1] > 0 ? c_h : -c_h;
        n = normalized(pc - sp);
        pos = n * s_r + sp;
        depth = s_r - norm(sp - pc);
    }
    n = mv(cyl.xf.basis, n);
    pos = cyl.xf.apply(pos);
    add_contact_point(w, res, n, pos - n * R(0), -depth + 2 * R(0));
    return true;
}

// ---------------------------------------------------------------- box-cylinder
// box_cylinder_collision_algorithm.cpp
template <class R> inline bool point_in_box(const V3<R>& he, const V3<R>& q) {  // :68-74
    for (int i = 0; i < 3; i++)
        if (q[i] > he[i] || q[i] < -he[i]) return false;
    return true;
}
template <class R> inline void add_point_on_box(const World<R>& w, ContactResult<R>& res, const V3<R>& pc, int face,
                                                const V3<R>& he, const Xf<R>& to) {  // :79-114
    if (!(face >= -3 && face <= 3 && face != 0)) return;
    if (!point_in_box(he, pc)) return;
    V3<R> p = pc, n(0, 0, 0);
    R pen;
    if (face > 0) { const int i = face - 1; p[i] = he[i]; n[i] = 1; pen = pc[i] - he[i]; }
    else { const int i = -face - 1; p[i] = -he[i]; n[i] = -1; pen = -pc[i] - he[i]; }
    const V3<R> nw = mv(to.basis, n);
    const V3<R> pw = to.apply(p);
    add_contact_point(w, res, nw, pw - nw * R(0), pen + 2 * R(0));
}
template <class R> inline void add_point_on_cyl(const World<R>& w, ContactResult<R>& res, const V3<R>& p,
                                                const V3<R>& pc, const V3<R>& ax, R rad, R h,
                                                const Xf<R>& to) {  // :116-139
    const V3<R> r = p - pc;
    const R dy = dot(r, ax);
    const R dxz = std::sqrt(sqnorm(r) - dy * dy);
    const R d_side = rad - dxz, d_top = h - std::abs(dy);
    V3<R> n;
    R dist;
    if (d_side < d_top) { dist = d_side; n = normalized(ax * dy - r); }
    else { dist = d_top; n = dy < 0 ? ax : -ax; }
    n = mv(to.basis, n);
    const V3<R> pw = to.apply(p);
    add_contact_point(w, res, n, pw - n * R(0), -dist + 2 * R(0));
}
template <class R> inline int closest_box_face(const V3<R>& he, const V3<R>& q) {  // :141-157
    int code = 0;
    R dist = Lim<R>::max();
    for (int i = 0; i < 3; i++) {
        const R pd = std::abs(q[i] - he[i]), nd = std::abs(q[i] + he[i]);
        if (pd < dist) { code = i + 1; dist = pd; }
        if (nd < dist) { code = -(i + 1); dist = nd; }
    }
    return code;
}
template <class R> inline bool line_plane(const V3<R>& s, const V3<R>& d, const V3<R>& pp, const V3<R>& pn, R tol,
                                          R& t) {  // :159-169
    const R nd = dot(pn, d);
    if (std::abs(nd) < tol) return false;
    t = dot(pn, pp - s) / nd;
    return true;
}
template <class R> inline bool seg_box(const V3<R>& he, const V3<R>& pc, const V3<R>& ax, R h, R tol, R& tmin,
                                       R& tmax) {  // :171-231
    tmin = Lim<R>::min(); tmax = Lim<R>::max();
    for (int k = 0; k < 3; k++) {
        if (std::abs(ax[k]) < tol) {
            if (std::abs(pc[k]) > he[k]) return false;
        } else {
            const R t1 = (-he[k] - pc[k]) / ax[k];
            const R t2 = (+he[k] - pc[k]) / ax[k];
            const R lo = t1 < t2 ? t1 : t2, hi = t1 > t2 ? t1 : t2;
            tmin = tmin > lo ? tmin : lo;
            tmax = tmax < hi ? tmax : hi;
            if (tmin > tmax) return false;
        }
    }
    if ((tmin < -h && tmax < -h) || (tmin > +h && tmax > +h)) return false;
    auto clamp = [](R x, R lo, R hi) { const R m = x < hi ? x : hi; return m > lo ? m : lo; };
    tmin = clamp(tmin, -h, +h);
    tmax = clamp(tmax, -h, +h);
    return true;
}
template <class R> inline bool seg_cyl(const V3<R>& s, const V3<R>& sd, R len, const V3<R>& pc, const V3<R>& ax,
                                       R h, R rad, R tol, R& tmin, R& tmax) {  // :233-304
    tmin = Lim<R>::min(); tmax = Lim<R>::max();
    const V3<R> v = s - pc;
    const R cd_sd = dot(ax, sd), v_cd = dot(v, ax), v_sd = dot(v, sd), v_v = dot(v, v);
    const R a = R(1.0) - cd_sd * cd_sd;
    const R b = v_sd - v_cd * cd_sd;
    const R c = v_v - v_cd * v_cd - rad * rad;
    if (std::abs(a) < tol) {
        const R d2 = sqnorm(v - v_cd * ax);
        if (d2 > rad * rad) return false;
        tmin = -len; tmax = +len;
    } else {
        R delta = b * b - a * c;
        if (delta < 0) return false;
        delta = std::sqrt(delta);
        tmin = (-b - delta) / a;
        tmax = (-b + delta) / a;
    }
    R t1, t2;
    const bool c1 = line_plane(s, sd, pc + h * ax, ax, tol, t1);
    const bool c2 = line_plane(s, sd, pc - h * ax, ax, tol, t2);
    if (c1 && c2) {
        if (t1 < t2) { tmin = tmin > t1 ? tmin : t1; tmax = tmax < t2 ? tmax : t2; }
        else { tmin = tmin > t2 ? tmin : t2; tmax = tmax < t1 ? tmax : t1; }
        if (tmax < tmin) return false;
    } else {
        const R d1 = std::abs(dot(ax, pc + h * ax - s));
        const R d2 = std::abs(dot(ax, pc - h * ax - s));
        if (d1 > 2 * h || d2 > 2 * h) return false;
    }
    if ((tmin < -len && tmax < -len) || (tmin > +len && tmax > +len)) return false;
    auto clamp = [](R x, R lo, R hi) { const R m = x < hi ? x : hi; return m > lo ? m : lo; };
    tmin = clamp(tmin, -len, +len);
    tmax = clamp(tmax, -len, +len);
    return true;
}
// BoxCylinderCollisionAlgorithm::getClosestPoints (:306-388) (A side = cylinder, B side = box)
template <class R> inline bool box_cylinder(const World<R>& w, ContactResult<R>& res, const Body<R>& box,
                                            const Body<R>& cyl) {
    const R rad = cyl.radius, h = cyl.height / R(2.0);
    const V3<R> he(box.size[0] / R(2.0), box.size[1] / R(2.0), box.size[2] / R(2.0));
    const Xf<R> rel = xf_mul(xf_inverse(box.xf), cyl.xf);
    const V3<R> ax = rel.basis.col(1);
    const V3<R> pc = rel.origin;
    const R eps = R(0.00001);
    for (int id = 0; id < 3; id++) {
        V3<R> nd(0, 0, 0);
        nd[id] = 1;
        if (std::abs(ax[id] - 1) < eps || std::abs(ax[id] + 1) < eps) continue;
        const V3<R> r = normalized(cross(cross(nd, ax), ax)) * rad;
        const R dirs[2] = {-1, 1};
        for (R jd : dirs) {
            const V3<R> cs = pc + jd * r;
            R tmin, tmax;
            if (seg_box(he, cs, ax, h, eps, tmin, tmax)) {
                const V3<R> pmin = cs + ax * tmin;
                const V3<R> pmax = cs + ax * tmax;
                const V3<R> pmid = cs + ax * ((tmin + tmax) / 2);
                const int face = closest_box_face(he, pmid);
                add_point_on_box(w, res, pmin, face, he, box.xf);
                add_point_on_box(w, res, pmax, face, he, box.xf);
            }
        }
    }
    for (int id = 0; id < 3; id++) {
        V3<R> ed(0, 0, 0);
        ed[id] = 1;
        const R eh = he[id];
        const int jd = (id + 1) % 3, kd = (id + 2) % 3;
        for (int j = -1; j <= 1; j += 2) {
            for (int k = -1; k <= 1; k += 2) {
                V3<R> ec;
                ec[id] = 0;
                ec[jd] = R(j) * he[jd];
                ec[kd] = R(k) * he[kd];
                R tmin, tmax;
                if (seg_cyl(ec, ed, eh, pc, ax, h, rad, eps, tmin, tmax)) {
                    const V3<R> pmid = ec + ed * ((tmin + tmax) / 2);
                    add_point_on_cyl(w, res, pmid, pc, ax, rad, h, box.xf);
                }
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------- cylinder-mesh
// cylinder_convex_collision_algorithm.cpp:53-107
template <class R> inline bool seg_face(const Poly<R>& m, int fi, const V3<R>& ps, const V3<R>& sd, R l, R& t1,
                                        R& t2, R& d1, R& d2) {
    const auto& f = m.f[fi];
    const V3<R>& fn = m.fn[fi];
    const V3<R> org = m.v[f[0]];
    const V3<R> end = ps + sd * l;
    const V3<R> r1 = ps - org, r2 = end - org;
    const R dist1 = dot(r1, fn), dist2 = dot(r2, fn);
    if ((dist1 > 0 && dist2 > 0) || (dist1 <= 0 && dist2 <= 0)) return false;
    const R t = dist1 / (dist1 - dist2) * l;
    const V3<R> v = ps + sd * t;
    R tmin = 0, tmax = l;
    R mind = Lim<R>::max();
    const int nf = (int)f.size();
    for (int i = 0; i < nf; i++) {
        const V3<R>& v0 = m.v[f[i]];
        const V3<R>& v1 = m.v[f[(i + 1) % nf]];
        const V3<R> tc = normalized(cross(fn, v1 - v0));
        const R ddc = dot(sd, tc);
        const R ed = std::abs(dot(v0 - v, tc));
        mind = mind < ed ? mind : ed;
        if (!(std::abs(ddc - 0) < R(0.00001))) {
            if (ddc > 0) { const R x = dot(v0 - ps, tc) / ddc; tmin = tmin > x ? tmin : x; }
            else { const R x = dot(v0 - ps, tc) / ddc; tmax = tmax < x ? tmax : x; }
        } else {
            if (dot(cross(v0 - v, v1 - v), fn) < 0) return false;
        }
    }
    if (t < tmin || t > tmax) return false;
    t1 = dist1 <= 0 ? tmin : t;
    t2 = dist2 <= 0 ? tmax : t;
    d1 = dist1 <= 0 ? dist1 * (t - tmin) / t : 0;
    d2 = dist2 <= 0 ? dist2 * (tmax - t) / (l - t) : 0;
    if (-d1 > mind + R(0) || -d2 > mind + R(0)) return false;
    return true;
}
// CylinderConvexCollisionAlgorithm::getClosestPoints (:109-203) with the mesh of a convex or
// concave body (A side = cylinder, B side = mesh)
template <class R> inline bool cylinder_mesh(const World<R>& w, ContactResult<R>& res, const Body<R>& mb,
                                             const Poly<R>& m, const Body<R>& cyl) {
    const R rad = cyl.radius, h = cyl.height / R(2.0);
    const Xf<R> rel = xf_mul(xf_inverse(mb.xf), cyl.xf);
    const V3<R> ax = rel.basis.col(1);
    const V3<R> pc = rel.origin;
    std::vector<int> hit;
    {
        Body<R> tmp = cyl;
        tmp.xf = rel;
        body_aabb(tmp);
        bvh_search(m, tmp.aabb_min, tmp.aabb_max, hit);
    }
    const Xf<R>& to = mb.xf;
    auto add_seg = [&](int fi, const V3<R>& ps, const V3<R>& sd, R l) {
        R t1, t2, d1, d2;
        if (seg_face(m, fi, ps, sd, l, t1, t2, d1, d2)) {
            const V3<R>& fn = m.fn[fi];
            const V3<R> p1 = to.apply(ps + sd * t1 - fn * d1);
            const V3<R> p2 = to.apply(ps + sd * t2 - fn * d2);
            const V3<R> n = mv(to.basis, fn);
            add_contact_point(w, res, n, p1 - n * R(0), d1 + R(0) * 2);
            add_contact_point(w, res, n, p2 - n * R(0), d2 + R(0) * 2);
        }
    };
    static const int dx[4] = {1, -1, -1, 1};
    static const int dy[4] = {1, 1, -1, -1};
    for (int fi : hit) {
        const V3<R>& fn = m.fn[fi];
        const R adc = dot(ax, fn);
        V3<R> ux, uy = ax, uz;
        bool vert = false;
        if (std::abs(std::abs(adc) - 1) < R(0.00001)) {
            uz = normalized(cross(ax, V3<R>(1, 0, 0)));
            ux = normalized(cross(ax, uz));
            vert = true;
        } else {
            ux = normalized(cross(cross(fn, ax), ax));
        }
        for (int i = 0; i < 4; i++) {
            const V3<R> ps = pc + ux * R(dx[i]) * rad + uy * R(dy[i]) * h;
            const V3<R> sd = (dx[i] == dy[i] ? uy : ux) * R(-dx[i]);
            const R l = (dx[i] == dy[i] ? h : rad) * 2;
            add_seg(fi, ps, sd, l);
        }
        if (vert) {
            for (int i = 0; i < 3; i += 2) {
                const V3<R> ps = pc + uz * R(dx[i]) * rad + uy * R(dy[i]) * h;
                const V3<R> sd = uy * R(-dx[i]);
                const R l = h * 2;
                add_seg(fi, ps, sd, l);
            }
        }
    }
    for (int fi : hit) {
        const auto& f = m.f[fi];
        const int nf = (int)f.size();
        for (int i = 0; i < nf; i++) {
            const V3<R>& v1 = m.v[f[i]];
            const V3<R>& v2 = m.v[f[(i + 1) % nf]];
            R t1, t2;
            const V3<R> ss = (v1 + v2) / R(2);
            const R l = norm(v2 - v1) / R(2);
            const V3<R> sd = (v2 - v1) / (l * 2);
            if (seg_cyl(ss, sd, l, pc, ax, h, rad, R(0), t1, t2)) {
                const V3<R> p1 = ss + sd * t1;
                const V3<R> p2 = ss + sd * t2;
                const V3<R> pm = (p1 + p2) / R(2);
                add_point_on_cyl(w, res, pm, pc, ax, rad, h, to);
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------- sphere-mesh
// sphere_convex_collision_algorithm.cpp:36-211 (A side = sphere, B side = mesh)
template <class R> inline R seg_sqdist(const V3<R>& from, const V3<R>& to, const V3<R>& p, V3<R>& nearest) {  // :36-57
    V3<R> diff = p - from;
    const V3<R> v = to - from;
    R t = dot(v, diff);
    if (t > 0) {
        const R vv = dot(v, v);
        if (t < vv) { t /= vv; diff = diff - t * v; }
        else { t = 1; diff = diff - v; }
    } else {
        t = 0;
    }
    nearest = from + t * v;
    return dot(diff, diff);
}
template <class R> inline bool point_in_tri(const V3<R>* vt, const V3<R>& n, const V3<R>& p) {  // :59-83
    const V3<R> e1 = vt[1] - vt[0], e2 = vt[2] - vt[1], e3 = vt[0] - vt[2];
    const V3<R> p1 = p - vt[0], p2 = p - vt[1], p3 = p - vt[2];
    const R r1 = dot(cross(e1, n), p1);
    const R r2 = dot(cross(e2, n), p2);
    const R r3 = dot(cross(e3, n), p3);
    return (r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0);
}
template <class R> inline bool sphere_tri(const V3<R>& c, R rad, const V3<R>* vt, V3<R>& point, V3<R>& rn,
                                          R& depth) {  // :91-163
    const R eps = R(0.00001);
    V3<R> n = cross(vt[1] - vt[0], vt[2] - vt[0]);
    const R l2 = sqnorm(n);
    bool has = false;
    V3<R> cp;
    if (l2 >= eps * eps) {
        n = n / std::sqrt(l2);
        const V3<R> p1c = c - vt[0];
        R dist = dot(p1c, n);
        if (dist < 0) { dist *= R(-1.); n = n * R(-1.); }
        if (dist < rad) {
            if (point_in_tri(vt, n, c)) {
                has = true;
                cp = c - n * dist;
            } else {
                R mind = rad * rad;
                V3<R> ne;
                for (int i = 0; i < 3; i++) {
                    const R d2 = seg_sqdist(vt[i], vt[(i + 1) % 3], c, ne);
                    if (d2 < mind) { mind = d2; has = true; cp = ne; }
                }
            }
        }
    }
    if (has) {
        const V3<R> cc = c - cp;
        const R d2 = sqnorm(cc);
        if (d2 < rad * rad) {
            if (d2 > eps * eps) {
                const R d = std::sqrt(d2);
                rn = normalized(cc);
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
template <class R> inline bool sphere_mesh(const World<R>& w, ContactResult<R>& res, const Body<R>& sph,
                                           const Body<R>& mb, const Poly<R>& m) {  // :165-211
    const V3<R> c = mb.xf.inv_apply(sph.xf.origin);
    const R rad = sph.radius;
    const V3<R> aa = c - V3<R>(rad, rad, rad), bb = c + V3<R>(rad, rad, rad);
    std::vector<int> hit;
    bvh_search(m, aa, bb, hit);
    bool ret = false;
    for (int fi : hit) {
        const auto& f = m.f[fi];
        for (int i = 0; i < (int)f.size() - 2; i++) {
            V3<R> vt[3] = {m.v[f[0]], m.v[f[i + 1]], m.v[f[i + 2]]};
            const V3<R> cs = mb.xf.inv_apply(sph.xf.origin);
            V3<R> p, n;
            R depth = 0;
            if (sphere_tri(cs, rad, vt, p, n, depth)) {
                n = mv(mb.xf.basis, n);
                p = mb.xf.apply(p);
                add_contact_point(w, res, n, p - n * R(0), depth + 2 * R(0));
                ret = true;
            }
        }
    }
    return ret;
}

// ---------------------------------------------------------------- concave (triangle mesh) + polyhedron
// ConcaveConvexCollisionAlgorithm::getClosestPoints (concave_convex_collision_algorithm.cpp:38-68):
// every triangle whose BVH leaf overlaps the polyhedron's AABB (in the mesh frame) becomes a
// one-face mesh shape (no unique edges) and goes through the convex routine with the
// polyhedron as A.
template <class R> inline bool concave_poly(const World<R>& w, ContactResult<R>& res, const Body<R>& mb,
                                            const Poly<R>& m, const Part<R>& A, const Body<R>& ab, R ref_scale) {
    const Xf<R> rel = xf_mul(xf_inverse(mb.xf), ab.xf);
    Body<R> tmp = ab;
    tmp.xf = rel;
    body_aabb(tmp);
    const V3<R> aa = tmp.aabb_min, bb = tmp.aabb_max;
    std::vector<int> hit;
    bvh_search(m, aa, bb, hit);
    for (int fi : hit) {
        Poly<R> tri;
        const auto& f = m.f[fi];
        std::vector<int> idx;
        for (int i = 0; i < (int)f.size(); i++) { tri.v.push_back(m.v[f[i]]); idx.push_back(i); }
        tri.f.push_back(idx);
        tri.fn.push_back(m.fn[fi]);
        const Part<R> B{ST_CONVEX, &mb, &tri};
        convex_convex(w, res, A, B, ab.xf, mb.xf, ref_scale);
    }
    return true;
}

// unit-mass local inertia of a mesh body
template <class R> M3<R> poly_inertia(const Poly<R>& p) { return p.inertia; }
// ConvexMeshShape::getAABB (convex_mesh_shape.cpp:138-150)
template <class R> void poly_aabb(const Poly<R>& p, const M3<R>& rot, const V3<R>& pos, V3<R>& mn, V3<R>& mx) {
    mn = V3<R>(Lim<R>::max(), Lim<R>::max(), Lim<R>::max());
    mx = V3<R>(Lim<R>::min(), Lim<R>::min(), Lim<R>::min());
    for (const auto& q : p.v) {
        const V3<R> x = mv(rot, q);
        for (int k = 0; k < 3; k++) {
            mn[k] = mn[k] < x[k] ? mn[k] : x[k];
            mx[k] = mx[k] > x[k] ? mx[k] : x[k];
        }
    }
    mn = mn + pos;
    mx = mx + pos;
}

}  // namespace pe_oracle
