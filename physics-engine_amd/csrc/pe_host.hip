// Host-side set-up shared by the pipelines (see pe_host.hpp).
#include "pe_host.hpp"
#include "pe_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>

namespace pe {

namespace {

struct V3f { float x, y, z; };
inline V3f vsub(V3f a, V3f b) { return V3f{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float vget(const V3f& a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Eigen's 3x3 inverse in FP32 (InverseImpl.h, cofactor form): det = c0 m00 + (c1 m10 + c2 m20),
// entry (i, j) = cofactor(j, i) / det.
void inverse3(const float m[9], float r[9]) {
    auto M = [&](int i, int j) { return m[3 * i + j]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
    };
    const float c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
    const float det = c0 * M(0, 0) + (c1 * M(1, 0) + c2 * M(2, 0));
    const float inv = 1.0f / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[3 * i + j] = cof(j, i) * inv;
}

struct Face { std::vector<int> idx; float n[3]; };

// utils/bvh.cpp:45-83 over faces[l..r]: node bounds, leaf when at most nmax faces, otherwise
// the faces are ordered with std::sort by their centroid along the longest axis of the node
// and split at the middle.  Nodes are numbered in creation order (root 0, then the left
// subtree, then the right subtree).  With sorted = true the faces are already in the order a
// build left them (a shape after setMesh): every node's range then holds the faces the build put
// there (later sorts only permute inside halves), so the same tree follows without sorting.
struct BvhNode { float aa[3], bb[3]; int left = -1, right = -1, n = 0, index = 0; };
int build_bvh(std::vector<Face>& faces, const std::vector<V3f>& v, int l, int r, int nmax, std::vector<BvhNode>& out,
              bool sorted = false) {
    if (l > r) return -1;
    const int id = (int)out.size();
    out.emplace_back();
    BvhNode nd;
    for (int k = 0; k < 3; k++) { nd.aa[k] = REAL_MAX; nd.bb[k] = REAL_MIN; }
    for (int i = l; i <= r; i++) {
        float mn[3] = {REAL_MAX, REAL_MAX, REAL_MAX}, mx[3] = {REAL_MIN, REAL_MIN, REAL_MIN};
        for (int q : faces[i].idx)
            for (int k = 0; k < 3; k++) {
                const float c = vget(v[q], k);
                if (c < mn[k]) mn[k] = c;
                if (c > mx[k]) mx[k] = c;
            }
        for (int k = 0; k < 3; k++) {
            if (mn[k] < nd.aa[k]) nd.aa[k] = mn[k];
            if (mx[k] > nd.bb[k]) nd.bb[k] = mx[k];
        }
    }
    if (r - l + 1 <= nmax) {
        nd.n = r - l + 1;
        nd.index = l;
        out[id] = nd;
        return id;
    }
    const float lx = nd.bb[0] - nd.aa[0], ly = nd.bb[1] - nd.aa[1], lz = nd.bb[2] - nd.aa[2];
    const int axis = (lx >= ly && lx >= lz) ? 0 : (ly >= lz ? 1 : 2);
    auto centre = [&](const Face& f) {
        float s = 0;
        for (int q : f.idx) s += vget(v[q], axis);
        s /= (float)f.idx.size();
        return s;
    };
    if (!sorted)
        std::sort(faces.begin() + l, faces.begin() + r + 1, [&](const Face& a, const Face& b) { return centre(a) < centre(b); });
    const int mid = (l + r) / 2;
    nd.left = build_bvh(faces, v, l, mid, nmax, out, sorted);
    nd.right = build_bvh(faces, v, mid + 1, r, nmax, out, sorted);
    out[id] = nd;
    return id;
}

// pe::Vector3Hash (phys_general.h:133-143) of the coordinates, as the reference computes it
uint32_t vhash(const V3f& p) {
    auto ui = [](float c) { return (uint32_t)(int64_t)((double)c * 1e5); };
    uint32_t h = 0x995af;
    h ^= ui(p.x) + 0x9e3779b9u + (h << 6) + (h >> 2);
    h ^= ui(p.y) + 0x9e3779b9u + (h << 6) + (h >> 2);
    h ^= ui(p.z) + 0x9e3779b9u + (h << 6) + (h >> 2);
    return h;
}

}  // namespace

void cylinder_unit_table(float out[CYL_UNIT_FLOATS]) {
    // Ring vertices: 0.5 cos / -0.5 sin of the ring angle (top i: 18 i degrees, bottom i:
    // 9 + 18 i degrees), rounded to 6 decimals as the reference's table lists them.
    const double deg = 3.14159265358979323846 / 180.0;
    auto r6 = [](double x) { return std::round(x * 1e6) / 1e6 + 0.0; };
    for (int i = 0; i < 20; i++) {
        out[i] = (float)r6(0.5 * std::cos(18.0 * i * deg));
        out[20 + i] = (float)r6(-0.5 * std::sin(18.0 * i * deg));
        out[40 + i] = (float)r6(0.5 * std::cos((9.0 + 18.0 * i) * deg));
        out[60 + i] = (float)r6(-0.5 * std::sin((9.0 + 18.0 * i) * deg));
    }
    // Side-face normals as the reference's table lists them (6 decimals): magnitudes by the
    // face's ring angle reduced to the first quadrant, signs from the angle; the table's face
    // 8 (ring angle 117 degrees) has x = -0.493982.
    static const double up[5] = {0.98767, 0.89099, 0.707093, 0.453982, 0.156432};
    static const double dn[6] = {0.999981, 0.951038, 0.809002, 0.587774, 0.309011, 0.0};
    for (int f = 0; f < 40; f++) {
        const bool is_up = f < 20;
        const int k = is_up ? f : f - 20;
        const double a = is_up ? 9.0 + 18.0 * k : 18.0 * k;
        const double c = std::cos(a * deg), s = std::sin(a * deg);
        double q = std::fmod(a, 180.0);
        if (q > 90.0) q = 180.0 - q;
        double mc, ms;
        if (is_up) { const int j = (int)std::lround((q - 9.0) / 18.0); mc = up[j]; ms = up[4 - j]; }
        else { const int j = (int)std::lround(q / 18.0); mc = dn[j]; ms = dn[5 - j]; }
        double n0 = std::fabs(c) < 1e-9 ? 0.0 : (c > 0 ? mc : -mc);
        const double n1 = is_up ? 0.006156 : -0.006156;
        const double n2 = std::fabs(s) < 1e-9 ? 0.0 : (s > 0 ? -ms : ms);
        if (is_up && k == 6) n0 = -0.493982;
        out[80 + 3 * f] = (float)n0;
        out[80 + 3 * f + 1] = (float)n1;
        out[80 + 3 * f + 2] = (float)n2;
    }
}

int HostMeshes::add(const pe_mesh_desc& md, bool concave, float* relocation, float* volume_inertia, bool prepared) {
    if (md.nv <= 0 || md.nf <= 0 || !md.vertices || !md.face_start || !md.face_index || !md.face_normal) {
        pe_set_error("mesh: empty or null arrays");
        return -1;
    }
    std::vector<V3f> v(md.nv);
    for (int i = 0; i < md.nv; i++)
        v[i] = V3f{(float)md.vertices[3 * i], (float)md.vertices[3 * i + 1], (float)md.vertices[3 * i + 2]};
    std::vector<Face> faces(md.nf);
    for (int f = 0; f < md.nf; f++) {
        const int a = md.face_start[f], b = md.face_start[f + 1];
        if (b - a < 3 || (concave && b - a != 3)) {
            pe_set_error(concave ? "concave mesh: every face must be a triangle" : "mesh: a face has fewer than 3 vertices");
            return -1;
        }
        if (b - a > MAX_FACE_VERTS) {
            pe_set_error("mesh: a face has more than " + std::to_string(MAX_FACE_VERTS) + " vertices (the clipper's limit)");
            return -1;
        }
        for (int k = a; k < b; k++) {
            const int q = md.face_index[k];
            if (q < 0 || q >= md.nv) { pe_set_error("mesh: face index out of range"); return -1; }
            faces[f].idx.push_back(q);
        }
        for (int k = 0; k < 3; k++) faces[f].n[k] = (float)md.face_normal[3 * f + k];
    }
    float vi[10] = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    float cm[3] = {0, 0, 0};
    int nmax;
    if (!concave) {
        // convex_mesh_shape.cpp:29-104: polyhedral mass-property integrals over the faces'
        // fans, centre of mass, unit-mass inertia about it; vertices moved to the centre.
        const float mult[10] = {1.0f / 6, 1.0f / 24, 1.0f / 24, 1.0f / 24, 1.0f / 60,
                                1.0f / 60, 1.0f / 60, 1.0f / 120, 1.0f / 120, 1.0f / 120};
        float in[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        auto terms = [](float w0, float w1, float w2, float& f1, float& f2, float& f3, float& g0, float& g1, float& g2) {
            const float t0 = w0 + w1;
            f1 = t0 + w2;
            const float t1 = w0 * w0;
            const float t2 = t1 + w1 * t0;
            f2 = t2 + w2 * f1;
            f3 = w0 * t1 + w1 * t2 + w2 * f2;
            g0 = f2 + w0 * (f1 + w0);
            g1 = f2 + w1 * (f1 + w1);
            g2 = f2 + w2 * (f1 + w2);
        };
        for (const Face& fc : faces) {
            const V3f& p0 = v[fc.idx[0]];
            for (int i = 0; i + 2 < (int)fc.idx.size(); i++) {
                const V3f& p1 = v[fc.idx[i + 1]];
                const V3f& p2 = v[fc.idx[i + 2]];
                const float a1 = p1.x - p0.x, b1 = p1.y - p0.y, c1 = p1.z - p0.z;
                const float a2 = p2.x - p0.x, b2 = p2.y - p0.y, c2 = p2.z - p0.z;
                const float d0 = b1 * c2 - b2 * c1, d1 = a2 * c1 - a1 * c2, d2 = a1 * b2 - a2 * b1;
                float f1x, f2x, f3x, g0x, g1x, g2x, f1y, f2y, f3y, g0y, g1y, g2y, f1z, f2z, f3z, g0z, g1z, g2z;
                terms(p0.x, p1.x, p2.x, f1x, f2x, f3x, g0x, g1x, g2x);
                terms(p0.y, p1.y, p2.y, f1y, f2y, f3y, g0y, g1y, g2y);
                terms(p0.z, p1.z, p2.z, f1z, f2z, f3z, g0z, g1z, g2z);
                in[0] += d0 * f1x;
                in[1] += d0 * f2x;
                in[2] += d1 * f2y;
                in[3] += d2 * f2z;
                in[4] += d0 * f3x;
                in[5] += d1 * f3y;
                in[6] += d2 * f3z;
                in[7] += d0 * (p0.y * g0x + p1.y * g1x + p2.y * g2x);
                in[8] += d1 * (p0.z * g0y + p1.z * g1y + p2.z * g2y);
                in[9] += d2 * (p0.x * g0z + p1.x * g1z + p2.x * g2z);
            }
        }
        for (int i = 0; i < 10; i++) in[i] *= mult[i];
        const float V = in[0];
        if (!(V > 0.0f)) { pe_set_error("convex mesh: non-positive volume (faces must be counter-clockwise seen from outside)"); return -1; }
        cm[0] = in[1] / V; cm[1] = in[2] / V; cm[2] = in[3] / V;
        float I[9];
        I[0] = in[5] + in[6] - V * (cm[1] * cm[1] + cm[2] * cm[2]);
        I[4] = in[4] + in[6] - V * (cm[2] * cm[2] + cm[0] * cm[0]);
        I[8] = in[4] + in[5] - V * (cm[0] * cm[0] + cm[1] * cm[1]);
        I[1] = I[3] = -in[7] + V * cm[0] * cm[1];
        I[5] = I[7] = -in[8] + V * cm[1] * cm[2];
        I[2] = I[6] = -in[9] + V * cm[2] * cm[0];
        vi[0] = V;
        for (int k = 0; k < 9; k++) vi[1 + k] = I[k] / V;
        if (!prepared) for (auto& p : v) p = V3f{p.x - cm[0], p.y - cm[1], p.z - cm[2]};
        nmax = std::max(md.nf / 8191, 1);
    } else {
        vi[0] = 0.0f;   // concave_mesh_shape.cpp:9-10: identity inertia, zero volume
        nmax = std::max(md.nf / 2047, 1);
    }
    std::vector<BvhNode> nodes;
    build_bvh(faces, v, 0, md.nf - 1, nmax, nodes, prepared);
    {
        // the device walk pushes both children of an inner node and holds at most one pending
        // sibling per level above it, so a tree of depth D needs D + 2 stack entries
        std::vector<int> depth(nodes.size(), 0);
        int dmax = 0;
        for (size_t k = 0; k < nodes.size(); k++) {
            dmax = std::max(dmax, depth[k]);
            if (nodes[k].left >= 0) depth[nodes[k].left] = depth[k] + 1;
            if (nodes[k].right >= 0) depth[nodes[k].right] = depth[k] + 1;
        }
        if (dmax + 2 > BVH_STACK) { pe_set_error("mesh: BVH deeper than the device walk's stack"); return -1; }
    }
    // unique edges (convex_mesh_shape.cpp:110-128), convex meshes only
    std::vector<std::pair<V3f, V3f>> edges;
    if (!concave) {
        std::vector<V3f> vm;
        std::vector<uint32_t> vh;
        auto id_of = [&](const V3f& p) {
            const uint32_t h = vhash(p);
            for (size_t i = 0; i < vm.size(); i++)
                if (vh[i] == h && std::fabs(vm[i].x - p.x) < EPS && std::fabs(vm[i].y - p.y) < EPS && std::fabs(vm[i].z - p.z) < EPS)
                    return (int)i;
            vm.push_back(p);
            vh.push_back(h);
            return (int)vm.size() - 1;
        };
        std::map<std::pair<int, int>, bool> seen;
        for (const Face& fc : faces) {
            const int nfv = (int)fc.idx.size();
            for (int i = 0; i < nfv; i++) {
                const V3f& a = v[fc.idx[i]];
                const V3f& b = v[fc.idx[(i + 1) % nfv]];
                int i0 = id_of(a), i1 = id_of(b);
                if (i0 > i1) std::swap(i0, i1);
                if (seen.emplace(std::make_pair(i0, i1), true).second) edges.push_back({a, b});
            }
        }
    }
    // arena layout
    const int32_t base = (int32_t)arena.size();
    int nidx = 0;
    for (const Face& fc : faces) nidx += (int)fc.idx.size();
    int32_t r[MRW];
    r[MR_NV] = md.nv; r[MR_NF] = md.nf; r[MR_NE] = (int)edges.size(); r[MR_NB] = (int)nodes.size(); r[MR_PAD] = 0;
    int32_t o = base;
    r[MR_V] = o; o += md.nv;
    r[MR_FN] = o; o += md.nf;
    r[MR_FACE] = o; o += (md.nf + 1) / 2;
    r[MR_FI] = o; o += (nidx + 3) / 4;
    r[MR_E] = o; o += 2 * (int)edges.size();
    r[MR_NODE] = o; o += 2 * (int)nodes.size();
    r[MR_LEAF] = o; o += ((int)nodes.size() + 1) / 2;
    arena.resize(o, make_float4(0, 0, 0, 0));
    float4* A = arena.data();
    for (int i = 0; i < md.nv; i++) A[r[MR_V] + i] = make_float4(v[i].x, v[i].y, v[i].z, 0.0f);
    int2* face = reinterpret_cast<int2*>(A + r[MR_FACE]);
    int32_t* fi = reinterpret_cast<int32_t*>(A + r[MR_FI]);
    int q = 0;
    for (int f = 0; f < md.nf; f++) {
        A[r[MR_FN] + f] = make_float4(faces[f].n[0], faces[f].n[1], faces[f].n[2], 0.0f);
        face[f] = make_int2(q, (int)faces[f].idx.size());
        for (int x : faces[f].idx) fi[q++] = x;
    }
    for (size_t e = 0; e < edges.size(); e++) {
        A[r[MR_E] + 2 * e] = make_float4(edges[e].first.x, edges[e].first.y, edges[e].first.z, 0.0f);
        A[r[MR_E] + 2 * e + 1] = make_float4(edges[e].second.x, edges[e].second.y, edges[e].second.z, 0.0f);
    }
    int2* leaf = reinterpret_cast<int2*>(A + r[MR_LEAF]);
    for (size_t k = 0; k < nodes.size(); k++) {
        const BvhNode& nd = nodes[k];
        int32_t lb, rb;
        std::memcpy(&lb, &nd.left, 4);
        std::memcpy(&rb, &nd.right, 4);
        float lf, rf;
        std::memcpy(&lf, &lb, 4);
        std::memcpy(&rf, &rb, 4);
        A[r[MR_NODE] + 2 * k] = make_float4(nd.aa[0], nd.aa[1], nd.aa[2], lf);
        A[r[MR_NODE] + 2 * k + 1] = make_float4(nd.bb[0], nd.bb[1], nd.bb[2], rf);
        leaf[k] = make_int2(nd.n, nd.index);
    }
    rec.insert(rec.end(), r, r + MRW);
    if (relocation) for (int k = 0; k < 3; k++) relocation[k] = cm[k];
    if (volume_inertia) std::memcpy(volume_inertia, vi, sizeof(vi));
    props.insert(props.end(), vi, vi + 10);
    return count() - 1;
}

void HostMeshes::rebase(int32_t off) {
    for (int m = 0; m < count(); m++)
        for (int k : {MR_V, MR_FN, MR_FACE, MR_FI, MR_E, MR_NODE, MR_LEAF}) rec[MRW * m + k] += off;
}

int check_scene(const pe_scene_desc* sc) {
    if (!sc || sc->n <= 0) { pe_set_error("empty scene"); return PE_ERR_INVALID; }
    for (int i = 0; i < sc->n; i++) {
        const int t = sc->shape_type[i];
        if (t < PE_SHAPE_BOX || t > PE_SHAPE_COMPOUND) {
            pe_set_error("body " + std::to_string(i) + ": unknown shape type " + std::to_string(t));
            return PE_ERR_UNSUPPORTED;
        }
        if (t == PE_SHAPE_COMPOUND) {
            // CompoundShape::addShape (compound_shape.cpp:7-20): no compound or concave children, mass
            // ratios above PE_EPS
            if (!sc->compound_start || !sc->compound_children || sc->compound_start[i + 1] <= sc->compound_start[i]) {
                pe_set_error("body " + std::to_string(i) + ": compound body without children");
                return PE_ERR_INVALID;
            }
            for (int k = sc->compound_start[i]; k < sc->compound_start[i + 1]; k++) {
                const pe_compound_child& ch = sc->compound_children[k];
                const bool mesh_ok = ch.shape_type != PE_SHAPE_CONVEX ||
                                     (sc->meshes && ch.mesh_index >= 0 && ch.mesh_index < sc->n_meshes);
                if (ch.shape_type < PE_SHAPE_BOX || ch.shape_type > PE_SHAPE_CONVEX || !mesh_ok || !(ch.mass_ratio >= 1e-5)) {
                    pe_set_error("body " + std::to_string(i) + ": compound child " + std::to_string(k) +
                                 " is not a box, sphere, cylinder or convex mesh with a mass ratio above 1e-5");
                    return PE_ERR_INVALID;
                }
            }
        }
        if (t == PE_SHAPE_CONVEX || t == PE_SHAPE_CONCAVE) {
            if (!sc->mesh_index || !sc->meshes || sc->mesh_index[i] < 0 || sc->mesh_index[i] >= sc->n_meshes) {
                pe_set_error("body " + std::to_string(i) + ": convex / concave body without a valid mesh");
                return PE_ERR_INVALID;
            }
            if (t == PE_SHAPE_CONCAVE && !sc->kinematic[i]) {
                pe_set_error("body " + std::to_string(i) + ": a concave mesh body must be kinematic");
                return PE_ERR_UNSUPPORTED;
            }
        }
    }
    return PE_OK;
}

namespace {
// Unit-mass inertia of a box / sphere / cylinder (box_shape.cpp:16-22, sphere_shape.cpp:5-12,
// cylinder_shape.cpp:23-31), diagonal.
void primitive_inertia(int t, const float sp[3], float I[9]) {
    for (int k = 0; k < 9; k++) I[k] = 0.0f;
    if (t == PE_SHAPE_BOX) {
        const float x2 = sp[0] * sp[0], y2 = sp[1] * sp[1], z2 = sp[2] * sp[2];
        I[0] = (y2 + z2) / 12.0f; I[4] = (x2 + z2) / 12.0f; I[8] = (x2 + y2) / 12.0f;
    } else if (t == PE_SHAPE_SPHERE) {
        I[0] = I[4] = I[8] = (float)(2.0 / 5.0) * sp[0] * sp[0];
    } else {
        const float r_2 = sp[0] * sp[0], h_2 = sp[1] * sp[1];
        I[4] = r_2 * 0.5f;
        I[0] = I[8] = r_2 / 4.0f + h_2 / 12.0f;
    }
}
float sum3h(float a, float b, float c) { return a + (b + c); }   // Eigen's FP32 3-term sum

// CompoundShape::addShape for each child of body i, then CompoundShape::init (compound_shape.cpp:
// 24-47) in FP32 and the reference's evaluation order: the children's origins move by minus their
// mass-ratio-weighted centroid, and the unit-mass inertia is the sum over the children of
// ((R I_child) R^T + |p|^2 1 - p p^T) ratio / total.  Appends the children to the compound table
// and returns the compound id (or -1).
int stage_compound(const pe_scene_desc* sc, int i, HostMeshes& meshes, std::map<std::pair<int, int>, int>& ids,
                   float I[9]) {
    const int k0 = sc->compound_start[i], k1 = sc->compound_start[i + 1];
    float total = 0.0f;
    for (int k = k0; k < k1; k++) total += (float)sc->compound_children[k].mass_ratio;
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int k = k0; k < k1; k++) {
        const pe_compound_child& ch = sc->compound_children[k];
        const float r = (float)ch.mass_ratio;
        for (int q = 0; q < 3; q++) c[q] = c[q] + (float)ch.position[q] * r;
    }
    for (int q = 0; q < 3; q++) c[q] = c[q] / total;
    for (int q = 0; q < 9; q++) I[q] = 0.0f;
    const int first = (int)meshes.kids.size() / (CHW / 4);
    for (int k = k0; k < k1; k++) {
        const pe_compound_child& ch = sc->compound_children[k];
        float rec[CHW] = {};
        float sp[3] = {(float)ch.shape_param[0], (float)ch.shape_param[1], (float)ch.shape_param[2]};
        float si[9];
        if (ch.shape_type == PE_SHAPE_CONVEX) {
            const auto key = std::make_pair(ch.mesh_index, 0);
            auto it = ids.find(key);
            if (it == ids.end()) {
                const int id = meshes.add(sc->meshes[ch.mesh_index], false);
                if (id < 0) return -1;
                it = ids.emplace(key, id).first;
            }
            const int id = it->second;
            for (int q = 0; q < 9; q++) si[q] = meshes.props[10 * id + 1 + q];
            int32_t bits = id;
            std::memcpy(&sp[0], &bits, 4);
            sp[1] = sp[2] = 0.0f;
        } else {
            primitive_inertia(ch.shape_type, sp, si);
        }
        float R[9], p[3];
        for (int q = 0; q < 9; q++) R[q] = (float)ch.rotation[q];
        for (int q = 0; q < 3; q++) p[q] = (float)ch.position[q] - c[q];
        const float pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        float T[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) T[3 * r + q] = sum3h(R[3 * r] * si[q], R[3 * r + 1] * si[3 + q], R[3 * r + 2] * si[6 + q]);
        const float w = (float)ch.mass_ratio / total;
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
                const float a = sum3h(T[3 * r] * R[3 * q], T[3 * r + 1] * R[3 * q + 1], T[3 * r + 2] * R[3 * q + 2]);
                const float t = (r == q ? 1.0f : 0.0f) * pp - p[r] * p[q];
                I[3 * r + q] += (a + t) * w;
            }
        int32_t tb = ch.shape_type;
        std::memcpy(&rec[CH_TYPE], &tb, 4);
        for (int q = 0; q < 3; q++) rec[CH_SP + q] = sp[q];
        for (int q = 0; q < 9; q++) rec[CH_R + q] = R[q];
        for (int q = 0; q < 3; q++) rec[CH_P + q] = p[q];
        for (int q = 0; q < CHW / 4; q++) meshes.kids.push_back(make_float4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]));
    }
    meshes.krange.push_back(make_int2(first, k1 - k0));
    return (int)meshes.krange.size() - 1;
}
}  // namespace

int stage_bodies(const pe_scene_desc* sc, std::vector<float>& hf, std::vector<int32_t>& hi, size_t stride, int off,
                 HostMeshes& meshes) {
    std::map<std::pair<int, int>, int> ids;   // (scene mesh, concave) -> mesh id
    for (int i = 0; i < sc->n; i++) {
        const size_t col = (size_t)off + i;
        auto F = [&](int f) -> float& { return hf[(size_t)f * stride + col]; };
        const int t = sc->shape_type[i];
        hi[(size_t)I_TYPE * stride + col] = t;
        hi[(size_t)I_KIN * stride + col] = sc->kinematic[i] ? 1 : 0;
        for (int k = 0; k < 3; k++) F(F_SP + k) = (float)sc->shape_param[4 * i + k];
        // the fourth shape parameter carries the body's collision flag (COLL_*)
        F(F_SP + 3) = (sc->ignore_collision && sc->ignore_collision[i]) ? COLL_IGNORE : COLL_NORMAL;
        for (int k = 0; k < 3; k++) {
            F(F_POS + k) = (float)sc->position[3 * i + k];
            F(F_V + k) = (float)sc->linear_velocity[3 * i + k];
            F(F_W + k) = (float)sc->angular_velocity[3 * i + k];
        }
        for (int k = 0; k < 9; k++) F(F_ROT + k) = (float)sc->rotation[9 * i + k];
        const float mass = (float)sc->mass[i];
        F(F_MASS) = mass;
        F(F_IM) = mass > EPS ? 1.0f / mass : 0.0f;
        // shape inertia at unit mass: box_shape.cpp:16-22, sphere_shape.cpp:5-12,
        // cylinder_shape.cpp:23-31, convex_mesh_shape.cpp:93-100, concave_mesh_shape.cpp:9-10
        float I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (t == PE_SHAPE_BOX) {
            const float x2 = F(F_SP) * F(F_SP), y2 = F(F_SP + 1) * F(F_SP + 1), z2 = F(F_SP + 2) * F(F_SP + 2);
            I[0] = (y2 + z2) / 12.0f; I[4] = (x2 + z2) / 12.0f; I[8] = (x2 + y2) / 12.0f;
        } else if (t == PE_SHAPE_SPHERE) {
            const float r = F(F_SP);
            I[0] = I[4] = I[8] = (float)(2.0 / 5.0) * r * r;
        } else if (t == PE_SHAPE_CYLINDER) {
            const float r = F(F_SP), h = F(F_SP + 1);
            const float r_2 = r * r, h_2 = h * h;
            I[4] = r_2 * 0.5f;
            I[0] = I[8] = r_2 / 4.0f + h_2 / 12.0f;
        } else if (t == PE_SHAPE_COMPOUND) {
            const int id = stage_compound(sc, i, meshes, ids, I);
            if (id < 0) return PE_ERR_INVALID;
            int32_t bits = id;
            std::memcpy(&F(F_SP), &bits, 4);
            F(F_SP + 1) = F(F_SP + 2) = 0.0f;
        } else {
            const bool concave = t == PE_SHAPE_CONCAVE;
            const auto key = std::make_pair(sc->mesh_index[i], concave ? 1 : 0);
            auto it = ids.find(key);
            if (it == ids.end()) {
                const int id = meshes.add(sc->meshes[key.first], concave);
                if (id < 0) return PE_ERR_INVALID;
                it = ids.emplace(key, id).first;
            }
            const int id = it->second;
            for (int k = 0; k < 9; k++) I[k] = meshes.props[10 * id + 1 + k];
            int32_t bits = id;
            std::memcpy(&F(F_SP), &bits, 4);
            F(F_SP + 1) = F(F_SP + 2) = 0.0f;
        }
        float L[9];
        for (int k = 0; k < 9; k++) L[k] = I[k] * mass;
        for (int k = 0; k < 9; k++) F(F_LINV + k) = 0.0f;
        if (mass > EPS) {
            float Li[9];
            inverse3(L, Li);
            for (int k = 0; k < 9; k++) F(F_LINV + k) = Li[k];
        }
        F(F_MU) = (float)sc->friction[i];
        F(F_E) = (float)sc->restitution[i];
        F(F_LD) = sc->linear_damping ? (float)sc->linear_damping[i] : 0.0f;
        F(F_AD) = sc->angular_damping ? (float)sc->angular_damping[i] : 0.0f;
    }
    return PE_OK;
}

int DevMeshes::upload(const HostMeshes& h, hipStream_t s) {
    release();
    float unit[CYL_UNIT_FLOATS];
    cylinder_unit_table(unit);
    const size_t nrec = std::max<size_t>(h.rec.size(), MRW), nar = std::max<size_t>(h.arena.size(), 1);
    if (hipMalloc(&rec, nrec * sizeof(int32_t)) != hipSuccess || hipMalloc(&arena, nar * sizeof(float4)) != hipSuccess ||
        hipMalloc(&cyl, CYL_UNIT_FLOATS * sizeof(float)) != hipSuccess) {
        pe_set_error("mesh arena allocation failed");
        return PE_ERR_DEVICE;
    }
    if ((!h.rec.empty() && hipMemcpyAsync(rec, h.rec.data(), h.rec.size() * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess) ||
        (!h.arena.empty() && hipMemcpyAsync(arena, h.arena.data(), h.arena.size() * sizeof(float4), hipMemcpyHostToDevice, s) != hipSuccess) ||
        hipMemcpyAsync(cyl, unit, sizeof(unit), hipMemcpyHostToDevice, s) != hipSuccess) {
        pe_set_error("mesh arena upload failed");
        return PE_ERR_DEVICE;
    }
    if (!h.krange.empty()) {
        if (hipMalloc(&kids, h.kids.size() * sizeof(float4)) != hipSuccess ||
            hipMalloc(&krange, h.krange.size() * sizeof(int2)) != hipSuccess ||
            hipMemcpyAsync(kids, h.kids.data(), h.kids.size() * sizeof(float4), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(krange, h.krange.data(), h.krange.size() * sizeof(int2), hipMemcpyHostToDevice, s) != hipSuccess) {
            pe_set_error("compound table upload failed");
            return PE_ERR_DEVICE;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        pe_set_error("mesh arena upload failed");
        return PE_ERR_DEVICE;
    }
    return PE_OK;
}

void DevMeshes::release() {
    if (rec) hipFree(rec);
    if (arena) hipFree(arena);
    if (cyl) hipFree(cyl);
    if (kids) hipFree(kids);
    if (krange) hipFree(krange);
    rec = nullptr;
    arena = nullptr;
    cyl = nullptr;
    kids = nullptr;
    krange = nullptr;
}

}  // namespace pe

// ------------------------------------------------------------------ C ABI: mesh preparation
extern "C" int pe_mesh_prepare(const pe_mesh_desc* mesh, int32_t concave, int32_t* counts, float* vertices,
                               int32_t* face_start, int32_t* face_index, float* face_normal, float* edges,
                               float* volume_inertia, float* relocation) {
    if (!mesh || !counts) { pe_set_error("pe_mesh_prepare: null argument"); return PE_ERR_INVALID; }
    pe::HostMeshes h;
    float vi[10], cm[3];
    if (h.add(*mesh, concave != 0, cm, vi) < 0) return PE_ERR_INVALID;
    const int32_t* r = h.rec.data();
    const float4* A = h.arena.data();
    int nidx = 0;
    const int2* face = reinterpret_cast<const int2*>(A + r[pe::MR_FACE]);
    for (int f = 0; f < r[pe::MR_NF]; f++) nidx += face[f].y;
    counts[0] = r[pe::MR_NV]; counts[1] = r[pe::MR_NF]; counts[2] = nidx; counts[3] = r[pe::MR_NE];
    if (vertices)
        for (int i = 0; i < r[pe::MR_NV]; i++) {
            const float4 p = A[r[pe::MR_V] + i];
            vertices[3 * i] = p.x; vertices[3 * i + 1] = p.y; vertices[3 * i + 2] = p.z;
        }
    const int32_t* fi = reinterpret_cast<const int32_t*>(A + r[pe::MR_FI]);
    for (int f = 0; f < r[pe::MR_NF]; f++) {
        if (face_start) face_start[f] = face[f].x;
        if (face_normal) {
            const float4 n = A[r[pe::MR_FN] + f];
            face_normal[3 * f] = n.x; face_normal[3 * f + 1] = n.y; face_normal[3 * f + 2] = n.z;
        }
    }
    if (face_start) face_start[r[pe::MR_NF]] = nidx;
    if (face_index) for (int k = 0; k < nidx; k++) face_index[k] = fi[k];
    if (edges)
        for (int e = 0; e < r[pe::MR_NE]; e++) {
            const float4 a = A[r[pe::MR_E] + 2 * e], b = A[r[pe::MR_E] + 2 * e + 1];
            const float q[6] = {a.x, a.y, a.z, b.x, b.y, b.z};
            std::memcpy(edges + 6 * e, q, sizeof(q));
        }
    if (volume_inertia) std::memcpy(volume_inertia, vi, sizeof(vi));
    if (relocation) std::memcpy(relocation, cm, sizeof(cm));
    return PE_OK;
}
