// The reference engine's World stages over the library's stage entry points (pe_amd_stages.hpp).
#include "pe_amd_stages.hpp"

#include "rigid/object/rigidbody.h"
#include "rigid/shape/box_shape.h"
#include "rigid/shape/sphere_shape.h"
#include "rigid/shape/cylinder_shape.h"
#include "rigid/shape/convex_mesh_shape.h"
#include "rigid/shape/concave_mesh_shape.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pe_amd_integration {

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void pe_check(int rc, const char* what) {
    if (rc != PE_OK) throw std::runtime_error(std::string(what) + ": " + pe_last_error());
}
float f32(pe::Real x) { return static_cast<float>(x); }
float int_bits(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// Record field offsets (include/pe_amd.h, "stage entry points")
constexpr int CB_R = 0, CB_P = 9, CB_SCALE = 12, CB_SP = 13, CB_TYPE = 16, CB_KIN = 17, CB_MU = 18, CB_E = 19;
constexpr int SB_IM = 6, SB_KIN = 7, SB_IW = 8, SB_V = 17, SB_W = 20;

}  // namespace

// ------------------------------------------------------------------ device buffers
void DeviceBuffer::reserve(size_t need, void* stream) {
    if (need <= bytes) return;
    const size_t cap = std::max(need, bytes * 2);
    void* p = nullptr;
    hip_check(hipMalloc(&p, cap), "hipMalloc");
    if (ptr) {
        hip_check(hipMemcpyAsync(p, ptr, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        hipFree(ptr);
    }
    ptr = p;
    bytes = cap;
}
void DeviceBuffer::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

DeviceContext::DeviceContext() {
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    _stream = s;
}
DeviceContext::~DeviceContext() {
    if (_stream) hipStreamSynchronize((hipStream_t)_stream);
    for (DeviceBuffer* b : {&aabb, &pair_a, &pair_b, &result, &body, &solver, &points, &npoints, &swap, &solve_points})
        b->release();
    if (_stage) pe_stage_destroy(_stage);
    if (_stream) hipStreamDestroy((hipStream_t)_stream);
}

pe_stage* DeviceContext::stage(int n, int m) {
    if (!_stage || n > _max_bodies || m > _max_pairs) {
        sync();
        if (_stage) pe_stage_destroy(_stage);
        _stage = nullptr;
        _max_bodies = std::max({n, 2 * _max_bodies, 64});
        _max_pairs = std::max({m, 2 * _max_pairs, 1024});
        pe_check(pe_stage_create(_max_bodies, _max_pairs, &_stage), "pe_stage_create");
        _meshes_dirty = !_meshes.empty();
    }
    if (_meshes_dirty) register_meshes();
    return _stage;
}
void DeviceContext::sync() { hip_check(hipStreamSynchronize((hipStream_t)_stream), "hipStreamSynchronize"); }
void DeviceContext::upload(DeviceBuffer& b, const void* src, size_t bytes) {
    b.reserve(std::max<size_t>(bytes, 16), _stream);   // an empty array still gets a valid device pointer
    if (bytes && src) hip_check(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, (hipStream_t)_stream), "hipMemcpyAsync");
}
void DeviceContext::download(void* dst, const DeviceBuffer& b, size_t bytes) {
    if (bytes) hip_check(hipMemcpyAsync(dst, b.ptr, bytes, hipMemcpyDeviceToHost, (hipStream_t)_stream), "hipMemcpyAsync");
    sync();
}

int32_t DeviceContext::mesh_id(const pe_phys_shape::Shape* shape) {
    auto it = _mesh_ids.find(shape);
    if (it != _mesh_ids.end()) return it->second;
    const int32_t id = (int32_t)_meshes.size();
    _mesh_ids[shape] = id;
    _meshes.push_back(shape);
    _meshes_dirty = true;
    return id;
}

// The registered shapes' meshes as the shapes hold them after setMesh (prepared = 1: vertices
// about the centre of mass, faces in the BVH build's order, convex_mesh_shape.cpp:12-132).
void DeviceContext::register_meshes() {
    if (!_stage) return;
    std::vector<std::vector<double>> v(_meshes.size()), fn(_meshes.size());
    std::vector<std::vector<int32_t>> fs(_meshes.size()), fi(_meshes.size());
    std::vector<pe_mesh_desc> desc(_meshes.size());
    std::vector<int32_t> concave(_meshes.size());
    for (size_t k = 0; k < _meshes.size(); k++) {
        const auto* cs = dynamic_cast<const pe_phys_shape::ConvexMeshShape*>(_meshes[k]);
        const pe::Mesh& m = cs->getMesh();
        for (const auto& x : m.vertices)
            for (int c = 0; c < 3; c++) v[k].push_back((double)x.position[c]);
        fs[k].push_back(0);
        for (const auto& f : m.faces) {
            for (auto i : f.indices) fi[k].push_back((int32_t)i);
            fs[k].push_back((int32_t)fi[k].size());
            for (int c = 0; c < 3; c++) fn[k].push_back((double)f.normal[c]);
        }
        desc[k] = pe_mesh_desc{(int32_t)m.vertices.size(), v[k].data(), (int32_t)m.faces.size(), fs[k].data(), fi[k].data(),
                               fn[k].data()};
        concave[k] = _meshes[k]->getType() == pe_phys_shape::ShapeType::ST_ConcaveMesh ? 1 : 0;
    }
    sync();
    pe_check(pe_stage_set_meshes(_stage, (int32_t)desc.size(), desc.data(), concave.data(), 1), "pe_stage_set_meshes");
    _meshes_dirty = false;
}

void collision_record(const pe_phys_object::RigidBody* rb, DeviceContext& ctx, float* o) {
    const pe::Transform& t = rb->getTransform();
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) o[CB_R + 3 * r + c] = f32(t.getBasis()(r, c));
    for (int c = 0; c < 3; c++) o[CB_P + c] = f32(t.getOrigin()[c]);
    o[CB_SCALE] = f32(rb->getAABBScale());
    o[CB_SP] = o[CB_SP + 1] = o[CB_SP + 2] = 0.0f;
    const auto* shape = rb->getCollisionShape();
    int32_t type = -1;
    switch (shape->getType()) {
        case pe_phys_shape::ShapeType::ST_Box: {
            const auto& s = dynamic_cast<const pe_phys_shape::BoxShape*>(shape)->getSize();
            for (int c = 0; c < 3; c++) o[CB_SP + c] = f32(s[c]);
            type = PE_SHAPE_BOX;
            break;
        }
        case pe_phys_shape::ShapeType::ST_Sphere:
            o[CB_SP] = f32(dynamic_cast<const pe_phys_shape::SphereShape*>(shape)->getRadius());
            type = PE_SHAPE_SPHERE;
            break;
        case pe_phys_shape::ShapeType::ST_Cylinder: {
            const auto* s = dynamic_cast<const pe_phys_shape::CylinderShape*>(shape);
            o[CB_SP] = f32(s->getRadius());
            o[CB_SP + 1] = f32(s->getHeight());
            type = PE_SHAPE_CYLINDER;
            break;
        }
        case pe_phys_shape::ShapeType::ST_ConvexMesh:
            o[CB_SP] = int_bits(ctx.mesh_id(shape));
            type = PE_SHAPE_CONVEX;
            break;
        case pe_phys_shape::ShapeType::ST_ConcaveMesh:
            o[CB_SP] = int_bits(ctx.mesh_id(shape));
            type = PE_SHAPE_CONCAVE;
            break;
        default:
            throw std::runtime_error("compound shapes are not part of the device narrowphase");
    }
    o[CB_TYPE] = int_bits(type);
    o[CB_KIN] = int_bits(rb->isKinematic() ? 1 : 0);
    o[CB_MU] = f32(rb->getFrictionCoeff());
    o[CB_E] = f32(rb->getRestitutionCoeff());
}

// ------------------------------------------------------------------ broadphase
void AmdBroadPhase::calcCollisionPairs(pe::Array<pe_phys_object::RigidBody*> objects,
                                       pe::Array<pe_phys_collision::CollisionPair>& pairs) {
    const int n = (int)objects.size();
    if (n < 2) return;   // broad_phase_sweep_and_prune.cpp:8: the pair list is left as it is
    _rec.assign((size_t)n * PE_AABB_FLOATS, 0.0f);
    for (int i = 0; i < n; i++) {
        const auto* rb = objects[i];
        float* r = _rec.data() + (size_t)i * PE_AABB_FLOATS;
        for (int c = 0; c < 3; c++) {
            r[c] = f32(rb->getAABBMin()[c]);
            r[4 + c] = f32(rb->getAABBMax()[c]);
        }
        r[3] = rb->isKinematic() ? 1.0f : 0.0f;
        r[7] = (rb->isIgnoreCollision() || !rb->getCollisionShape()) ? 1.0f : 0.0f;
    }
    _ctx.upload(_ctx.aabb, _rec.data(), _rec.size() * sizeof(float));
    int32_t res[3] = {0, 0, 0};
    for (int attempt = 0;; attempt++) {
        pe_stage* st = _ctx.stage(n, _cap);
        _ctx.pair_a.reserve((size_t)_cap * 4, _ctx.stream());
        _ctx.pair_b.reserve((size_t)_cap * 4, _ctx.stream());
        _ctx.result.reserve(64, _ctx.stream());
        pe_check(pe_stage_broadphase(st, n, (const float*)_ctx.aabb.ptr, _axis, (int32_t*)_ctx.pair_a.ptr,
                                     (int32_t*)_ctx.pair_b.ptr, _cap, (int32_t*)_ctx.result.ptr, _ctx.stream()),
                 "pe_stage_broadphase");
        _ctx.download(res, _ctx.result, sizeof(res));
        if (res[2] == 0) break;
        if (attempt > 4) throw std::runtime_error("AmdBroadPhase: the pair list does not fit the device workspace");
        // the reference never truncates: grow the output (bit 0) or the sweep's workspace (bit 2)
        // and sweep again from the same AABBs
        _cap = std::max(2 * _cap, res[0] + 1024);
    }
    const int m = res[0];
    _a.resize(m);
    _b.resize(m);
    _ctx.download(_a.data(), _ctx.pair_a, (size_t)m * 4);
    _ctx.download(_b.data(), _ctx.pair_b, (size_t)m * 4);
    pairs.clear();
    for (int k = 0; k < m; k++) {
        auto* a = objects[_a[k]];
        auto* b = objects[_b[k]];
        if (a->getGlobalId() == b->getGlobalId() || a->ignoreCollisionId(b->getGlobalId()) ||
            b->ignoreCollisionId(a->getGlobalId()))
            continue;
        pairs.emplace_back(a, b);
    }
    _axis = res[1];
}

// ------------------------------------------------------------------ narrowphase
AmdNarrowPhase::~AmdNarrowPhase() {
    for (auto* r : _owned) delete r;
}

void AmdNarrowPhase::calcContactResults(const pe::Array<pe_phys_collision::CollisionPair>& pairs,
                                        pe::Array<pe_phys_collision::ContactResult*>& results) {
    const int np = (int)pairs.size();
    // one ContactResult per pair, reused by index (simple_narrow_phase.cpp:10-22)
    while ((int)_owned.size() < np) _owned.push_back(new pe_phys_collision::ContactResult());
    results.assign(_owned.begin(), _owned.begin() + np);
    // the bodies of the pairs, numbered in order of first appearance
    std::unordered_map<const pe_phys_object::RigidBody*, int> index;
    std::vector<pe_phys_object::RigidBody*> bodies;
    std::vector<std::pair<pe_phys_object::RigidBody*, pe_phys_object::RigidBody*>> ab(np);
    _a.resize(np);
    _b.resize(np);
    for (int i = 0; i < np; i++) {
        auto* a = pairs[i].first;
        auto* b = pairs[i].second;
        if (b->getTag() == "wheel") std::swap(a, b);   // simple_narrow_phase.cpp:28
        ab[i] = {a, b};
        for (int s = 0; s < 2; s++) {
            auto* rb = s == 0 ? a : b;
            auto it = index.find(rb);
            int id;
            if (it == index.end()) {
                id = (int)bodies.size();
                index[rb] = id;
                bodies.push_back(rb);
            } else {
                id = it->second;
            }
            (s == 0 ? _a : _b)[i] = id;
        }
    }
    const int n = std::max(1, (int)bodies.size());
    _body.assign((size_t)n * PE_BODY_FLOATS, 0.0f);
    for (size_t i = 0; i < bodies.size(); i++) collision_record(bodies[i], _ctx, _body.data() + i * PE_BODY_FLOATS);
    pe_stage* st = _ctx.stage(n, std::max(np, _prev_pairs));
    // the slots keep the previous call's distances (the device copy of every result's cache)
    _ctx.points.reserve((size_t)std::max(np, _prev_pairs) * 8 * PE_CONTACT_FLOATS * 4, _ctx.stream());
    _ctx.upload(_ctx.body, _body.data(), _body.size() * 4);
    _ctx.upload(_ctx.pair_a, _a.data(), (size_t)np * 4);
    _ctx.upload(_ctx.pair_b, _b.data(), (size_t)np * 4);
    _ctx.npoints.reserve((size_t)std::max(np, 1) * 4, _ctx.stream());
    _ctx.swap.reserve((size_t)std::max(np, 1) * 4, _ctx.stream());
    _ctx.result.reserve(64, _ctx.stream());
    pe_check(pe_stage_narrowphase(st, n, (const float*)_ctx.body.ptr, np, (const int32_t*)_ctx.pair_a.ptr,
                                  (const int32_t*)_ctx.pair_b.ptr, _prev_pairs, (float*)_ctx.points.ptr,
                                  (int32_t*)_ctx.npoints.ptr, (int32_t*)_ctx.swap.ptr, (int32_t*)_ctx.result.ptr,
                                  _ctx.stream()),
             "pe_stage_narrowphase");
    int32_t ovf = 0;
    _ctx.download(&ovf, _ctx.result, 4);
    if (ovf) throw std::runtime_error("AmdNarrowPhase: a clipped contact polygon outgrew its buffer");
    _prev_pairs = np;
    _npts.resize(np);
    _swap.resize(np);
    _points.resize((size_t)np * 8 * PE_CONTACT_FLOATS);
    _ctx.download(_npts.data(), _ctx.npoints, (size_t)np * 4);
    _ctx.download(_swap.data(), _ctx.swap, (size_t)np * 4);
    _ctx.download(_points.data(), _ctx.points, _points.size() * 4);
    for (int i = 0; i < np; i++) {
        auto* cr = results[i];
        cr->clearContactPoints();
        cr->setObjects(ab[i].first, ab[i].second);
        for (int k = 0; k < PE_CONTACT_SLOTS; k++) {
            const float* c = _points.data() + ((size_t)i * 8 + k) * PE_CONTACT_FLOATS;
            auto& cp = cr->getContactPoint(k);
            if (k < _npts[i]) {
                cp = pe_phys_collision::ContactPoint(pe::Vector3(c[0], c[1], c[2]), pe::Vector3(c[3], c[4], c[5]),
                                                     ab[i].first, ab[i].second, pe::Vector3(c[6], c[7], c[8]),
                                                     pe::Vector3(c[9], c[10], c[11]), c[12]);
            } else {
                cp.invalidate();
                cp.setDistance(c[12]);
            }
        }
        cr->sortContactPoints();   // the slots are already in sorted order: this sets the point count
        cr->setSwapFlag(_swap[i] != 0);
    }
    // neighbour counts for the sleep test (simple_narrow_phase.cpp:43-62)
    for (int i = 0; i < np; i++) {
        auto* cr = results[i];
        if (cr->getPointSize() == 0) continue;
        auto* a = cr->getObjectA();
        auto* b = cr->getObjectB();
        if (!b->isKinematic()) {
            if (a->isSleep() || a->isKinematic()) b->incStaticCount();
            else b->incDynamicCount();
        }
        if (!a->isKinematic()) {
            if (b->isSleep() || b->isKinematic()) a->incStaticCount();
            else a->incDynamicCount();
        }
    }
}

// ------------------------------------------------------------------ solver
void AmdSequentialImpulseSolver::setupSolver(pe::Real dt, const pe::Vector3& gravity,
                                             const pe::Array<pe_phys_object::RigidBody*>& objects,
                                             const pe::Array<pe_phys_collision::ContactResult*>& contact_results,
                                             const pe::Array<pe_phys_constraint::Constraint*>& constraints) {
    (void)gravity;
    if (!constraints.empty()) throw std::runtime_error("AmdSequentialImpulseSolver: joint constraints are not supported");
    _dt = dt;
    _objects = objects;
    const int n = (int)objects.size();
    std::unordered_map<const pe_phys_object::RigidBody*, int> index;
    _body.assign((size_t)std::max(n, 1) * PE_BODY_FLOATS, 0.0f);
    _solver.assign((size_t)std::max(n, 1) * PE_SOLVER_BODY_FLOATS, 0.0f);
    for (int i = 0; i < n; i++) {
        const auto* rb = objects[i];
        index[rb] = i;
        collision_record(rb, _ctx, _body.data() + (size_t)i * PE_BODY_FLOATS);
        float* s = _solver.data() + (size_t)i * PE_SOLVER_BODY_FLOATS;
        s[SB_IM] = f32(rb->getKinematicInvMass());
        s[SB_KIN] = rb->isKinematic() ? 1.0f : 0.0f;
        const pe::Matrix3 iw = rb->getKinematicWorldInvInertia();
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) s[SB_IW + 3 * r + c] = f32(iw(r, c));
        for (int c = 0; c < 3; c++) {
            s[SB_V + c] = f32(rb->getLinearVelocity()[c]);
            s[SB_W + c] = f32(rb->getAngularVelocity()[c]);
        }
    }
    _npairs = (int)contact_results.size();
    _a.resize(_npairs);
    _b.resize(_npairs);
    _npts.resize(_npairs);
    _points.assign((size_t)_npairs * 8 * PE_CONTACT_FLOATS, 0.0f);
    for (int p = 0; p < _npairs; p++) {
        const auto* cr = contact_results[p];
        _a[p] = index.at(cr->getObjectA());
        _b[p] = index.at(cr->getObjectB());
        _npts[p] = cr->getPointSize();
        for (int k = 0; k < _npts[p]; k++) {
            const auto& cp = cr->getContactPoint(k);
            float* o = _points.data() + ((size_t)p * 8 + k) * PE_CONTACT_FLOATS;
            for (int c = 0; c < 3; c++) {
                o[c] = f32(cp.getWorldPos()[c]);
                o[3 + c] = f32(cp.getWorldNormal()[c]);
                o[6 + c] = f32(cp.getLocalPosA()[c]);
                o[9 + c] = f32(cp.getLocalPosB()[c]);
            }
            o[12] = f32(cp.getDistance());
        }
    }
}

void AmdSequentialImpulseSolver::solve() {
    const int n = std::max(1, (int)_objects.size());
    pe_stage* st = _ctx.stage(n, _npairs);
    auto up = [&](DeviceBuffer& b, const void* src, size_t bytes) { _ctx.upload(b, src, bytes); };
    up(_ctx.body, _body.data(), _body.size() * 4);
    up(_ctx.solver, _solver.data(), _solver.size() * 4);
    up(_ctx.pair_a, _a.data(), _a.size() * 4);
    up(_ctx.pair_b, _b.data(), _b.size() * 4);
    up(_ctx.npoints, _npts.data(), _npts.size() * 4);
    up(_ctx.solve_points, _points.data(), _points.size() * 4);
    _ctx.result.reserve(64, _ctx.stream());
    pe_check(pe_stage_solve(st, (int32_t)_objects.size(), (const float*)_ctx.body.ptr, (float*)_ctx.solver.ptr, _npairs,
                            (const int32_t*)_ctx.pair_a.ptr, (const int32_t*)_ctx.pair_b.ptr,
                            (const int32_t*)_ctx.npoints.ptr, (const float*)_ctx.solve_points.ptr, f32(_dt),
                            getIteration(), nullptr, (int32_t*)_ctx.result.ptr, _ctx.stream()),
             "pe_stage_solve");
    _ctx.download(_solver.data(), _ctx.solver, _solver.size() * 4);
    // syncTempVelocity (rigidbody.cpp:154-158): kinematic bodies keep theirs
    for (size_t i = 0; i < _objects.size(); i++) {
        auto* rb = _objects[i];
        if (rb->isKinematic()) continue;
        const float* s = _solver.data() + i * PE_SOLVER_BODY_FLOATS;
        const pe::Vector3 v(s[SB_V], s[SB_V + 1], s[SB_V + 2]), w(s[SB_W], s[SB_W + 1], s[SB_W + 2]);
        rb->setTempLinearVelocity(v);
        rb->setTempAngularVelocity(w);
        rb->setLinearVelocity(v);
        rb->setAngularVelocity(w);
    }
}

}  // namespace pe_amd_integration
