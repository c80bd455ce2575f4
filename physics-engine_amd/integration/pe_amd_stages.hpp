// The reference engine's three replaceable World stages, implemented over this library's C ABI
// (include/pe_amd.h, "stage entry points"), for a host that keeps the reference's own World
// (src/intf/world.h) and swaps stages with World::setBroadPhase / setNarrowPhase /
// setConstraintSolver (world.h:97-100):
//
//     pe_amd_integration::DeviceContext ctx;                 // one per World
//     world.setBroadPhase(new pe_amd_integration::AmdBroadPhase(ctx));
//     world.setNarrowPhase(new pe_amd_integration::AmdNarrowPhase(ctx));
//     world.setConstraintSolver(new pe_amd_integration::AmdSequentialImpulseSolver(ctx));
//
// Each class marshals the RigidBody objects it is handed into the records the entry points take,
// runs the stage on the GPU and writes the results back where the reference's own stage leaves
// them (the pair array, the ContactResult objects, the bodies' velocities).  Compiled against the
// reference's headers (src/, ThirdParty/, SimpleViewer/include) and this repository's include/,
// linked with libpe_amd.so and the HIP runtime.  The arithmetic is FP32: with the reference built in
// FP32 (no PE_USE_DOUBLE) results match the reference's own stages as the device pipeline does.
// Errors of the C ABI are thrown as std::runtime_error, the reference's own convention
// (rigidbody.cpp:12-14).
#pragma once

#include "rigid/collision/broad_phase/broad_phase_base.h"
#include "rigid/collision/narrow_phase/narrow_phase_base.h"
#include "rigid/constraint/constraint_solver/constraint_solver.h"

#include "pe_amd.h"

#include <cstdint>
#include <unordered_map>
#include <vector>

namespace pe_amd_integration {

// A device buffer that grows on demand (keeping its contents).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t need, void* stream);
    void release();
};

// The device side one World's stages share: the stage workspace, a HIP stream and the buffers
// the records travel in.  Not thread-safe (neither is World, world.h:21).
class DeviceContext {
public:
    DeviceContext();
    ~DeviceContext();
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;

    // The workspace for at least n bodies and m pairs (recreated larger when needed; the meshes
    // are registered again then).
    pe_stage* stage(int n, int m);
    void* stream() const { return _stream; }
    void sync();
    void upload(DeviceBuffer& b, const void* src, size_t bytes);
    void download(void* dst, const DeviceBuffer& b, size_t bytes);
    // Mesh id of a convex / concave shape's mesh (as the shape holds it after setMesh), registering
    // it with the workspace on first sight.
    int32_t mesh_id(const pe_phys_shape::Shape* shape);

    // points carries the narrowphase's contact cache from one step to the next; the solver
    // stages its own copy of the contact points in solve_points
    DeviceBuffer aabb, pair_a, pair_b, result, body, solver, points, npoints, swap, solve_points;

private:
    void register_meshes();
    pe_stage* _stage = nullptr;
    int _max_bodies = 0, _max_pairs = 0;
    void* _stream = nullptr;
    std::unordered_map<const pe_phys_shape::Shape*, int32_t> _mesh_ids;
    std::vector<const pe_phys_shape::Shape*> _meshes;
    bool _meshes_dirty = false;
};

// BroadPhaseSweepAndPrune (broad_phase_sweep_and_prune.cpp:6-41) on the GPU: the same pairs, in
// the same sweep order, the same next axis.  Ties of the sort key are ordered by position in the
// object array (the reference's std::sort leaves them unspecified).  Pairs excluded by a body's
// ignore-collision id set (validateCollisionPair, broad_phase_base.cpp:12-13) are dropped from the
// GPU's list on the host, keeping the order of the others.
class AmdBroadPhase : public pe_phys_collision::BroadPhaseBase {
public:
    explicit AmdBroadPhase(DeviceContext& ctx) : _ctx(ctx) {}
    void calcCollisionPairs(pe::Array<pe_phys_object::RigidBody*> objects,
                            pe::Array<pe_phys_collision::CollisionPair>& pairs) override;
    int targetAxis() const { return _axis; }
    void setTargetAxis(int a) { _axis = a; }

private:
    DeviceContext& _ctx;
    int _axis = 0;
    int _cap = 1024;
    std::vector<float> _rec;
    std::vector<int32_t> _a, _b;
};

// SimpleNarrowPhase (simple_narrow_phase.cpp:7-63) on the GPU: one ContactResult per pair (owned
// here, reused by pair index as the reference reuses its pooled ones, with invalid slots keeping
// their distances), the "wheel" tag swap, and the static / dynamic neighbour counts the sleep
// test reads.
class AmdNarrowPhase : public pe_phys_collision::NarrowPhaseBase {
public:
    explicit AmdNarrowPhase(DeviceContext& ctx) : _ctx(ctx) {}
    ~AmdNarrowPhase() override;
    void calcContactResults(const pe::Array<pe_phys_collision::CollisionPair>& pairs,
                            pe::Array<pe_phys_collision::ContactResult*>& results) override;

private:
    DeviceContext& _ctx;
    std::vector<pe_phys_collision::ContactResult*> _owned;
    int _prev_pairs = 0;
    std::vector<float> _body, _points;
    std::vector<int32_t> _a, _b, _npts, _swap;
};

// SequentialImpulseSolver (sequential_impulse_solver.cpp:11-81) on the GPU, the pairs swept in
// colour order (pairs of one colour share no dynamic body; the sequential sweep over the pairs
// grouped by colour).  Joint constraints are not part of the device solver: a non-empty
// constraint array throws.
class AmdSequentialImpulseSolver : public pe_phys_constraint::ConstraintSolver {
public:
    explicit AmdSequentialImpulseSolver(DeviceContext& ctx) : _ctx(ctx) { _iteration = 10; }
    pe_phys_constraint::ConstraintSolverType getType() const override {
        return pe_phys_constraint::ConstraintSolverType::CST_SEQUENTIAL_IMPULSE;
    }
    void setupSolver(pe::Real dt, const pe::Vector3& gravity, const pe::Array<pe_phys_object::RigidBody*>& objects,
                     const pe::Array<pe_phys_collision::ContactResult*>& contact_results,
                     const pe::Array<pe_phys_constraint::Constraint*>& constraints) override;
    void solve() override;

private:
    DeviceContext& _ctx;
    pe::Real _dt = 0;
    pe::Array<pe_phys_object::RigidBody*> _objects;
    int _npairs = 0;
    std::vector<float> _body, _solver, _points;
    std::vector<int32_t> _a, _b, _npts;
};

// The collision record (PE_BODY_FLOATS floats) of a body, as the stage entry points take it.
void collision_record(const pe_phys_object::RigidBody* rb, DeviceContext& ctx, float* out);

}  // namespace pe_amd_integration
