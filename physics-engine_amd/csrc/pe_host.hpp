// Host-side set-up shared by the single-world, batched and served pipelines: scene validation,
// staging of bodies into the SoA store with their mass properties, and the polyhedral shape
// data (ConvexMeshShape::setMesh / ConcaveMeshShape::setMesh and the unit cylinder mesh).
#pragma once
#include "pe_common.hpp"
#include "../../include/pe_amd.h"

#include <map>
#include <vector>

namespace pe {

// Prepared meshes of one or more scenes, laid out as the device MeshSet expects (arena offsets
// are relative to the start of this arena; rebase() shifts them when the arena is placed at an
// offset inside a larger device arena).
struct HostMeshes {
    std::vector<int32_t> rec;      // [MRW per mesh]
    std::vector<float4> arena;
    std::vector<float> props;      // [10 per mesh] volume, unit-mass inertia (row-major)
    std::vector<float4> kids;      // compound children (CHW floats each)
    std::vector<int2> krange;      // per compound: first child, count
    int count() const { return (int)rec.size() / MRW; }
    // ConvexMeshShape::setMesh (concave = false) or ConcaveMeshShape::setMesh (concave = true)
    // of a mesh; returns its mesh id, or -1 (with pe_last_error set) if the mesh is malformed.
    // Optional outputs: relocation (the centre of mass the vertices were moved by) and the
    // volume followed by the unit-mass inertia (10 floats).  prepared = true takes a mesh as a
    // shape holds it after setMesh (vertices already about the centre of mass, faces already in
    // the BVH build's order): nothing is moved or reordered, the BVH and edges follow from it.
    int add(const pe_mesh_desc& md, bool concave, float* relocation = nullptr, float* volume_inertia = nullptr,
            bool prepared = false);
    void rebase(int32_t arena_offset);
    int nv(int m) const { return rec[MRW * m + MR_NV]; }
};

// The unit cylinder mesh (shape/default_mesh.h:59-273) in the MeshSet layout.
void cylinder_unit_table(float out[CYL_UNIT_FLOATS]);

// Validates a scene for the device pipelines; PE_OK or a PE_ERR_* code with pe_last_error set.
int check_scene(const pe_scene_desc* sc);

// Stages the bodies of one scene into the host copies of the body store (columns off .. off+n-1
// of a [NF][stride] float and [NI][stride] int store): pose, velocities, material, shape
// parameters and mass properties as RigidBody::setMass + setCollisionShape compute them
// (rigidbody.cpp:11-41).  The scene's meshes are prepared into `meshes` (one shape per
// mesh and kind, shared by the bodies that name it); a convex / concave body's mesh id is
// stored in its first shape parameter as int bits.  Returns PE_OK or a PE_ERR_* code.
int stage_bodies(const pe_scene_desc* sc, std::vector<float>& hf, std::vector<int32_t>& hi, size_t stride, int off,
                 HostMeshes& meshes);

// Device copy of a HostMeshes plus the unit cylinder table; view() is the MeshSet kernels take.
struct DevMeshes {
    int32_t* rec = nullptr;
    float4* arena = nullptr;
    float* cyl = nullptr;
    float4* kids = nullptr;
    int2* krange = nullptr;
    int upload(const HostMeshes& h, hipStream_t s);
    void release();
    MeshSet view() const { return MeshSet{rec, arena, cyl, kids, krange}; }
};

}  // namespace pe
