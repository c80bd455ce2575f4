// Host/device shared definitions: SoA field layout of the body store, error handling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace pe {

// Body store: one float array of NF fields x stride (stride = n rounded up to 64), field-major,
// so every per-field access of consecutive bodies is one coalesced load.
enum BodyField : int {
    F_POS = 0,     // 3: origin
    F_ROT = 3,     // 9: basis, row-major
    F_V = 12,      // 3: linear velocity
    F_W = 15,      // 3: angular velocity
    F_IW = 18,     // 9: world inverse inertia
    F_AMIN = 27,   // 3: AABB min
    F_AMAX = 30,   // 3: AABB max
    F_IM = 33,     // inverse mass
    F_MASS = 34,   // mass
    F_LINV = 35,   // 9: local inverse inertia
    F_SP = 44,     // 4: shape parameters (box size / sphere radius / cylinder radius, height / mesh id),
                   //    then the collision flag (COLL_*)
    F_MU = 48,     // friction
    F_E = 49,      // restitution
    F_LD = 50,     // linear damping
    F_AD = 51,     // angular damping
    F_TV = 52,     // 3: temp linear velocity (solver)
    F_TW = 55,     // 3: temp angular velocity (solver)
    F_SCALE = 58,  // AABB scale of the current pose
    NF = 59
};
enum BodyIntField : int { I_TYPE = 0, I_KIN = 1, NI = 2 };
// Collision flag of a body, in its fourth shape parameter and in the last float of its AABB record:
// COLL_IGNORE (RigidBody::setIgnoreCollision) puts it in no pair; COLL_PARTNERS marks a body that some
// other body ignores (addIgnoreCollisionId), whose candidate pairs are looked up in the world's
// ignore-pair list.
constexpr float COLL_NORMAL = 0.0f;
constexpr float COLL_IGNORE = 1.0f;
constexpr float COLL_PARTNERS = 2.0f;

// The narrowphase and the solver visit bodies, contact slots and rows in pair order, i.e. at
// random; for them each record is contiguous (array of structures, float4-aligned) so one
// body or one row costs one or two cache lines instead of one line per field.
//
// Collision-body record, written by the integrate stage: pose, AABB scale, shape, material.
enum CBField : int { CB_R = 0, CB_P = 9, CB_SCALE = 12, CB_SP = 13, CB_TYPE = 16, CB_KIN = 17, CB_MU = 18, CB_E = 19, CBW = 20 };
// Solver-body record: temp velocities (updated by the sweeps), inverse mass, kinematic flag,
// world inverse inertia, and the velocities the solver setup reads.
enum SBField : int { SB_TV = 0, SB_TW = 3, SB_IM = 6, SB_KIN = 7, SB_IW = 8, SB_V = 17, SB_W = 20, SBW = 24 };
// Contact slot record (8 per pair): world_pos (on B), normal (B -> A), local A, local B.
// Distances live in their own array (read for all 8 slots of a pair at the narrowphase start).
enum ContactField : int { C_WPOS = 0, C_N = 3, C_LA = 6, C_LB = 9, CW = 12, CF = 13 };
// Solver row record (one per contact slot), FrictionContactConstraint::ConstraintInfo
// (friction_contact_constraint.h:15-29).
enum RowField : int {
    R_RA = 0, R_RB = 3, R_N = 6, R_T0 = 9, R_T1 = 12, R_RHS = 15, R_MN = 16, R_MT0 = 17, R_MT1 = 18,
    R_MU = 19, R_LN = 20, R_LT0 = 21, R_LT1 = 22, RF = 23, RW = 24
};

// Polyhedral shape data of the convex / concave meshes of a world or batch (prepared on the
// host as ConvexMeshShape::setMesh / ConcaveMeshShape::setMesh do) in one float4 arena, one
// record of arena offsets per mesh, plus the unit cylinder mesh shared by all cylinders.
struct MeshSet {
    const int32_t* rec;   // [MRW per mesh]
    const float4* arena;
    const float* cyl;     // unit cylinder: top ring x[20] z[20], bottom ring x[20] z[20], side normals [40][3]
    const float4* kids = nullptr;    // compound children, CHW / 4 float4 each (a single world's compounds)
    const int2* krange = nullptr;    // per compound: first child, child count
};
// Compound child record (CompoundShape::ShapeEle after CompoundShape::init): shape type (int bits),
// shape parameters (a convex child's mesh id as int bits in the first), local basis (row-major), local
// origin (moved by the children's centroid).
enum ChildField : int { CH_TYPE = 0, CH_SP = 1, CH_R = 4, CH_P = 13, CHW = 16 };
// Mesh record: arena offsets (in float4 units) and counts.
//   vertices: float4 xyz;  face normals: float4 xyz;  faces: int2 (first index, count) per face;
//   face indices: int32, relative to the mesh's first vertex;  unique edges: 2 float4 (ends);
//   BVH nodes: 2 float4 ((aa.xyz, left), (bb.xyz, right), children as int bits);  leaves: int2
//   (face count, 0 = inner node; first face) per node.
enum MeshRec : int { MR_V = 0, MR_NV, MR_FN, MR_NF, MR_FACE, MR_FI, MR_E, MR_NE, MR_NODE, MR_NB, MR_LEAF, MR_PAD, MRW };
constexpr int CYL_UNIT_FLOATS = 80 + 120;
// Limits of the polyhedral narrowphase, checked when a mesh is prepared: a face has at most
// MAX_FACE_VERTS vertices (an incident face clipped by a reference face's side planes then has at
// most 2 * MAX_FACE_VERTS, the clip buffer), and the BVH walk's stack holds BVH_STACK entries (the
// median-split tree of F faces is about log2(F) deep; the walk needs at most depth + 2 entries).
constexpr int MAX_FACE_VERTS = 32;
constexpr int BVH_STACK = 48;

struct BodyView {
    float* f;          // [NF][stride] field-major state (coalesced per-body kernels)
    int32_t* i;        // [NI][stride]
    float* cb;         // [n][CBW] collision-body records
    float* sb;         // [n][SBW] solver-body records
    float* ab;         // [n][8] AABB records: min xyz, kinematic flag, max xyz, 0
    MeshSet ms;
    int n;
    int stride;
    __host__ __device__ float& at(int field, int b) const { return f[(size_t)field * stride + b]; }
    __host__ __device__ int32_t& iat(int field, int b) const { return i[(size_t)field * stride + b]; }
    __host__ __device__ float* cbr(int b) const { return cb + (size_t)b * CBW; }
    __host__ __device__ float* sbr(int b) const { return sb + (size_t)b * SBW; }
    __host__ __device__ float* abr(int b) const { return ab + (size_t)b * 8; }
};

}  // namespace pe

// ------------------------------------------------------------------ host error helpers
void pe_set_error(const std::string& msg);
// HIP keeps the last error of any runtime call per host thread until hipGetLastError reads it.  An
// entry point that checks hipGetLastError after its own launches first drops what earlier calls
// left there (their own return codes reported it), so it reports only its own launches.
inline void pe_clear_stale_error() { (void)hipGetLastError(); }
#define PE_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            pe_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return PE_ERR_DEVICE;                                                        \
        }                                                                                \
    } while (0)
