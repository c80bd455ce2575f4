/*
 * pe_amd.h -- C ABI of the MI355X (gfx950) rigid-body step.
 *
 * One world's step is the reference engine's World::step pipeline
 * (src/intf/world.cpp:267-300): integrate + gravity, sweep-and-prune broadphase,
 * narrowphase with the 8-slot contact cache, sequential-impulse friction-contact solve.
 * Every entry point returns 0 on success or a negative PE_ERR_* code; the message of the
 * last error on the calling thread is available from pe_last_error().  No C++ exception
 * crosses this boundary.  Handles are opaque; a handle may be used by one host thread at
 * a time ("NOT thread-safe", as src/intf/world.h:21 says of World).
 *
 * All device state is FP32 structure-of-arrays and stays resident in HBM between calls.
 */
#ifndef PE_AMD_H
#define PE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define PE_OK               0
#define PE_ERR_INVALID    (-1)  /* bad argument (null pointer, size out of range) */
#define PE_ERR_DEVICE     (-2)  /* HIP runtime error */
#define PE_ERR_CAPACITY   (-3)  /* a fixed device capacity (pairs, contacts) overflowed */
#define PE_ERR_UNSUPPORTED (-4) /* shape pair or feature not available */
#define PE_ERR_BUSY       (-5)  /* server: no slot or HBM budget for this request now */
#define PE_ERR_NOT_FOUND  (-6)  /* server: unknown request id */

/* ------------------------------------------------------------------ shapes
 * Values of pe_phys_shape::ShapeType (src/rigid/shape/shape.h:7-9). */
#define PE_SHAPE_BOX       0
#define PE_SHAPE_SPHERE    1
#define PE_SHAPE_CYLINDER  2
#define PE_SHAPE_CONVEX    3
#define PE_SHAPE_CONCAVE   4
#define PE_SHAPE_COMPOUND  5   /* children of box / sphere / cylinder / convex shapes (single world) */

/* A polygon mesh as a reference user hands it to ConvexMeshShape::setMesh or
 * ConcaveMeshShape::setMesh (pe::Mesh, common/mesh.h: vertex positions, faces as vertex index
 * lists with unit normals).  The library runs setMesh itself: a convex body's mesh is moved to
 * its centre of mass and its faces are reordered by the BVH build
 * (convex_mesh_shape.cpp:12-132); a concave body's mesh keeps its frame (concave_mesh_shape.cpp:6-17).
 * Limits of the device narrowphase, checked here (PE_ERR_INVALID beyond them, where the reference
 * has none): at most 32 vertices per face, and a BVH at most 46 levels deep (always true below
 * 2^44 faces: the tree is a median split). */
typedef struct pe_mesh_desc {
    int32_t nv;                       /* vertices */
    const double* vertices;           /* [3 nv] */
    int32_t nf;                       /* faces */
    const int32_t* face_start;        /* [nf + 1] offsets into face_index */
    const int32_t* face_index;        /* [face_start[nf]] vertex indices, counter-clockwise seen from outside */
    const double* face_normal;        /* [3 nf] outward unit normals */
} pe_mesh_desc;

/* Scene description (host memory, structure of arrays, FP64 so every consumer rounds the
 * same value to its own precision).  Mirrors what the reference sets on a RigidBody before
 * World::addRigidBody (src/rigid/object/rigidbody.h): setMass, setCollisionShape,
 * setTransform, setLinearVelocity, setAngularVelocity, setFrictionCoeff,
 * setRestitutionCoeff, setLinearDamping, setAngularDamping, setKinematic.  Bodies that name
 * the same mesh share one shape, as RigidBody::setCollisionShape allows. */
typedef struct pe_scene_desc {
    int32_t n;                        /* number of bodies; body i is the i-th addRigidBody */
    const int32_t* shape_type;        /* [n] PE_SHAPE_* */
    const double* shape_param;        /* [4n] box: size x,y,z (BoxShape(size)); sphere: radius;
                                         cylinder: radius, height; convex / concave: unused */
    const double* mass;               /* [n] */
    const double* position;           /* [3n] transform origin */
    const double* rotation;           /* [9n] transform basis, row-major */
    const double* linear_velocity;    /* [3n] */
    const double* angular_velocity;   /* [3n] */
    const double* friction;           /* [n] */
    const double* restitution;        /* [n] */
    const double* linear_damping;     /* [n] or NULL (0) */
    const double* angular_damping;    /* [n] or NULL (0) */
    const int32_t* kinematic;         /* [n] 0/1 */
    int32_t n_meshes;                 /* meshes of convex / concave bodies (0 if none) */
    const pe_mesh_desc* meshes;       /* [n_meshes] */
    const int32_t* mesh_index;        /* [n] mesh of body i if it is convex / concave; NULL if n_meshes = 0 */
    const int32_t* ignore_collision;  /* [n] RigidBody::setIgnoreCollision (0/1: the body is in no pair), or NULL */
    const double* lifetime;           /* [n] RigidBody::setLifeTime, seconds (the body leaves the world once it
                                         has been stepped that long; >= 1e30 never), or NULL (never) */
    const int32_t* compound_start;    /* [n + 1] or NULL: the children of compound body i are
                                         compound_children[compound_start[i] .. compound_start[i + 1]) */
    const struct pe_compound_child* compound_children;
} pe_scene_desc;

/* A child of a compound body (CompoundShape::addShape, compound_shape.cpp:7-20): a box, sphere,
 * cylinder or convex mesh at a transform relative to the body, with its mass ratio (> 1e-5).  The
 * library applies CompoundShape::init (compound_shape.cpp:24-47) as pe_world_create runs: the
 * children move so that their mass-ratio-weighted centroid is the body's origin, and the body's
 * inertia is the children's (each about its own centre, shifted, weighted by its ratio) times the
 * body's mass.  The body's transform is taken as given (the reference's compound demo sets the
 * transform's origin minus the centroid init returns, examples/compound_demo/main.cpp:77-78). */
typedef struct pe_compound_child {
    int32_t shape_type;               /* PE_SHAPE_BOX / SPHERE / CYLINDER / CONVEX */
    int32_t mesh_index;               /* convex: the scene mesh (pe_scene_desc.meshes) */
    double shape_param[3];            /* as pe_scene_desc.shape_param */
    double position[3];               /* local transform origin */
    double rotation[9];               /* local transform basis, row-major */
    double mass_ratio;
} pe_compound_child;

/* Joint constraints (src/rigid/constraint/constraint/{ball,hinge,slider}_joint_constraint.h), as a
 * reference user configures them before World::addConstraint (world.h:90): the two bodies
 * (Constraint::setObjectA / setObjectB), the body-local anchors and axes, and for a hinge its
 * angle limits and motor.  Values of pe_phys_constraint::ConstraintType. */
#define PE_JOINT_BALL   0
#define PE_JOINT_HINGE  1
#define PE_JOINT_SLIDER 2
typedef struct pe_joint_desc {
    int32_t type;                     /* PE_JOINT_* */
    int32_t body_a, body_b;           /* body indices */
    int32_t use_limits, use_motor;    /* hinge: setUseLimits, setUseMotor */
    int32_t reserved;
    double anchor_a[3], anchor_b[3];  /* setAnchorA / setAnchorB (body-local) */
    double axis_a[3], axis_b[3];      /* setAxisA / setAxisB (hinge, slider; body-local) */
    double min_angle, max_angle;      /* hinge: setMinAngle / setMaxAngle (radians) */
    double target_speed;              /* hinge: setTargetSpeed */
} pe_joint_desc;

/* World settings: World::setDt / setGravity (src/intf/world.h:24-25) and
 * SequentialImpulseSolver::setIteration (sequential_impulse_solver.cpp:8). */
typedef struct pe_world_params {
    double dt;
    double gravity[3];
    int32_t iterations;
} pe_world_params;

/* Layout of one contact point as returned by pe_world_get_contacts (13 floats):
 * world_pos[3] (on B), world_normal[3] (B -> A), local_pos_a[3], local_pos_b[3], distance
 * -- the fields of pe_phys_collision::ContactPoint (contact_result.h:10-74). */
#define PE_CONTACT_FLOATS 13
#define PE_CONTACT_SLOTS  8   /* PE_CONTACT_CACHE_SIZE, contact_result.h:6 */

const char* pe_last_error(void);
int pe_device_count(int32_t* count);

/* ConvexMeshShape::setMesh (concave = 0, convex_mesh_shape.cpp:12-132) or
 * ConcaveMeshShape::setMesh (concave = 1, concave_mesh_shape.cpp:6-17) as the library applies
 * it to a body's mesh, on the host (no device needed).  counts[4] receives the prepared
 * shape's vertex, face, face-index and unique-edge counts; each other output may be NULL or
 * hold: vertices [3 nv] (moved to the centre of mass for a convex mesh), face_start [nf + 1],
 * face_index, face_normal [3 nf] (faces in the BVH build's order), edges [6 ne] (the unique
 * edges' end points), volume_inertia [10] (volume, then the unit-mass inertia row-major) and
 * relocation [3] (the centre of mass setMesh returns).  FP32, as the device uses them. */
int pe_mesh_prepare(const pe_mesh_desc* mesh, int32_t concave, int32_t* counts, float* vertices,
                    int32_t* face_start, int32_t* face_index, float* face_normal, float* edges,
                    float* volume_inertia, float* relocation);

/* ------------------------------------------------------------------ plain single world */
typedef struct pe_world pe_world;

/* Creates a device-resident world.  max_pairs bounds the broadphase output (0 = 16 * n). */
int pe_world_create(const pe_scene_desc* scene, const pe_world_params* params,
                    int32_t max_pairs, pe_world** out);
int pe_world_destroy(pe_world* w);

/* World::step (src/intf/world.cpp:267-300), nsteps times, on the world's stream. */
int pe_world_step(pe_world* w, int32_t nsteps);

/* pe_world_step enqueued on the caller's stream (a hipStream_t, NULL for the null stream; the
 * type is declared below) instead of the world's own: the steps run after the caller's earlier
 * work on that stream and after the world's pending work, and the world's later calls (state
 * getters, pe_world_sync, further steps) wait for them.  Nothing is synchronised here. */
typedef void* pe_stream_t;
int pe_world_step_async(pe_world* w, int32_t nsteps, pe_stream_t stream);

/* Stage entry points; one call of each, in this order, is one World::step. */
/* World::updateObjectStatus + applyExternalForce (world.cpp:35-127) */
int pe_world_integrate(pe_world* w);
/* BroadPhaseBase::calcCollisionPairs, sweep-and-prune (broad_phase_sweep_and_prune.cpp:6-41).
 * With npairs non-null it waits for the stage and reports capacity errors as pe_world_sync does. */
int pe_world_broadphase(pe_world* w, int32_t* npairs);
/* NarrowPhaseBase::calcContactResults (simple_narrow_phase.cpp:7-63) */
int pe_world_narrowphase(pe_world* w);
/* ConstraintSolver::setupSolver + solve (sequential_impulse_solver.cpp:11-81) */
int pe_world_solve(pe_world* w);

/* World::addConstraint (world.h:90) for each of n joints (appended to the world's joints, in this
 * order): they are set up with the contacts and swept after them in every Gauss-Seidel iteration
 * (sequential_impulse_solver.cpp:49-51, 66-68), in an order grouped by a greedy colouring of the joints
 * (joint k takes the lowest colour no earlier joint sharing a dynamic body of it holds; joints of one
 * colour share no dynamic body, so the sweep equals the sequential one over that order).  At most
 * 65536 joints per world. */
int pe_world_add_joints(pe_world* w, int32_t n, const pe_joint_desc* joints);
/* RigidBody::addIgnoreCollisionId for n body pairs [2n]: the two bodies of a pair never form a
 * broadphase pair (broad_phase_base.cpp:12-13), as for the bodies a joint connects. */
int pe_world_add_ignore_pairs(pe_world* w, int32_t n, const int32_t* pairs);

/* ---- sleeping, lifetimes, removal, collision reports (single world) ----
 * World::updateObjectStatus (world.cpp:35-119) with the sequential loop's rules: a body below the
 * sleep thresholds (scaled by (static + 1) / (static + dynamic + 1) neighbour counts of the last
 * narrowphase, simple_narrow_phase.cpp:43-62) for `time` seconds falls asleep (velocities zeroed) and
 * is neither moved nor damped until a velocity reaches its scaled threshold; a body leaves the world
 * once it has lived its lifetime (pe_scene_desc.lifetime; RigidBody::step, rigidbody.cpp:215-218).
 * The batched and served paths keep the reference's defaults (no sleeping, no lifetimes). */

/* World::setSleepLinVel2Threshold, setSleepAngVel2Threshold, setSleepTimeThreshold (world.h:26-28);
 * all zero (the reference's default) never puts a body to sleep. */
int pe_world_set_sleep(pe_world* w, double lin_vel2, double ang_vel2, double time);
/* World::removeRigidBody (world.cpp:223-231): the body leaves the sweep and is no longer stepped
 * (its last state stays readable).  Not for a body with joints. */
int pe_world_remove_body(pe_world* w, int32_t body);
/* Per body [n] (each pointer may be NULL): flags (bit 0 RigidBody::isSleep, bit 1 removed: its
 * lifetime ended or pe_world_remove_body), the static and dynamic neighbour counts of the last
 * narrowphase (cleared by the next integrate), RigidBody::getSleepTime.  Synchronises. */
int pe_world_get_body_status(pe_world* w, int32_t* flags, int32_t* static_count, int32_t* dynamic_count,
                             float* sleep_time);
/* RigidBody::addCollisionCallback (on = 1) / clearCollisionCallbacks (on = 0) for n bodies: from the
 * next narrowphase, each contact result with points of which one of them is a body reports one
 * pe_collision_event (World::execCollisionCallbacks, world.cpp:129-192). */
int pe_world_set_collision_reports(pe_world* w, int32_t n, const int32_t* bodies, int32_t on);
/* One contact result's report: the host calls body_a's callbacks with (body_a, body_b,
 * position + normal * depth, -normal, -velocity) and body_b's with (body_b, body_a, position, normal,
 * velocity), as World::execCollisionCallbacks does. */
typedef struct pe_collision_event {
    int32_t body_a, body_b;   /* the result's objects A and B (the pair's order) */
    float position[3];        /* mean contact point (world, on B) */
    float normal[3];          /* normalised sum of the contact normals (B -> A) */
    float depth;              /* mean contact distance */
    float velocity[3];        /* mean of A's minus B's velocity at the contact points' local positions */
} pe_collision_event;
/* The reports of the last narrowphase, in pair order: at most cap into out, *count = all of them.
 * Synchronises. */
int pe_world_get_collision_events(pe_world* w, int32_t cap, pe_collision_event* out, int32_t* count);

/* ---- raycast ----
 * DefaultRaycastSolver::performRaycast (default_raycast_solver.cpp:11-86) for each of nrays rays
 * against the world's bodies as they are now (boxes, spheres, cylinders, convex and concave meshes:
 * raycast_box.cpp, raycast_sphere.cpp, raycast_cylinder.cpp, raycast_mesh.cpp).  rays is host memory,
 * [nrays][7]: start (3), direction (3), length; ignore (host, nignore body indices) is the call's
 * ignore set for every ray (the reference's ignores); ignore-collision and removed bodies are skipped.
 * As in the reference a query ends with hit = 0 at the first mesh body (in body order) whose AABB
 * the unbounded ray misses, the result then holding the best hit of the bodies before it.  Results
 * to host memory out[nrays].  Synchronises. */
typedef struct pe_ray_hit {
    int32_t hit;          /* performRaycast's return value */
    int32_t body;         /* RaycastResult::m_collisionObject's index, -1 if none */
    float distance;       /* m_distance (PE_REAL_MAX = 1e30 if none; may be negative: a sphere behind the start) */
    float point[3];       /* m_hitPoint */
    float normal[3];      /* m_hitNormal */
} pe_ray_hit;
int pe_world_raycast(pe_world* w, int32_t nrays, const float* rays, int32_t nignore, const int32_t* ignore,
                     pe_ray_hit* out);

/* Synchronises the world's stream, then reports capacity errors raised by the steps since the
 * last report (PE_ERR_CAPACITY, the causes in pe_last_error): more broadphase pairs than max_pairs
 * (the extra pairs are dropped), more sweep candidates than the sweep's chunk capacity, or a
 * clipped contact polygon beyond its buffer.  The flags are cleared once reported. */
int pe_world_sync(pe_world* w);

/* State access (FP32, device precision).  Any pointer may be NULL. */
int pe_world_get_state(pe_world* w, float* position /*3n*/, float* rotation /*9n row-major*/,
                       float* linear_velocity /*3n*/, float* angular_velocity /*3n*/);
int pe_world_set_state(pe_world* w, const float* position, const float* rotation,
                       const float* linear_velocity, const float* angular_velocity);
int pe_world_get_aabbs(pe_world* w, float* aabb_min /*3n*/, float* aabb_max /*3n*/);
/* Pairs of the last broadphase in emission order: (A, B) body indices. */
int pe_world_get_pairs(pe_world* w, int32_t* pairs /*2*cap*/, int32_t cap, int32_t* npairs);
/* Contact results of the last narrowphase, one per pair: point count, swap flag and the
 * 8 slots in cache order (PE_CONTACT_FLOATS floats each). */
int pe_world_get_contacts(pe_world* w, int32_t* npoints /*cap*/, float* points /*cap*8*13*/,
                          int32_t cap);
/* Accumulated impulses of the last solve, per pair and contact slot (slots in cache order, zero
 * beyond the point count): normal, tangent 0, tangent 1 -- ConstraintInfo::n_applied_impulse,
 * t0_applied_impulse, t1_applied_impulse (friction_contact_constraint.h:25-27). */
int pe_world_get_impulses(pe_world* w, float* impulses /*cap*8*3*/, int32_t cap);
/* Solve order of the last solve: pair indices grouped by colour, colour start offsets. */
int pe_world_get_solve_order(pe_world* w, int32_t* order /*cap*/, int32_t cap, int32_t* nordered,
                             int32_t* color_start /*ccap+1*/, int32_t ccap, int32_t* ncolors);
/* The contact cache the next narrowphase starts from: the pair count of the last narrowphase and
 * the distances of its pairs' 8 slots ([prev_npairs][8], stale ones included).  The reference
 * reuses its ContactResult objects by pair index and keeps invalid slots' distances
 * (simple_narrow_phase.cpp:10-22, contact_result.cpp:158-164), so these decide which points the
 * next step keeps; together with the body state (pe_world_get_state / set_state) and the SAP axis
 * they are everything a step reads from the steps before it, i.e. a checkpoint. */
int pe_world_get_cache(pe_world* w, float* distances /*8*cap*/, int32_t cap, int32_t* prev_npairs);
int pe_world_set_cache(pe_world* w, const float* distances /*8*prev_npairs*/, int32_t prev_npairs);
/* SAP axis that the next broadphase sorts on (BroadPhaseSweepAndPrune::_target_axis). */
int pe_world_get_axis(pe_world* w, int32_t* axis);
int pe_world_set_axis(pe_world* w, int32_t axis);
/* Cumulative device time per stage in ms: integrate, broadphase, narrowphase, solver --
 * the four timers of World (world.h:58-61).  enable = 1 records hipEvents around stages. */
int pe_world_set_timing(pe_world* w, int32_t enable);
int pe_world_get_stage_times(pe_world* w, double* ms /*4*/);
/* Workgroups of the grid-wide Gauss-Seidel sweep (0 = the default: 4 to 16 by body count, about one
 * colour's pair per thread; clamped to the co-resident limit).  No counterpart in the reference; results
 * do not depend on it. */
int pe_world_set_solver_blocks(pe_world* w, int32_t blocks);
/* How the solver's colouring runs (0 = by body count: one workgroup's claims in LDS up to 3 000
 * bodies, rounds over the GPU beyond; 1 = one workgroup, at most 12 288 bodies; 2 = rounds).  Both
 * give the same colour classes.  No counterpart in the reference; results do not depend on it. */
int pe_world_set_colouring(pe_world* w, int32_t mode);
/* The per-body records the stage entry points take (layouts under "stage entry points" below),
 * as the world's last integrate stage wrote them: AABB records [n][8], collision records [n][20],
 * solver records [n][24] (velocities after gravity, before the solve).  Any pointer may be NULL. */
int pe_world_get_records(pe_world* w, float* aabb, float* body, float* solver);
/* Counts of the last step: bodies, dynamic bodies, pairs, contact points, colours, rounds of the
 * colouring (0 when one workgroup coloured: see pe_world_set_colouring). */
int pe_world_get_counts(pe_world* w, int64_t* counts /*6*/);

/* ------------------------------------------------------------------ stage entry points
 * The reference's World calls three replaceable stages (World::setBroadPhase / setNarrowPhase /
 * setConstraintSolver, src/intf/world.h:97-100).  These entry points run one stage each on the
 * caller's device buffers, enqueued on the caller's stream (a hipStream_t; NULL is the null
 * stream) without synchronising, so a host that keeps the reference's World can replace a stage
 * at a time (the C++ classes implementing the stage interfaces over them are in INTEGRATION.md).
 * All buffers are device memory.  A pe_stage holds the working memory; one host thread uses a
 * pe_stage at a time.  Record layouts (FP32, 16-byte aligned rows):
 *   AABB record      [8]  min x y z, kinematic (0 / 1), max x y z, collision flag (0: normal,
 *                         1: setIgnoreCollision, in no pair; a world's records hold 2 for a body
 *                         with ignore partners, which the stage entry points treat as 0)
 *   collision record [20] basis (9, row-major), origin (3), AABB scale (RigidBody::getAABBScale),
 *                         shape parameters (3: box size / sphere radius / cylinder radius, height /
 *                         mesh id of pe_stage_set_meshes), shape type (PE_SHAPE_*, int bits),
 *                         kinematic (int bits), friction, restitution
 *   solver record    [24] temp velocity (3), temp angular velocity (3) (both ignored on input),
 *                         inverse mass, kinematic (0 / 1), world inverse inertia (9, row-major),
 *                         linear velocity (3), angular velocity (3), unused
 *   contact slot     [13] as pe_world_get_contacts */
#define PE_AABB_FLOATS 8
#define PE_BODY_FLOATS 20
#define PE_SOLVER_BODY_FLOATS 24

typedef struct pe_stage pe_stage;
/* Working memory for worlds of up to max_bodies bodies and max_pairs pairs. */
int pe_stage_create(int32_t max_bodies, int32_t max_pairs, pe_stage** out);
int pe_stage_destroy(pe_stage* s);
/* The meshes that collision records of convex / concave bodies name (mesh id = index).  With
 * prepared = 1 each mesh is taken as the reference's shape holds it after
 * ConvexMeshShape::setMesh / ConcaveMeshShape::setMesh (convex_mesh_shape.cpp:12-132: vertices
 * about the centre of mass, faces in the BVH build's order); with 0 setMesh is applied here, as
 * pe_world_create does.  Synchronous. */
int pe_stage_set_meshes(pe_stage* s, int32_t n_meshes, const pe_mesh_desc* meshes, const int32_t* concave,
                        int32_t prepared);
/* BroadPhaseSweepAndPrune::calcCollisionPairs (broad_phase_sweep_and_prune.cpp:6-41) over n AABB
 * records sorted on `axis`: pairs (A, B) in the reference's sweep order to pair_a / pair_b (at most
 * cap); result[3] (device) = pairs found (beyond cap when the output overflowed: then result[2] has
 * bit 0 and only cap pairs are written), the next axis, capacity-error bits (bit 2: more sweep
 * candidates than the workspace holds).  With n < 2 nothing is written but result, whose pair
 * count is 0 and axis unchanged (the reference leaves its pair list untouched then). */
int pe_stage_broadphase(pe_stage* s, int32_t n, const float* aabb, int32_t axis, int32_t* pair_a, int32_t* pair_b,
                        int32_t cap, int32_t* result, pe_stream_t stream);
/* SimpleNarrowPhase::calcContactResults (simple_narrow_phase.cpp:7-41): one contact result per
 * pair into points [npairs][8][13] (slots in cache order, invalid slots with the position at
 * 1e30 and their distance kept), npoints [npairs], swap [npairs] (ContactResult::swap_flag).
 * On input the distances of the slots of pairs below prev_npairs are the previous call's (the
 * reference reuses its ContactResult objects by pair index, contact_result.cpp:158-164).
 * result[1] (device) = capacity-error bits (bit 3: a clipped polygon outgrew its buffer). */
int pe_stage_narrowphase(pe_stage* s, int32_t n, const float* body, int32_t npairs, const int32_t* pair_a,
                         const int32_t* pair_b, int32_t prev_npairs, float* points, int32_t* npoints, int32_t* swap,
                         int32_t* result, pe_stream_t stream);
/* SequentialImpulseSolver::setupSolver + solve (sequential_impulse_solver.cpp:11-81) over the
 * contact results of the narrowphase, in colour order (pairs of one colour share no dynamic body;
 * the same as the reference's sequential sweep over the pairs grouped by colour): the solved
 * velocities go to the velocity fields of solver_body (kinematic bodies untouched), the
 * accumulated impulses to impulses [npairs][8][3] if not NULL; result[2] (device) = colours,
 * capacity-error bits. */
int pe_stage_solve(pe_stage* s, int32_t n, const float* body, float* solver_body, int32_t npairs, const int32_t* pair_a,
                   const int32_t* pair_b, const int32_t* npoints, const float* points, float dt, int32_t iterations,
                   float* impulses, int32_t* result, pe_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_H */
