/*
 * pe_amd_batch.h -- batched and served stepping of many independent worlds on one MI355X.
 *
 * These entry points have no counterpart in the reference engine, which steps one World at a
 * time on a CPU thread pool (src/intf/world.cpp:267-300).  Each world here is stepped by the
 * same algorithm as pe_world_step (include/pe_amd.h), one workgroup per world, so a world's
 * trajectory does not depend on what else is in the batch.
 */
#ifndef PE_AMD_BATCH_H
#define PE_AMD_BATCH_H

#include "pe_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PE_BATCH_MAX_WORLD_BODIES 4096

/* ------------------------------------------------------------------ batch */
typedef struct pe_batch pe_batch;

/* nworlds independent worlds, all with the same dt / gravity / iterations. */
int pe_batch_create(int32_t nworlds, const pe_scene_desc* scenes, const pe_world_params* params, pe_batch** out);
int pe_batch_destroy(pe_batch* b);
/* nsteps World::step of every world (replays the captured hipGraph when one exists). */
int pe_batch_step(pe_batch* b, int32_t nsteps);
/* Waits for the batch's work, then reports capacity errors raised since the last report: a world
 * whose broadphase found more pairs than its capacity (16 per body + 64; the extra pairs are
 * dropped) or whose clipper lost contact points returns PE_ERR_CAPACITY naming the worlds.  Each
 * world's flag stays set (pe_batch_get_counts counts them). */
int pe_batch_sync(pe_batch* b);
/* Captures one step (one launch per size class) into a hipGraph used by pe_batch_step. */
int pe_batch_capture_graph(pe_batch* b);
/* Size classes: class c holds worlds with upper[c-1] < n <= upper[c] and launches with
 * block[c] threads (64, 128, 256, 512 or 1024); at most 32 classes.  Each class's integrate +
 * broadphase, narrowphase and solver launch in sequence on the stream it was dealt, and the
 * classes of different streams overlap.  Drops a captured graph. */
int pe_batch_set_config(pe_batch* b, int32_t nclasses, const int32_t* upper, const int32_t* block);
/* Solver grouping per size class: the solve kernel takes consecutive worlds of a class (worlds
 * are laid out largest first) up to group_bodies[c] bodies into one workgroup and solves them as
 * one system, bodies, pairs and colouring in LDS; 0 solves one world per workgroup.  Worlds
 * never share a body, so every world's result is the one it gets alone.  One value per class
 * (0..2048).  Drops a captured graph. */
int pe_batch_set_solver_groups(pe_batch* b, int32_t nclasses, const int32_t* group_bodies);
int pe_batch_get_solver_groups(pe_batch* b, int32_t* nclasses, int32_t* group_bodies, int32_t cap);
int pe_batch_get_config(pe_batch* b, int32_t* nclasses, int32_t* upper, int32_t* block, int32_t* count, int32_t cap);
int pe_batch_get_world_state(pe_batch* b, int32_t world, float* position, float* rotation,
                             float* linear_velocity, float* angular_velocity);
/* bodies, dynamic bodies, pairs, contact points, max colours, worlds that ever exceeded a capacity, worlds */
int pe_batch_get_counts(pe_batch* b, int64_t* counts /*7*/);
/* Device bytes one world of this scene occupies (bodies, pair slots, solver rows). */
int pe_batch_world_footprint(const pe_scene_desc* scene, int64_t* bytes);

/* Runs nsteps of the batch and returns the device time of those steps (hipEvents recorded on
 * the batch's stream around them), in milliseconds. */
int pe_batch_step_timed(pe_batch* b, int32_t nsteps, double* ms);

/* Runs one step recording per-world phase timestamps; us[k] is the sum over worlds of phase k's
 * duration in microseconds: integrate, sort, sweep, narrowphase, solver setup, colouring,
 * Gauss-Seidel sweeps.  Drops a captured graph. */
int pe_batch_phase_profile(pe_batch* b, double* us /*8*/);
/* The same step's raw timestamps, ts[world][8] in ticks of 10 ns (0 where not recorded): start,
 * integrate, sort, sweep done (K1); solver start, rows set up, coloured, swept (K3; stamped on the
 * first world of each solver group only).  No counterpart in the reference; for tuning. */
int pe_batch_phase_trace(pe_batch* b, uint64_t* ts /*8 nworlds*/);

/* Runs nsteps eagerly, staged for measurement (every class's integrate + broadphase, then every
 * class's narrowphase, then every class's solver), with a hipEvent pair around every stage;
 * ms[k] is the summed device time of integrate + broadphase (k = 0, all size classes),
 * narrowphase (1, all classes) and solver (2, all size classes), each from before its first
 * launch to after its last (the size classes of stages 0 and 2 run concurrently on side streams),
 * without the gaps between stages. */
int pe_batch_stage_times(pe_batch* b, int32_t nsteps, double* ms /*3*/);

/* Runs nsteps eagerly with every size class launched alone, in turn; ms[2c] and ms[2c + 1] are
 * class c's summed integrate + broadphase + narrowphase and solver device milliseconds
 * (cap >= 2 * classes).  Each class's per-step total then sets where it runs in concurrent
 * steps: classes are dealt, longest first, onto the least-loaded of the batch's streams.  Drops
 * a captured graph. */
int pe_batch_class_times(pe_batch* b, int32_t nsteps, double* ms, int32_t cap);

/* No counterpart in the reference (it has no tuner): pe_batch_snapshot copies the whole device
 * state of every world (bodies, contact caches, solver rows, per-world counters) aside, on the
 * device; pe_batch_restore puts it back, so the next step is exactly the one that followed the
 * snapshot.  A tuner times its candidates from one snapshot and restores it afterwards, so tuning
 * neither advances the worlds nor times candidates on different states.  The first snapshot
 * allocates a second copy of the batch's state; the launch configuration is not part of it. */
int pe_batch_snapshot(pe_batch* b);
int pe_batch_restore(pe_batch* b);

/* ------------------------------------------------------------------ server
 * Continuous batching of world-simulation requests.  A request (a scene and a horizon in
 * steps) is queued by pe_server_submit.  Each pe_server_tick admits queued requests in
 * arrival order while their device footprint fits the HBM budget and arena space, runs one
 * step of every live world in one batched launch, and retires worlds that reached their
 * horizon.  Uploads of admitted requests are issued on a copy stream while the current step
 * runs.  Results are the final body states, bit-identical to stepping the same scene alone. */
typedef struct pe_server pe_server;
typedef struct pe_server_config {
    int64_t hbm_budget_bytes;    /* device bytes for world state (bodies, pairs, rows, staging) */
    int32_t max_worlds;          /* concurrent live worlds */
    int32_t max_world_bodies;    /* largest admissible request (<= PE_BATCH_MAX_WORLD_BODIES) */
    pe_world_params params;      /* dt, gravity, iterations shared by all requests */
} pe_server_config;

int pe_server_create(const pe_server_config* cfg, pe_server** out);
int pe_server_destroy(pe_server* s);
/* Device bytes a request of this scene holds while live. */
int pe_server_footprint(const pe_scene_desc* scene, int64_t* bytes);
/* Queues a request; returns PE_ERR_BUSY if it can never fit (larger than the budget). */
int pe_server_submit(pe_server* s, const pe_scene_desc* scene, int32_t horizon, int64_t* request_id);
/* One continuous-batching step; reports live worlds and bodies after admission. */
int pe_server_tick(pe_server* s, int32_t* live_worlds, int64_t* live_bodies);
/* Ids of requests that finished since the last poll. */
int pe_server_poll(pe_server* s, int64_t* ids, int32_t cap, int32_t* count);
/* Final state of a finished request (n = its body count); frees the stored result.  A request
 * whose world exceeded a capacity while stepping (more pairs than 16 per body + 64, or a clipped
 * contact polygon beyond its buffer) returns PE_ERR_CAPACITY and no state. */
int pe_server_result(pe_server* s, int64_t id, float* position, float* rotation, float* linear_velocity,
                     float* angular_velocity);
/* Drops a queued or live request: a queued one at once, a live one (or one whose upload is in
 * flight) at the next tick, which reclaims its slot and memory.  A finished request is not
 * cancelled (PE_ERR_INVALID); its result stays available. */
int pe_server_cancel(pe_server* s, int64_t id);
/* ticks, p50 tick ms, p99 tick ms, mean tick ms, body-steps done, live worlds, queued, admitted.
 * A tick steps every live world once, so the tick time is each live world's step latency. */
int pe_server_stats(pe_server* s, double* out /*8*/);
/* Per-request figures: p50 and p99 admission wait (submission to admission, ms), p50 and p99
 * latency (submission to result available, ms), mean and maximum queue depth over ticks,
 * requests finished, requests cancelled, results refused with PE_ERR_CAPACITY. */
int pe_server_request_stats(pe_server* s, double* out /*9*/);

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_BATCH_H */
