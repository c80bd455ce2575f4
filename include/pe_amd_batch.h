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
int pe_batch_sync(pe_batch* b);
/* Captures one step (one launch per size class) into a hipGraph used by pe_batch_step. */
int pe_batch_capture_graph(pe_batch* b);
/* Size classes: class c holds worlds with upper[c-1] < n <= upper[c] and launches with
 * block[c] threads (64, 128, 256, 512 or 1024).  Drops a captured graph. */
int pe_batch_set_config(pe_batch* b, int32_t nclasses, const int32_t* upper, const int32_t* block);
int pe_batch_get_config(pe_batch* b, int32_t* nclasses, int32_t* upper, int32_t* block, int32_t* count, int32_t cap);
int pe_batch_get_world_state(pe_batch* b, int32_t world, float* position, float* rotation,
                             float* linear_velocity, float* angular_velocity);
/* bodies, dynamic bodies, pairs, contact points, max colours, worlds with overflow, worlds */
int pe_batch_get_counts(pe_batch* b, int64_t* counts /*7*/);
/* Device bytes one world of this scene occupies (bodies, pair slots, solver rows). */
int pe_batch_world_footprint(const pe_scene_desc* scene, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_BATCH_H */
