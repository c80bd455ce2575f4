/*
 * pe_amd_noalloc.h -- the sweep-and-prune stage for callers that cannot allocate device memory.
 */
#ifndef PE_AMD_NOALLOC_H
#define PE_AMD_NOALLOC_H

#include "pe_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PE_NOALLOC_MAX_BODIES 65536
#define PE_NOALLOC_SCRATCH_BYTES (10 << 20)

/* No counterpart in the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): pe_stage_broadphase without a
 * pe_stage, the same kernels and results (same arguments and result[3]), its working memory a
 * fixed device array of PE_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  It calls
 * no allocator.  n above PE_NOALLOC_MAX_BODIES is refused (PE_ERR_INVALID) before anything is
 * enqueued; more than 32 Mi sweep candidates sets bit 2 of result[2] as the workspace overflow of
 * pe_stage_broadphase does.  Every call on one GPU shares that GPU's array: calls that overlap in
 * time on one GPU (different streams or threads) race; calls in order on one stream do not. */
int pe_stage_broadphase_noalloc(int32_t n, const float* aabb, int32_t axis, int32_t* pair_a, int32_t* pair_b,
                                int32_t cap, int32_t* result, pe_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_NOALLOC_H */
