// Device primitives shared by the single-world and batched pipelines: wave/block scans,
// a stable LSD radix sort of (float key, index) pairs, single-block exclusive scans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pe {

constexpr int WAVE = 64;

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & (WAVE - 1);
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Exclusive block-wide scan of one int per thread.  `wsum` is LDS scratch of blockDim/64 + 1
// ints.  Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int* wsum, int* total) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    int x = v;
    // inclusive wave scan (Hillis-Steele over 64 lanes)
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) wsum[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; w++) { const int t = wsum[w]; wsum[w] = acc; acc += t; }
        wsum[nw] = acc;
    }
    __syncthreads();
    const int r = wsum[wid] + x - v;
    *total = wsum[nw];
    __syncthreads();
    return r;
}

// Float -> order-preserving uint32.  Callers add +0.0f first so -0 and +0 map alike.
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// The float float_key was made from.
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// ---------------------------------------------------------------- radix sort (multi-block)
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

// Per-tile histogram of one 8-bit digit, written digit-major: hist[d * nblocks + b].
static __global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                           uint32_t* __restrict__ hist, int nblocks) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; j++) {
        const int e = base + j * RS_THREADS + threadIdx.x;
        if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of `n` uint32 in place, one block of 1024 threads.
static __global__ void __launch_bounds__(1024) k_scan_u32(uint32_t* __restrict__ data, int n, uint32_t* __restrict__ total) {
    __shared__ int wsum[17];
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int e = base + threadIdx.x;
        const int v = e < n ? (int)data[e] : 0;
        int tot;
        const int ex = block_exclusive_scan(v, wsum, &tot);
        if (e < n) data[e] = (uint32_t)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = (uint32_t)carry;
}

// Grid-wide exclusive scan of n uint32 in place (n = *count when count is given, read on the
// device), in three launches: k_scan_tiles scans each tile of SCAN_TILE elements and writes the
// tile's total to partial[tile]; k_scan_u32 scans the partials (and writes the grand total);
// k_scan_add adds each tile's offset.  Blocks past n do nothing (their partial is 0).
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = 1024 * SCAN_ITEMS;
static __global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t* __restrict__ data, int n_static,
                                                           const uint32_t* __restrict__ count, uint32_t* __restrict__ partial) {
    __shared__ int wsum[17];
    const int n = count ? (int)*count : n_static;
    const int base = blockIdx.x * SCAN_TILE;
    if (base >= n) {   // the same for every thread of the block
        if (threadIdx.x == 0) partial[blockIdx.x] = 0;
        return;
    }
    const int e0 = base + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = e0 + k < n ? data[e0 + k] : 0u;
        sum += (int)v[k];
    }
    int tot;
    int run = block_exclusive_scan(sum, wsum, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (e0 + k < n) data[e0 + k] = (uint32_t)run;
        run += (int)v[k];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = (uint32_t)tot;
}
static __global__ void __launch_bounds__(1024) k_scan_add(uint32_t* __restrict__ data, int n_static,
                                                         const uint32_t* __restrict__ count, const uint32_t* __restrict__ partial) {
    const int n = count ? (int)*count : n_static;
    const int base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    const uint32_t add = partial[blockIdx.x];
    for (int k = threadIdx.x; k < SCAN_TILE; k += 1024)
        if (base + k < n) data[base + k] += add;
}


// Stable scatter of one digit pass.  Sub-tiles of 256 keys are ranked in order: lanes with
// equal digits are found with 8 ballots, waves are combined through LDS.
static __global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                              int shift, const uint32_t* __restrict__ offs, int nblocks) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[RS_THREADS / WAVE][256];
    const int wid = threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    running[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
    const int base = blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; j++) {
        for (int w = 0; w < RS_THREADS / WAVE; w++) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int e = base + j * RS_THREADS + threadIdx.x;
        const bool valid = e < n;
        const uint32_t k = valid ? kin[e] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bits = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bits : ~bits;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && rank == 0) wcnt[wid][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        // exclusive prefix over waves for digit = threadIdx.x
        {
            uint32_t acc = running[threadIdx.x];
            for (int w = 0; w < RS_THREADS / WAVE; w++) { const uint32_t t = wcnt[w][threadIdx.x]; wcnt[w][threadIdx.x] = acc; acc += t; }
            running[threadIdx.x] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dst = wcnt[wid][d] + rank;
            kout[dst] = k;
            vout[dst] = vin[e];
        }
        __syncthreads();
    }
    (void)lane;
}

}  // namespace pe
