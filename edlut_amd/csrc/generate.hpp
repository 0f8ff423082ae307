// generate.hpp -- synthetic recurrent connectivity built on the device.
//
// Fixed in-degree random wiring inside each group of a population (one group
// for a plain population): target t draws `in_degree` sources uniformly (with
// replacement) from the other neurons of its own group and a delay uniformly
// from [1, max_delay], both from a
// counter-based hash of (group seed, local target index, draw index), so a
// group's wiring does not depend on where the group sits in the index space.
// The result is the out-CSR the delivery kernels read: rows (source, delay),
// targets ascending inside a row (Network::FindOutConnections order,
// Network.cpp:45-123), synapse type by source (the first
// round(exc_fraction * n) sources of a group are excitatory), weights
// lognormal(mu, sigma).
//
//   k_gen_edges<false>  per-row synapse counts (atomic increments)
//   k_scan_*            exclusive prefix sum of the counts -> row_ptr
//   k_gen_edges<true>   every synapse drops (target, draw) into its row
//   k_gen_rows          each row sorts its entries and writes target|type, weight
#pragma once
#include "kernels.hpp"

namespace edl {

struct GenArgs {
    const int32_t *g_start;        // per group: first neuron
    const int32_t *g_size;         // per group: real neurons
    const int32_t *g_exc;          // per group: excitatory sources
    const uint32_t *chunk_group;   // group of each kGroupAlign chunk (null: one group)
    int64_t n;                     // neurons in the index space
    int in_degree;                 // <= 256
    int max_delay;
    float w_mu, w_sigma;
    uint64_t seed;                 // group g uses seed + g
    uint32_t *row_cnt;             // [n*max_delay] counts, then fill cursors
    int64_t *row_ptr;              // [n*max_delay + 1]
    uint64_t *key;                 // [synapses] target << 8 | draw
    uint32_t *tgt_type;
    float *w;
};

__device__ __forceinline__ uint32_t gen_hash(uint64_t seed, uint32_t local, uint32_t draw,
                                             uint32_t salt) {
    const uint32_t a = mix32((uint32_t)(seed >> 32) ^ mix32(local * 0x85ebca6bu + salt));
    return mix32(mix32((uint32_t)seed ^ a) + draw * 0x9E3779B9u);
}

// uniform integer in [0, m) from a 32-bit hash (multiply-high)
__device__ __forceinline__ uint32_t gen_below(uint32_t h, uint32_t m) {
    return (uint32_t)(((uint64_t)h * (uint64_t)m) >> 32);
}

__device__ __forceinline__ bool gen_target(const GenArgs &a, int64_t t, int &g, int32_t &start,
                                           int32_t &size) {
    g = a.chunk_group ? (int)a.chunk_group[t / kGroupAlign] : 0;
    start = a.g_start[g];
    size = a.g_size[g];
    return t - start < size;
}

template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_gen_edges(GenArgs a) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n) return;
    int g;
    int32_t start, size;
    if (!gen_target(a, t, g, start, size)) return;
    const uint64_t seed = a.seed + (uint64_t)g;
    const uint32_t local = (uint32_t)(t - start);
    for (int k = 0; k < a.in_degree; k++) {
        // uniform over the group without the target itself
        uint32_t sl = gen_below(gen_hash(seed, local, k, 1u), (uint32_t)size - 1u);
        sl += sl >= local ? 1u : 0u;
        const uint32_t src = (uint32_t)start + sl;
        const uint32_t d = 1u + gen_below(gen_hash(seed, local, k, 2u), (uint32_t)a.max_delay);
        const int64_t row = (int64_t)src * a.max_delay + (d - 1);
        const uint32_t c = atomicAdd(&a.row_cnt[row], 1u);
        if (FILL) a.key[a.row_ptr[row] + c] = ((uint64_t)t << 8) | (uint64_t)k;
    }
}

__global__ __launch_bounds__(kBlock) void k_gen_rows(GenArgs a) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= a.n * a.max_delay) return;
    const int64_t b = a.row_ptr[r], e = a.row_ptr[r + 1];
    // insertion sort: rows hold in_degree/max_delay entries on average
    for (int64_t j = b + 1; j < e; j++) {
        const uint64_t v = a.key[j];
        int64_t i = j - 1;
        while (i >= b && a.key[i] > v) {
            a.key[i + 1] = a.key[i];
            --i;
        }
        a.key[i + 1] = v;
    }
    const int64_t src = r / a.max_delay;
    for (int64_t q = b; q < e; q++) {
        const uint64_t v = a.key[q];
        const int64_t t = (int64_t)(v >> 8);
        const uint32_t k = (uint32_t)(v & 255u);
        int g;
        int32_t start, size;
        gen_target(a, t, g, start, size);
        const uint64_t seed = a.seed + (uint64_t)g;
        const uint32_t local = (uint32_t)(t - start);
        const uint32_t type = (src - start) < (int64_t)a.g_exc[g] ? 0u : 1u;
        // Box-Muller from two hashed uniforms in (0, 1)
        const float u1 = ((float)(gen_hash(seed, local, k, 3u) >> 8) + 0.5f) * (1.0f / 16777216.0f);
        const float u2 = (float)(gen_hash(seed, local, k, 4u) >> 8) * (1.0f / 16777216.0f);
        const float z = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307f * u2);
        a.tgt_type[q] = (uint32_t)t | (type << 30);
        a.w[q] = expf(a.w_mu + a.w_sigma * z);
    }
}

// ---------------------------------------------------------------------------
// Exclusive prefix sum of uint32 counts into int64 offsets (out[n] = total),
// in three passes: per-chunk sums, a one-workgroup scan of the sums, and a
// per-chunk scan plus offset.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanChunk = kBlock * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *total) {
    __shared__ int64_t s[kBlock];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int64_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    const int64_t incl = s[threadIdx.x];
    *total = s[kBlock - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(const uint32_t *in, int64_t n, int64_t *sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int64_t v = 0;
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + (int64_t)j * kBlock + threadIdx.x;
        if (i < n) v += in[i];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    __shared__ int64_t s_w[kWavesPerBlock];
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int q = 0; q < kWavesPerBlock; q++) t += s_w[q];
        sums[blockIdx.x] = t;
    }
}

// one workgroup: in-place exclusive scan of the chunk sums, total at sums[m]
__global__ __launch_bounds__(kBlock) void k_scan_top(int64_t *sums, int64_t m) {
    int64_t carry = 0;
    for (int64_t b = 0; b < m; b += kBlock) {
        const int64_t i = b + threadIdx.x;
        const int64_t v = i < m ? sums[i] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(v, &tot);
        if (i < m) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[m] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t *in, int64_t n,
                                                       const int64_t *sums, int64_t *out) {
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    int64_t mine = 0;
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + j;
        v[j] = i < n ? in[i] : 0u;
        mine += v[j];
    }
    int64_t tot;
    int64_t run = sums[blockIdx.x] + block_exclusive_scan(mine, &tot);
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

}  // namespace edl
