// device_common.hpp -- wave-level helpers and the ordered spike compaction
// shared by every update kernel (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edl {

constexpr int kWave = 64;
constexpr int kBlock = 256;               // threads per update workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

// ---------------------------------------------------------------------------
// Ordered compaction of spiking neuron indices, in two passes without atomics.
//
// Pass 1 (inside every update kernel): tile t = workgroup t covers neurons
// [t*T, (t+1)*T) with T = kBlock*ITEMS.  A wave ballot per item gives each
// spiking lane its rank, the wave counts are combined in LDS, and the tile
// writes its count and its spiking indices (ascending) into its slot of a
// staging array.
//
// Pass 2 (k_compact): each workgroup sums the counts of the tiles before its
// range, scans its own range and copies every tile's indices to the output,
// which is therefore ascending.  The last workgroup writes the total.
// ---------------------------------------------------------------------------
// Update index of a launch: an immediate value, or a device-memory base
// plus the launch's offset within a captured chunk (graph replays advance the
// base once per chunk).
struct StepRef {
    const int32_t *ptr;
    int32_t imm;
    __device__ __forceinline__ int32_t get() const { return ptr ? *ptr + imm : imm; }
};

struct TileOut {
    int32_t *tile_count;   // [n_tiles]
    int32_t *staging;      // [n_tiles * tile_size]
};

// Tiles are wave-major: item j of lane l in wave w of tile `tile` is neuron
// tile*kBlock*ITEMS + w*ITEMS*kWave + j*kWave + l, so ascending neuron order
// is wave, then item, then lane, and each wave covers one contiguous span of
// ITEMS*kWave neurons.
template <int ITEMS>
__device__ __forceinline__ int tile_neuron(int tile, int j) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    return tile * kBlock * ITEMS + wid * ITEMS * kWave + j * kWave + lane;
}

template <int ITEMS>
__device__ __forceinline__ void tile_spikes(const TileOut &o, int tile, const bool *spike,
                                            const int *idx) {
    __shared__ uint32_t s_wave[ITEMS * kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    uint64_t ballot[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        ballot[j] = __ballot(spike[j]);
        if (lane == 0) s_wave[wid * ITEMS + j] = (uint32_t)__popcll(ballot[j]);
    }
    __syncthreads();
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < ITEMS * kWavesPerBlock; q++) local += s_wave[q];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (spike[j]) {
            uint32_t base = 0;
            for (int q = 0; q < wid * ITEMS + j; q++) base += s_wave[q];
            const uint32_t rank = (uint32_t)__popcll(ballot[j] & ((1ull << lane) - 1ull));
            o.staging[(int64_t)tile * kBlock * ITEMS + base + rank] = idx[j];
        }
    }
    if (threadIdx.x == 0) o.tile_count[tile] = (int32_t)local;
}

// The same output for the item-major layout of the BDF kernels (item j of
// thread t of tile `tile` is neuron tile*kBlock*ITEMS + j*kBlock + t), whose
// ascending order is item, then wave, then lane: a population smaller than
// a workgroup gets one neuron per thread.
template <int ITEMS>
__device__ __forceinline__ void tile_spikes_item_major(const TileOut &o, int tile, const bool *spike,
                                                       const int *idx) {
    __shared__ uint32_t s_wave[ITEMS * kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    uint64_t ballot[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        ballot[j] = __ballot(spike[j]);
        if (lane == 0) s_wave[j * kWavesPerBlock + wid] = (uint32_t)__popcll(ballot[j]);
    }
    __syncthreads();
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < ITEMS * kWavesPerBlock; q++) local += s_wave[q];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (spike[j]) {
            uint32_t base = 0;
            for (int q = 0; q < j * kWavesPerBlock + wid; q++) base += s_wave[q];
            const uint32_t rank = (uint32_t)__popcll(ballot[j] & ((1ull << lane) - 1ull));
            o.staging[(int64_t)tile * kBlock * ITEMS + base + rank] = idx[j];
        }
    }
    if (threadIdx.x == 0) o.tile_count[tile] = (int32_t)local;
}

struct CompactArgs {
    const int32_t *tile_count;
    const int32_t *staging;
    int n_tiles;
    int tile_size;         // neurons per tile (staging stride)
    // output slot (step % ring) of a ring of `ring` lists of `slot_len` indices
    int32_t *out_idx;      // ascending indices
    int32_t *out_count;    // total
    int ring;
    int64_t slot_len;
    // optional spike log; the entry count is double-buffered by step parity:
    // read log_count[step & 1], write log_count[(step + 1) & 1].  With
    // log_idx null only the count is kept.
    int32_t *log_step;
    int32_t *log_idx;
    int64_t *log_count;
    int64_t log_cap;
    uint32_t *err;         // bit 1: log overflow
    StepRef step;
    // timed runs: the earliest workgroup start of this launch, per step in
    // [clk_first, clk_first + clk_count), on the device wall clock (null:
    // untimed).  It ends the previous update kernel's launch as a kernel
    // trace sees it (back-to-back dispatches).
    unsigned long long *clk_next;
    int32_t clk_first;
    int32_t clk_count;
};

__device__ __forceinline__ void clock_next_kernel(const CompactArgs &c) {
    if (c.clk_next && threadIdx.x == 0) {
        const int32_t s = c.step.get() - c.clk_first;
        if (s >= 0 && s < c.clk_count) atomicMin(&c.clk_next[s], (unsigned long long)wall_clock64());
    }
}

constexpr int kCompactTiles = kBlock;   // tiles per compaction workgroup

// Body of the compaction for workgroup `bid` of `nb` (also run by the fused
// compaction + input-count kernel).
__device__ __forceinline__ void compact_body(const CompactArgs &a, int bid, int nb) {
    const int32_t step = a.step.get();
    const int slot = step % a.ring;
    int32_t *const out_idx = a.out_idx + (int64_t)slot * a.slot_len;
    __shared__ int64_t s_red[kWavesPerBlock];
    __shared__ int32_t s_scan[kBlock];
    __shared__ int64_t s_base;
    const int t0 = bid * kCompactTiles;
    // 1. spikes in the tiles before this range
    // tile_count is zero-padded to a multiple of 8*kBlock entries, and t0 is
    // a multiple of kBlock: the eight loads of a round are unconditional.
    int64_t before = 0;
    for (int t = threadIdx.x; t < t0; t += 8 * kBlock) {
        int32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a.tile_count[t + u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (t + u * kBlock < t0) before += v[u];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) before += __shfl_xor(before, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = before;
    // 2. inclusive scan of this range's tile counts
    const int my = t0 + (int)threadIdx.x;
    s_scan[threadIdx.x] = my < a.n_tiles ? a.tile_count[my] : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t b = 0;
        for (int q = 0; q < kWavesPerBlock; q++) b += s_red[q];
        s_base = b;
    }
    for (int off = 1; off < kBlock; off <<= 1) {
        const int32_t v = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    const int64_t base = s_base;
    const int64_t log_base = a.log_count ? a.log_count[step & 1] : 0;
    // 3. copy: one thread per tile (most tiles hold no spike)
    {
        const int q = (int)threadIdx.x;
        const int t = t0 + q;
        const int32_t lo = q > 0 ? s_scan[q - 1] : 0;
        const int32_t c = t < a.n_tiles ? s_scan[q] - lo : 0;
        const int64_t start = base + lo;
        for (int j = 0; j < c; j++) {
            const int32_t id = a.staging[(int64_t)t * a.tile_size + j];
            out_idx[start + j] = id;
            if (a.log_idx) {
                const int64_t lp = log_base + start + j;
                if (lp < a.log_cap) {
                    a.log_idx[lp] = id;
                    a.log_step[lp] = step;
                }
            }
        }
    }
    if (bid == nb - 1 && threadIdx.x == 0) {
        const int64_t total = base + s_scan[kBlock - 1];
        a.out_count[slot] = (int32_t)total;
        if (a.log_count) {
            const int64_t nl = log_base + total;
            if (a.log_idx && nl > a.log_cap) atomicOr(a.err, 2u);
            a.log_count[(step + 1) & 1] = nl;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_compact(CompactArgs a) {
    clock_next_kernel(a);
    compact_body(a, blockIdx.x, gridDim.x);
}

// Advances a device step base (last node of a captured chunk).
__global__ void k_advance_step(int32_t *step, int32_t by) {
    if (threadIdx.x == 0) *step = *step + by;
}

// Counter-based hash of (seed, step, neuron) for the external drive.
// mix32 is a 32-bit multiply-xorshift finaliser; the (seed, step) part is the
// same for every neuron of a launch and is computed once.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__host__ __device__ __forceinline__ uint32_t drive_step_key(uint64_t seed, int64_t step) {
    return mix32((uint32_t)seed ^ mix32((uint32_t)(seed >> 32) ^ mix32((uint32_t)step)));
}
__host__ __device__ __forceinline__ uint32_t drive_hash_keyed(uint32_t step_key, uint32_t neuron) {
    return mix32(neuron * 0x9E3779B9u + step_key);
}
__host__ __device__ __forceinline__ uint32_t drive_hash(uint64_t seed, int64_t step,
                                                        uint32_t neuron) {
    return drive_hash_keyed(drive_step_key(seed, step), neuron);
}

}  // namespace edl
