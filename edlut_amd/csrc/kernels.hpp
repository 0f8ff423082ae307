// kernels.hpp -- the per-step kernels of the time-driven neuron update path.
//
//   k_update     one step of every neuron: optional external drive, ODE step,
//                conductance decay, threshold/reset, NaN reset and ordered
//                spike compaction (one fused pass over the SoA state).
//   k_poisson    one step of a Poisson generator device (per-source LCG).
//   k_fold_*     ordered conductance accumulation of a step's input events.
#pragma once
#include <type_traits>
#include "device_common.hpp"
#include "device_models.hpp"
#include "eventq.hpp"

namespace edl {

// kBif*: the GPU Bifixed methods -- UpdateArgs::n_sub steps of the base
// method (INTEG - kBifixed) per update, at the sub-step's dt.
enum Integ { kEuler = 0, kRK2 = 1, kRK4 = 2, kBifEuler = 3, kBifRK2 = 4, kBifRK4 = 5 };
constexpr int kBifixed = 3;

// ---------------------------------------------------------------------------
// Ordered conductance accumulation.
//
// Every input event of a step is (key = type*stride + target, order, w); the
// CPU reference adds them one by one into the state, in `order`.  The fold
// reproduces that sequence exactly, for any event density, in two passes
// plus a tail in the update kernel:
//
//   count    per-key event counts (atomic increments)
//   scatter  keys with one event add it directly -- no ordering question;
//            for the others the first arrival reserves a contiguous segment
//            of the event buffer, every event drops (order, w) into it, and
//            the target's tile is flagged
//   update   in flagged tiles each neuron sorts its keys' segments by order
//            and folds them sequentially before integrating, then clears the
//            key's counters
//
// The event buffer can hold every event of a step (one per synapse) in the
// network mode, so it cannot overflow there.
// ---------------------------------------------------------------------------
constexpr int kItems = 4;                    // neurons per thread in the update kernels
constexpr int kUpdTile = kBlock * kItems;    // neurons per update tile (kItems per thread)
constexpr int kFlagTile = 512;               // neurons per fold flag (smallest update tile)
constexpr int kGroupAlign = 256;             // grouped populations: groups start on this
constexpr uint64_t kSegEmpty = ~0ull;
constexpr uint64_t kSegBusy = ~0ull - 1ull;

// One queued event of the fold: its order key and weight in one 16-byte
// entry, so that scattering an event writes one line.
struct FoldEntry {
    uint64_t ord;
    float w;
    uint32_t pad;
};

struct FoldTarget {
    float *state;
    int64_t stride;
    int ndiff;
    uint32_t *cnt;       // per key: events this step
    uint32_t *fill;      // per key: events scattered so far
    uint64_t *seg;       // per key: segment start (kSegEmpty when none)
    uint32_t *tile_flag; // per kFlagTile neurons: some key there has >= 2 events
    uint64_t *bump;      // [2] event-buffer allocator, by step parity
    int par;             // parity of this step, unless par_by_step
    int par_by_step;     // parity = update index & 1 (network mode)
    FoldEntry *b_ent;    // event buffer (order key, weight)
    uint64_t b_cap;
    uint32_t *err;       // bit 2: event buffer overflow, bit 3: spin limit
    // network mode: fixed segment of each key (prefix of the keys' synapse
    // counts; a synapse carries at most one event per step), null otherwise
    const uint64_t *seg_off;
    // network mode with the event queue's order: the rank tables (null:
    // canonical order)
    const struct RankTab *rtab;
};

// Rank of each group of the window that update k consumes, per source,
// [2][delays][cmax] by the parity of k (device memory, one per network).
struct RankTab {
    const uint32_t *rank[4];
    int32_t cmax[4];
    int32_t delays;
};

// Network-mode order key of an event: delay, source, chunk (the group of 1,024
// spiking sources it belongs to, or the input spike's position in its update),
// synapse position.  Sorting by it gives the canonical order (delay, source,
// spiking index, synapse); the event queue's order replaces (delay, source,
// chunk) by the group's rank in its window.
constexpr int kOrdSynBits = 36;
constexpr uint64_t kOrdSynMask = (1ull << kOrdSynBits) - 1ull;
constexpr int kOrdChunkBits = 18;
__host__ __device__ __forceinline__ uint64_t ord_key(int d, int p, uint32_t c, uint64_t q) {
    return ((uint64_t)(d - 1) << 56) | ((uint64_t)p << 54) | ((uint64_t)c << kOrdSynBits) | q;
}

__device__ __forceinline__ float *g_ptr(const FoldTarget &t, uint64_t key) {
    const uint64_t type = key / (uint64_t)t.stride;
    const uint64_t tgt = key - type * (uint64_t)t.stride;
    return t.state + (int64_t)(t.ndiff + (int)type) * t.stride + (int64_t)tgt;
}

enum FoldPass { kCount = 0, kScatter = 1, kStatic = 2 };

template <int PASS>
__device__ __forceinline__ void fold_event(const FoldTarget &t, int par, uint64_t key, uint64_t ord,
                                           float w) {
    if (PASS == kStatic) {
        // network mode, one pass: the key's fixed segment, next free entry
        const uint32_t pos = atomicAdd(&t.fill[key], 1u);
        const uint64_t at = t.seg_off[key] + pos;
        t.b_ent[at] = FoldEntry{ord, w, 0u};
        if (pos == 0u) t.tile_flag[(key % (uint64_t)t.stride) / kFlagTile] = 1u;
        return;
    }
    if (PASS == kCount) {
        atomicAdd(&t.cnt[key], 1u);
        return;
    }
    const uint32_t c = t.cnt[key];
    if (c == 1u) {
        float *g = g_ptr(t, key);
        *g = *g + w;
        t.cnt[key] = 0u;
        return;
    }
    // Reserve the key's segment (first arrival) -- every lane does this step
    // before any lane waits, so a wave never waits on one of its own lanes.
    const uint64_t prev = atomicCAS((unsigned long long *)&t.seg[key], kSegEmpty, kSegBusy);
    if (prev == kSegEmpty) {
        const uint64_t base = atomicAdd((unsigned long long *)&t.bump[par], (unsigned long long)c);
        __hip_atomic_store(&t.seg[key], base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t tgt = key % (uint64_t)t.stride;
        t.tile_flag[tgt / kFlagTile] = 1u;
    }
    uint64_t base = __hip_atomic_load(&t.seg[key], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (base == kSegBusy) {
        if (++spins > (1u << 22)) {
            atomicOr(t.err, 8u);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
        base = __hip_atomic_load(&t.seg[key], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t pos = base + atomicAdd(&t.fill[key], 1u);
    if (pos < t.b_cap) {
        t.b_ent[pos] = FoldEntry{ord, w, 0u};
    } else {
        atomicOr(t.err, 4u);
    }
}

// Network mode: folds the m >= 1 events of `key` (its fixed segment) into g
// in order and clears the key's fill count.
// Rank of the group an order key belongs to, in the window of the update of
// parity par.
__device__ __forceinline__ uint32_t key_rank(const RankTab *rt, int par, uint64_t o) {
    const int d = (int)(o >> 56) + 1;
    const int p = (int)((o >> 54) & 3u);
    const uint32_t c = (uint32_t)((o >> kOrdSynBits) & ((1u << kOrdChunkBits) - 1u));
    return rt->rank[p][((int64_t)par * rt->delays + (d - 1)) * rt->cmax[p] + c];
}

__device__ __attribute__((noinline)) float fold_static(uint32_t *fill, const uint64_t *seg_off, FoldEntry *e,
                                                       const RankTab *rt, int par, uint64_t key, uint32_t m,
                                                       float g) {
    const uint64_t b = seg_off[key];
    if (rt && m >= 2u) {
        // (delay, source, chunk) -> the group's rank in the event queue's window
        for (uint32_t j = 0; j < m; j++) {
            const uint64_t o = e[b + j].ord;
            e[b + j].ord = ((uint64_t)key_rank(rt, par, o) << kOrdSynBits) | (o & kOrdSynMask);
        }
    }
    for (uint32_t j = 1; j < m; j++) {
        const FoldEntry k = e[b + j];
        int64_t i = (int64_t)j - 1;
        while (i >= 0 && e[b + i].ord > k.ord) {
            e[b + i + 1] = e[b + i];
            --i;
        }
        e[b + i + 1] = k;
    }
    for (uint32_t j = 0; j < m; j++) g = g + e[b + j].w;
    fill[key] = 0u;
    return g;
}

// Folds the segment of `key` (m >= 2 events) into g in order and clears the
// key's counters.  Called by the update kernel for its own neuron, out of
// line so the common path keeps its small register budget; the segment is
// sorted in place in the event buffer (insertion sort: segments are short).
__device__ __attribute__((noinline)) float fold_segment(uint32_t *cnt, uint32_t *fill,
                                                        uint64_t *seg, FoldEntry *e, uint64_t b_cap,
                                                        uint32_t *err, uint64_t key,
                                                        uint32_t m, float g) {
    const uint64_t b = seg[key];
    if (b == kSegEmpty || b == kSegBusy || b + m > b_cap) {
        atomicOr(err, 4u);
    } else {
        for (uint32_t j = 1; j < m; j++) {
            const FoldEntry k = e[b + j];
            int64_t i = (int64_t)j - 1;
            while (i >= 0 && e[b + i].ord > k.ord) {
                e[b + i + 1] = e[b + i];
                --i;
            }
            e[b + i + 1] = k;
        }
        for (uint32_t j = 0; j < m; j++) g = g + e[b + j].w;
    }
    cnt[key] = 0u;
    fill[key] = 0u;
    seg[key] = kSegEmpty;
    return g;
}

struct ExtDrive {
    uint64_t seed;
    uint32_t thr[16];  // cumulative thresholds, decreasing
    int n_thr;
    float w;
};

// Grouped populations (many independent networks in one index space): group
// g owns neurons [start, start + n) and its own model parameters.  Groups
// start on multiples of kGroupAlign, so every wave's span of a tile lies in
// one group and the group's parameters are wave-uniform (scalar loads).
struct GroupInfo {
    ModelParams mp;
    int32_t start;
    int32_t n;
};
struct GroupTable {
    const GroupInfo *tab;
    const uint32_t *chunk_group;   // group of each kGroupAlign-neuron chunk
};

struct UpdateArgs {
    float *state;          // SoA, row stride `stride`
    int64_t stride;
    int n;                 // neurons
    int n_tiles;
    StepRef step;          // update index k (>= 1)
    ModelParams mp;
    StepConsts k;
    const float *chan;     // channel-rate table (HH, granule, Purkinje)
    const float *exp_lut;
    const float *nmda_lut;
    const float *refr;
    int refr_len;
    int32_t *last_spike;   // LIF, Purkinje: update (Bifixed: sub-step) index of the last spike
    int n_sub;             // Bifixed kernels: sub-steps per update
    BdfState bdf;          // BDF integrator history
    ExtDrive ext;
    TileOut out;
    FoldTarget fold;       // ordered input fold of this step (multi-event keys)
    GroupTable grp;        // grouped populations only
    // timed runs: every workgroup stores its start and end on the device
    // wall clock at clk[((step - clk_first) * gridDim.x + block) * 2 + {0,1}]
    // for steps in [clk_first, clk_first + clk_count) (clk null: untimed)
    unsigned long long *clk;
    int32_t clk_first;
    int32_t clk_count;
};

__device__ __forceinline__ unsigned long long *clock_slot(const UpdateArgs &a, int32_t step) {
    const int32_t k = step - a.clk_first;
    if (!a.clk || k < 0 || k >= a.clk_count) return nullptr;
    return a.clk + ((size_t)k * gridDim.x + blockIdx.x) * 2;
}
__device__ __forceinline__ void clock_begin(const UpdateArgs &a, int32_t step) {
    unsigned long long *c = clock_slot(a, step);
    if (c && threadIdx.x == 0) c[0] = (unsigned long long)wall_clock64();
}
__device__ __forceinline__ void clock_finish(const UpdateArgs &a, int32_t step) {
    unsigned long long *c = clock_slot(a, step);
    if (c) {
        __syncthreads();
        if (threadIdx.x == 0) c[1] = (unsigned long long)wall_clock64();
    }
}

// Reads (and clears) the fold flags covering a tile of T neurons.
template <int T>
__device__ __forceinline__ uint32_t take_tile_flag(uint32_t *flags, int tile) {
    constexpr int F = T / kFlagTile;
    uint32_t f = 0;
#pragma unroll
    for (int q = 0; q < F; q++) f |= flags[tile * F + q];
    return f;
}
template <int T>
__device__ __forceinline__ void clear_tile_flag(uint32_t *flags, int tile) {
    constexpr int F = T / kFlagTile;
#pragma unroll
    for (int q = 0; q < F; q++) flags[tile * F + q] = 0u;
}

// One tile = kBlock*ITEMS neurons in the wave-major layout of tile_neuron:
// each lane handles ITEMS neurons kWave apart (each item is one coalesced row
// access per wave).  All loads are issued first, the ITEMS neurons are
// integrated stage-major, and the stores come last, so every thread keeps
// ITEMS independent dependency chains in flight.  GROUPED takes the model
// parameters, the live range and the drive seed from the wave's group.
template <class M, int INTEG, unsigned MASK, bool EXT, int ITEMS, bool GROUPED>
__global__ __launch_bounds__(kBlock) void k_update(UpdateArgs a) {
    constexpr int T = kBlock * ITEMS;
    static_assert(T % kFlagTile == 0, "update tiles cover whole fold flags");
    static_assert(kGroupAlign % (ITEMS * kWave) == 0, "a wave's span lies in one group");
    const int tile = blockIdx.x;
    const int32_t step = a.step.get();
    clock_begin(a, step);
    const int par = a.fold.par_by_step ? (step & 1) : a.fold.par;
    __shared__ uint32_t s_flag;
    // the tile's fold flag is read alongside the state loads (the barrier
    // that publishes it comes after them)
    uint32_t flag = 0;
    if (threadIdx.x == 0) flag = take_tile_flag<T>(a.fold.tile_flag, tile);
    ModelParams gmp;
    int32_t gstart = 0, gn = a.n;
    uint64_t dseed = a.ext.seed;
    if (GROUPED) {
        const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
        const int first = tile * T + wid * ITEMS * kWave;
        const uint32_t g = a.grp.chunk_group[first / kGroupAlign];
        gmp = a.grp.tab[g].mp;
        gstart = a.grp.tab[g].start;
        gn = a.grp.tab[g].n;
        dseed = a.ext.seed + g;
    }
    const ModelParams &mp = GROUPED ? gmp : a.mp;
    // one row pointer per state variable (uniform), 32-bit neuron offsets
    float *__restrict__ row[M::NV];
#pragma unroll
    for (int j = 0; j < M::NV; j++) row[j] = a.state + (int64_t)j * a.stride;
    const int64_t ld = a.stride;
    const uint32_t step_key = EXT ? drive_step_key(dseed, step) : 0u;
    int idx[ITEMS];
    bool live[ITEMS];
    float s[ITEMS][M::NV];
    RhsCtx c[ITEMS];
#pragma unroll
    for (int n = 0; n < ITEMS; n++) {
        idx[n] = tile_neuron<ITEMS>(tile, n);
        live[n] = idx[n] - gstart < gn;
        // rows are padded to a multiple of kUpdTile, so the loads stay in bounds
        const uint32_t i = (uint32_t)idx[n];
#pragma unroll
        for (int j = 0; j < M::ND; j++) s[n][j] = row[j][i];
        s[n][M::EXC] = (MASK & 1u) ? row[M::EXC][i] : 0.0f;
        s[n][M::INH] = (MASK & 2u) ? row[M::INH][i] : 0.0f;
        s[n][M::NMDA] = (MASK & 4u) ? row[M::NMDA][i] : 0.0f;
        s[n][M::EXT] = (MASK & 8u) ? row[M::EXT][i] : 0.0f;
        c[n].chan = a.chan;
        c[n].exp_lut = a.exp_lut;
        c[n].nmda_lut = a.nmda_lut;
        c[n].refr = a.refr;
        c[n].refr_len = a.refr_len;
        c[n].since_spike = a.last_spike ? (step - 1 - a.last_spike[i]) : (1 << 30);
    }
    if (threadIdx.x == 0) {
        s_flag = flag;
        if (flag) clear_tile_flag<T>(a.fold.tile_flag, tile);
        // the next step's fold allocates from the other parity: clear it
        if (tile == 0) a.fold.bump[par ^ 1] = 0ull;
    }
    __syncthreads();
    if (s_flag) {
        // events of keys with several inputs this step, in delivery order
#pragma unroll
        for (int n = 0; n < ITEMS; n++) {
            if (!live[n]) continue;
#pragma unroll
            for (int ty = 0; ty < 3; ty++) {
                if (!(MASK & (1u << ty))) continue;
                const uint64_t key = (uint64_t)ty * (uint64_t)ld + (uint64_t)idx[n];
                if (a.fold.seg_off) {
                    const uint32_t m = a.fold.fill[key];
                    if (m >= 1u)
                        s[n][M::EXC + ty] = fold_static(a.fold.fill, a.fold.seg_off, a.fold.b_ent,
                                                        a.fold.rtab, step & 1, key, m, s[n][M::EXC + ty]);
                    continue;
                }
                const uint32_t m = a.fold.cnt[key];
                if (m >= 2u)
                    s[n][M::EXC + ty] = fold_segment(a.fold.cnt, a.fold.fill, a.fold.seg, a.fold.b_ent,
                                                     a.fold.b_cap, a.fold.err, key, m,
                                                     s[n][M::EXC + ty]);
            }
        }
    }
    if (EXT) {
        // External drive events of this interval, after the recurrent ones;
        // the hash takes the neuron's index within its group.  The event
        // count is #{j : u < thr[j]} (thresholds decrease); the loop walks
        // the thresholds in wave-uniform order and stops once no lane hits.
        uint32_t u[ITEMS];
        int cnt[ITEMS];
#pragma unroll
        for (int n = 0; n < ITEMS; n++) {
            u[n] = drive_hash_keyed(step_key, (uint32_t)(idx[n] - gstart));
            cnt[n] = 0;
        }
        for (int j = 0; j < a.ext.n_thr; j++) {
            const uint32_t t = a.ext.thr[j];
            bool hit = false;
#pragma unroll
            for (int n = 0; n < ITEMS; n++) {
                const bool h = u[n] < t;
                cnt[n] += h ? 1 : 0;
                hit = hit || h;
            }
            if (!__any(hit)) break;
        }
#pragma unroll
        for (int n = 0; n < ITEMS; n++)
            for (int q = 0; q < cnt[n]; q++) s[n][M::EXC] += a.ext.w;
    }
    // Fixed methods take one step.  The Bifixed methods take n_sub steps of
    // the base method, each followed by the spike condition, and report a
    // neuron once per update (Bifixed_RK4_GPU2.cuh:100-150); their spike clock
    // counts sub-steps, numbered from 1 over the run (update k covers
    // (k-1)*n_sub+1 .. k*n_sub), and the clock tables are built for the
    // sub-step.
    constexpr bool SUB = INTEG >= kBifixed;
    constexpr int BASE = SUB ? INTEG - kBifixed : INTEG;
    const int n_sub = SUB ? a.n_sub : 1;
    const int32_t g0 = SUB ? (step - 1) * n_sub : step - 1;
    bool spk[ITEMS];
    int32_t last[ITEMS];
#pragma unroll
    for (int n = 0; n < ITEMS; n++) {
        spk[n] = false;
        last[n] = (SUB && a.last_spike) ? a.last_spike[(uint32_t)idx[n]] : 0;
    }
    for (int it = 0; it < n_sub; it++) {
        const int32_t g = g0 + it + 1;
        float prev_v[ITEMS];
#pragma unroll
        for (int n = 0; n < ITEMS; n++) {
            if (SUB && a.last_spike) c[n].since_spike = g - 1 - last[n];
            prev_v[n] = s[n][0];
        }
        if (BASE == kEuler) integrate_euler<M, MASK, ITEMS>(mp, c, a.k, s);
        else if (BASE == kRK2) integrate_rk2<M, MASK, ITEMS>(mp, c, a.k, s);
        else integrate_rk4<M, MASK, ITEMS>(mp, c, a.k, s);
#pragma unroll
        for (int n = 0; n < ITEMS; n++) {
            const bool f = live[n] && M::spike(mp, prev_v[n], s[n]);
            if constexpr (M::kClockOverride)
                if (live[n]) M::clock_override(c[n], f, s[n]);
            if (f) {
                spk[n] = true;
                last[n] = g;
            }
        }
    }

#pragma unroll
    for (int n = 0; n < ITEMS; n++) {
        const uint32_t i = (uint32_t)idx[n];
        if (!live[n]) continue;
        if (spk[n] && a.last_spike) a.last_spike[i] = last[n];
        if (s[n][0] != s[n][0]) {
            // TimeDrivenNeuronModel.cpp:75-88: NaN -> whole state to initial.
#pragma unroll
            for (int j = 0; j < M::NV; j++) row[j][i] = M::init(mp, j);
        } else {
            // non-temporal stores: the next reader is the next step's
            // update, so the rows need not stay in L2 (and the end-of-kernel
            // write-back has less to do): 13.7 us per 1M-neuron AdEx launch
            // against 14.0 us with ordinary stores
#pragma unroll
            for (int j = 0; j < M::ND; j++) __builtin_nontemporal_store(s[n][j], &row[j][i]);
            if (MASK & 1u) __builtin_nontemporal_store(s[n][M::EXC], &row[M::EXC][i]);
            if (MASK & 2u) __builtin_nontemporal_store(s[n][M::INH], &row[M::INH][i]);
            if (MASK & 4u) __builtin_nontemporal_store(s[n][M::NMDA], &row[M::NMDA][i]);
        }
    }
    tile_spikes<ITEMS>(a.out, tile, spk, idx);
    clock_finish(a, step);
}


// Loads the state of neuron i at the start of a BDF step into sf: the row
// values, the ordered fold of its multi-event keys when its tile's fold flag
// is set, and the external drive.
template <class M, unsigned MASK, bool EXT>
__device__ __forceinline__ void bdf_load(const UpdateArgs &a, float *__restrict__ const *row, int32_t step,
                                         uint32_t step_key, uint32_t flag, uint32_t i, float *sf) {
    const int64_t ld = a.stride;
#pragma unroll
    for (int j = 0; j < M::ND; j++) sf[j] = row[j][i];
    sf[M::EXC] = (MASK & 1u) ? row[M::EXC][i] : 0.0f;
    sf[M::INH] = (MASK & 2u) ? row[M::INH][i] : 0.0f;
    sf[M::NMDA] = (MASK & 4u) ? row[M::NMDA][i] : 0.0f;
    sf[M::EXT] = (MASK & 8u) ? row[M::EXT][i] : 0.0f;
    if (flag) {
#pragma unroll
        for (int ty = 0; ty < 3; ty++) {
            if (!(MASK & (1u << ty))) continue;
            const uint64_t key = (uint64_t)ty * (uint64_t)ld + (uint64_t)i;
            if (a.fold.seg_off) {
                const uint32_t m = a.fold.fill[key];
                if (m >= 1u)
                    sf[M::EXC + ty] = fold_static(a.fold.fill, a.fold.seg_off, a.fold.b_ent,
                                                  a.fold.rtab, step & 1, key, m, sf[M::EXC + ty]);
                continue;
            }
            const uint32_t m = a.fold.cnt[key];
            if (m >= 2u)
                sf[M::EXC + ty] = fold_segment(a.fold.cnt, a.fold.fill, a.fold.seg, a.fold.b_ent,
                                               a.fold.b_cap, a.fold.err, key, m,
                                               sf[M::EXC + ty]);
        }
    }
    if (EXT) {
        const uint32_t u = drive_hash_keyed(step_key, i);
        int cnt = 0;
        while (cnt < a.ext.n_thr && u < a.ext.thr[cnt]) cnt++;
        for (int q = 0; q < cnt; q++) sf[M::EXC] += a.ext.w;
    }
}

// Stores neuron i's state after the step (TimeDrivenNeuronModel.cpp:75-88:
// a NaN membrane potential puts the whole state back to its initial values).
template <class M, unsigned MASK>
__device__ __forceinline__ void bdf_store(const UpdateArgs &a, float *__restrict__ const *row, uint32_t i,
                                          const float *out) {
    if (out[0] != out[0]) {
#pragma unroll
        for (int j = 0; j < M::NV; j++) row[j][i] = M::init(a.mp, j);
    } else {
#pragma unroll
        for (int j = 0; j < M::ND; j++) row[j][i] = out[j];
        if (MASK & 1u) row[M::EXC][i] = out[M::EXC];
        if (MASK & 2u) row[M::INH][i] = out[M::INH];
        if (MASK & 4u) row[M::NMDA][i] = out[M::NMDA];
    }
}

// BDF update: one neuron at a time per thread (data-dependent Newton
// iterations), ITEMS neurons per thread in the item-major layout of
// tile_spikes_item_major (a small population takes one neuron per thread).
// T = double runs the whole BDF step in double (FP64 Newton mode); the state
// is stored back in float.
template <class M, unsigned MASK, bool EXT, typename T, int ITEMS = kItems>
__global__ __launch_bounds__(kBlock) void k_update_bdf(UpdateArgs a) {
    constexpr int TILE = kBlock * ITEMS;
    static_assert(TILE % kFlagTile == 0, "update tiles cover whole fold flags");
    const int tile = blockIdx.x;
    const int32_t step = a.step.get();
    clock_begin(a, step);
    const int par = a.fold.par_by_step ? (step & 1) : a.fold.par;
    __shared__ uint32_t s_flag;
    if (threadIdx.x == 0) {
        s_flag = take_tile_flag<TILE>(a.fold.tile_flag, tile);
        if (s_flag) clear_tile_flag<TILE>(a.fold.tile_flag, tile);
        if (tile == 0) a.fold.bump[par ^ 1] = 0ull;
    }
    __syncthreads();
    float *__restrict__ row[M::NV];
#pragma unroll
    for (int j = 0; j < M::NV; j++) row[j] = a.state + (int64_t)j * a.stride;
    const int64_t ld = a.stride;
    const uint32_t step_key = EXT ? drive_step_key(a.ext.seed, step) : 0u;
    bool spk[ITEMS];
    int idx[ITEMS];
    for (int n = 0; n < ITEMS; n++) {
        idx[n] = tile * TILE + n * kBlock + (int)threadIdx.x;
        spk[n] = false;
        if (idx[n] >= a.n) continue;
        const uint32_t i = (uint32_t)idx[n];
        float sf[M::NV];
        bdf_load<M, MASK, EXT>(a, row, step, step_key, s_flag, i, sf);
        T S[M::NV];
#pragma unroll
        for (int j = 0; j < M::NV; j++) S[j] = sf[j];
        const T prev = S[0];
        RhsCtx c;
        c.chan = a.chan;
        c.exp_lut = a.exp_lut;
        c.nmda_lut = a.nmda_lut;
        c.refr = a.refr;
        c.refr_len = a.refr_len;
        c.since_spike = a.last_spike ? (step - 1 - a.last_spike[i]) : (1 << 30);
        int st = a.bdf.st[i];
        integrate_bdf<M, MASK, T>(a.mp, c, (T)a.k.dt, a.bdf, ld, i, st, S);
        const bool fired = M::spike(a.mp, prev, S);
        if constexpr (M::kClockOverride) M::clock_override(c, fired, S);
        if (fired && M::kResetsIntegrator) st = 0;
        a.bdf.st[i] = st;
        spk[n] = fired;
        if (fired && a.last_spike) a.last_spike[i] = step;
        float out[M::NV];
#pragma unroll
        for (int j = 0; j < M::NV; j++) out[j] = (float)S[j];
        bdf_store<M, MASK>(a, row, i, out);
    }
    tile_spikes_item_major<ITEMS>(a.out, tile, spk, idx);
    clock_finish(a, step);
}

// Granule BDF update (EgidioGranuleCell_TimeDriven with BDFn, the
// reference's per-neuron step).  The corrector runs a data-dependent number
// of Newton iterations per neuron (about 3 on average, but 10 for the few
// percent whose iterate ends in a one-ulp cycle), so a wave that gave each
// lane a fixed neuron would run each item for its slowest lane's count.
// Here each wave owns a contiguous range of neurons and a lane takes the
// range's next neuron as soon as its own corrector has ended; a neuron's
// arithmetic is the same either way.  The grid is at most one workgroup per
// CU (the Newton step fills the register file); the ranges are equal.
//   * The channel table (104 KB) is staged in LDS: every right-hand side
//     gathers a 26-float row from it.  A lane's first previous-state row
//     (15 floats) is copied to LDS when it takes a neuron, and the BDF
//     coefficients are staged there too.
//   * Spikes set bits of a.bdf.spk_bits; the last workgroup to finish turns
//     them into the per-tile spike lists (ascending, kFlagTile neurons per
//     tile) that the compaction reads, clears the bits and the tiles' fold
//     flags, and resets the workgroup count.
//   * A lane reads its neuron's fold flag when it takes the neuron.
constexpr int kGrBatch = 32;
template <bool EXT>
__global__ __launch_bounds__(kBlock) void k_update_granule_bdf(UpdateArgs a) {
    using G = GranuleModel;
    constexpr unsigned MASK = 15u;   // the synapse types are the run-time mask
    constexpr int TILE = kFlagTile;
    const int32_t step = a.step.get();
    clock_begin(a, step);
    const int par = a.fold.par_by_step ? (step & 1) : a.fold.par;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.fold.bump[par ^ 1] = 0ull;
        // the next step hands out working-storage slots from the other parity
        if (a.bdf.ws_count) a.bdf.ws_count[par ^ 1] = 0u;
    }
    float *__restrict__ row[G::NV];
#pragma unroll
    for (int j = 0; j < G::NV; j++) row[j] = a.state + (int64_t)j * a.stride;
    const int64_t ld = a.stride;
    const uint32_t step_key = EXT ? drive_step_key(a.ext.seed, step) : 0u;
    // each lane's copy of its neuron's first previous-state row, which every
    // corrector iteration reads (2.4% of the cerebellum step against reading
    // it from memory once lanes have drifted apart in their range)
    __shared__ float s_ph0[G::ND * kBlock];
    // the BDF coefficient table (7 x 7), read by every corrector iteration
    __shared__ float s_bdfc[49];
    extern __shared__ float4 s_chan[];
    {
        const float4 *g4 = (const float4 *)a.chan;
        for (int q = threadIdx.x; q < kGrTable * kGrVars / 4; q += kBlock) s_chan[q] = g4[q];
        if (threadIdx.x < 49) s_bdfc[threadIdx.x] = a.bdf.C[threadIdx.x];
    }
    __syncthreads();
    BdfState bl = a.bdf;
    bl.C = s_bdfc;
    RhsCtx c;
    c.chan = (const float *)s_chan;
    c.exp_lut = a.exp_lut;
    c.nmda_lut = a.nmda_lut;
    c.refr = a.refr;
    c.refr_len = a.refr_len;
    c.since_spike = 1 << 30;
    const float dt = a.k.dt;
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t below = (1ull << lane) - 1ull;
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave));
    const int64_t n_waves = (int64_t)gridDim.x * kWavesPerBlock;
    int cursor = (int)((int64_t)a.n * wave / n_waves);
    const int end = (int)((int64_t)a.n * (wave + 1) / n_waves);
    int i = -1;          // neuron on this lane (-1: none)
    int st = 0, k = 0;
    float epsi = 1.0f, prev = 0.0f;
    float S[G::NV], P[G::NV];
    for (;;) {
        bool busy = i >= 0 && epsi > 1e-6 && k < 10;   // the corrector's loop condition
        const uint64_t run = __ballot(busy);
        // lanes whose corrector has ended wait until kGrBatch lanes are free,
        // then all commit and take new neurons together
        if (__popcll(run) <= kWave - kGrBatch) {
            if (i >= 0 && !busy) {
                granule_bdf_commit<MASK>(a.mp, a.bdf, ld, (uint32_t)i, st, S, P);
                const bool fired = G::spike(a.mp, prev, S);
                a.bdf.st[i] = st;
                if (fired) {
                    if (a.last_spike) a.last_spike[i] = step;
                    atomicOr(&a.bdf.spk_bits[i >> 5], 1u << (i & 31));
                }
                bdf_store<G, MASK>(a, row, (uint32_t)i, S);
                i = -1;
            }
            if (cursor < end) {
                const uint64_t idle = __ballot(i < 0);
                if (i < 0) {
                    const int j = cursor + __popcll(idle & below);
                    if (j < end) {
                        i = j;
                        const uint32_t flag = a.fold.tile_flag[j / TILE];
                        bdf_load<G, MASK, EXT>(a, row, step, step_key, flag, (uint32_t)i, S);
                        prev = S[0];
                        st = a.bdf.st[i];
                        if (a.last_spike) c.since_spike = step - 1 - a.last_spike[i];
                        granule_bdf_predict<MASK>(a.mp, c, dt, a.bdf, ld, (uint32_t)i, st, S, P);
                        if (st > 1) {
                            const float *Ph = (const float *)a.bdf.prev;
#pragma unroll
                            for (int q = 0; q < G::ND; q++) s_ph0[q * kBlock + threadIdx.x] = Ph[(int64_t)q * ld + i];
                        }
                        epsi = 1.0f;
                        k = 0;
                        busy = true;
                    }
                }
                cursor += __popcll(idle);
            }
        }
        if (!__any(busy)) break;
        if (busy)
            granule_bdf_newton<MASK>(a.mp, c, dt, bl, par, ld, (uint32_t)i, st, S, P, epsi, k, s_ph0 + threadIdx.x);
    }
    __shared__ uint32_t s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(a.bdf.done, 1u) == gridDim.x - 1u;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        for (int t = threadIdx.x; t < a.n_tiles; t += kBlock) {
            int32_t cnt = 0;
            for (int w = 0; w < TILE / 32; w++) {
                uint32_t bits = atomicExch(&a.bdf.spk_bits[t * (TILE / 32) + w], 0u);
                while (bits) {
                    const int b = __builtin_ctz(bits);
                    bits &= bits - 1u;
                    a.out.staging[(int64_t)t * TILE + cnt++] = t * TILE + w * 32 + b;
                }
            }
            a.out.tile_count[t] = cnt;
            clear_tile_flag<TILE>(a.fold.tile_flag, t);
        }
        if (threadIdx.x == 0) *a.bdf.done = 0u;
    }
    clock_finish(a, step);
}

// ---------------------------------------------------------------------------
// Poisson generator device (PoissonGeneratorDeviceVector.cpp:56-70): per
// source, if rate > 0 draw the next frand() of its LCG
// (RandomGenerator.h:61-107) and fire when rate*step (double) exceeds it.
// ---------------------------------------------------------------------------
struct PoissonArgs {
    uint32_t *lcg;
    int n;
    int n_tiles;
    float rate;             // frequency of every generator (Hz) when rates is null
    const float *rates;     // per-generator frequencies (PoissonGeneratorDeviceVector)
    double step_s;          // time-driven step (s)
    TileOut out;
};

__global__ __launch_bounds__(kBlock) void k_poisson(PoissonArgs a) {
    const int tile = blockIdx.x;
    bool spk[1] = {false};
    int idx[1] = {tile * kBlock + (int)threadIdx.x};
    const int i = idx[0];
    const float f = (i < a.n && a.rates) ? a.rates[i] : a.rate;
    if (i < a.n && f > 0.0f) {
        uint32_t x = a.lcg[i];
        x = (uint32_t)(((uint64_t)x * 1103515245ull + 12345ull) & 0x7fffffffull);
        a.lcg[i] = x;
        const float u = ((float)x) / 2147483647.0f;
        spk[0] = (double)f * a.step_s > (double)u;
    }
    tile_spikes<1>(a.out, tile, spk, idx);
}

// ---------------------------------------------------------------------------
// Spike-train source: input neurons (InputSpikeNeuronModel driven by
// FileInputSpikeDriver.cpp:45-85 or Simulation_API::AddExternalSpikeActivity)
// whose spikes the host has placed on update steps.  Update k copies its
// ascending, duplicate-free list idx[ptr[k-1] .. ptr[k]) into ring slot
// k % ring; steps past the train have no spikes.
// ---------------------------------------------------------------------------
struct TrainArgs {
    const int64_t *ptr;     // [n_steps + 1]
    const int32_t *idx;
    int64_t n_steps;
    int32_t *ring_idx;      // [ring][n_src]
    int32_t *ring_cnt;      // [ring]
    int64_t n_src;
    int ring;
    StepRef step;
};

__global__ __launch_bounds__(kBlock) void k_train(TrainArgs a) {
    const int32_t k = a.step.get();
    const int slot = k % a.ring;
    int64_t b = 0, e = 0;
    if (k >= 1 && (int64_t)k <= a.n_steps) {
        b = a.ptr[k - 1];
        e = a.ptr[k];
    }
    int32_t *out = a.ring_idx + (int64_t)slot * a.n_src;
    for (int64_t j = b + (int64_t)blockIdx.x * kBlock + threadIdx.x; j < e; j += (int64_t)gridDim.x * kBlock)
        out[j - b] = a.idx[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ring_cnt[slot] = (int32_t)(e - b);
}

// Events pushed by the host: order = position in the host's list.  Events
// handed over in device memory are checked here (target in range, synapse
// type enabled); one that fails is dropped and sets err bit 8.
struct PushEvents {
    const int32_t *tgt;
    const uint8_t *type;
    const float *w;
    int n;
    int32_t n_tgt;
    uint32_t mask;       // enabled synapse types taking spikes (bits 0-2)
    uint32_t *err;
};

template <int PASS>
__global__ __launch_bounds__(kBlock) void k_fold_push(FoldTarget t, PushEvents e) {
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < e.n; q += gridDim.x * kBlock) {
        const int32_t tg = e.tgt[q];
        const uint32_t ty = e.type[q];
        if ((uint32_t)tg >= (uint32_t)e.n_tgt || ty > 2u || !((e.mask >> ty) & 1u)) {
            if (PASS == kCount) atomicOr(e.err, 256u);
            continue;
        }
        const uint64_t key = (uint64_t)ty * (uint64_t)t.stride + (uint64_t)tg;
        fold_event<PASS>(t, t.par, key, (uint64_t)q, e.w[q]);
    }
}

// Events generated on the device from the spike rings of the sources.
constexpr int kMaxSources = 4;
constexpr int kMaxDelay = 256;
// Additive STDP on the synapses of a source (STDPWeightChange / STDPState):
// per synapse a presynaptic and a postsynaptic trace and the time of its
// last update; weights change in place.
struct StdpSyn {
    float *pre;
    float *post;
    double *last;
    float a_plus, a_minus;           // max_LTP, max_LTD
    float inv_tau_plus, inv_tau_minus;
    float w_max;
};

struct RingSource {
    const int32_t *ring_idx;   // [ring][n_src]
    const int32_t *ring_cnt;   // [ring]
    const int64_t *row_ptr;    // [n_src*max_delay+1]
    const uint32_t *tgt_type;  // target | type << 30
    float *w;
    int64_t n_src;
    int stdp;                  // plastic synapses (StdpSyn below)
    StdpSyn syn;
    // groups of 1,024 spiking sources (the reference's PropagatedSpikeGroup):
    // the chunk of a list entry is its position / 1024 unless bnd_on[slot],
    // then the number of chunks whose last member comes before it
    // (bnd [ring][max_delay][cmax]); a spike train's events are one per spike
    int train;
    const int32_t *bnd;
    const uint8_t *bnd_on;
    int32_t cmax;
    int poisson;               // a Poisson device: its groups' times are on its own clock
};
struct RingEvents {
    RingSource src[kMaxSources];
    int n_sources;
    int max_delay;
    int ring;
    StepRef step;              // update index about to run
    // STDP only: update time of each ring slot (populations; Poisson
    // devices on their own 0.1 ms clock), delay of d steps in seconds
    const double *time_ring;   // [ring]
    const double *time_ring_p; // [ring]
    const double *delay_s;     // [max_delay + 1]
    const float *exp_lut;
    // a run that stops inside a window delivers the groups ranked before its
    // stop event at its end (rank_hi) and the rest with the next update:
    // that update's fold skips the groups ranked below skip[1] when the
    // update is skip[0] (device memory, set by the host); rtab null: all
    const RankTab *rtab;
    const uint32_t *skip;
    uint32_t rank_hi;
};

// STDPState::SetNewUpdateTime (STDPState.cpp:74-82): both traces decay by
// ExponentialTable factors of the float elapsed time.
__device__ __forceinline__ void stdp_decay(const StdpSyn &s, const float *exp_lut, int64_t q, double t) {
    const float elapsed = (float)(t - s.last[q]);
    s.pre[q] = s.pre[q] * lut_exp(exp_lut, -elapsed * s.inv_tau_plus);
    s.post[q] = s.post[q] * lut_exp(exp_lut, -elapsed * s.inv_tau_minus);
    s.last[q] = t;
}
// Interconnection::IncrementWeight (Interconnection.h:400-409)
__device__ __forceinline__ void stdp_increment(const StdpSyn &s, float *w, int64_t q, float inc) {
    float x = w[q] + inc;
    if (x > s.w_max) x = s.w_max;
    else if (x < 0.0f) x = 0.0f;
    w[q] = x;
}

// Work item = one spike of one source at emission step k-1-d; a group of
// LANES lanes walks its synapses of delay d.  Items are enumerated in
// (d, source, rank) order from a prefix over the per-(d, source) counts.
// The order key (ord_key) holds delay, source, chunk and synapse position;
// within a group the synapse position rises with the source index and then
// the target index, which is the reference's delivery order in a group.
template <int PASS, int LANES>
__device__ __forceinline__ void fold_ring_body(const FoldTarget &t, const RingEvents &r, int bid, int nb) {
    __shared__ int64_t s_pref[kMaxDelay * kMaxSources + 1];
    __shared__ int s_slot[kMaxDelay * kMaxSources];
    const int nseg = r.max_delay * r.n_sources;
    const int32_t step = r.step.get();
    const int par = t.par_by_step ? (step & 1) : t.par;
    const uint32_t rank_lo = (r.rtab && r.skip && r.skip[0] == (uint32_t)step) ? r.skip[1] : 0u;
    for (int g = threadIdx.x; g < nseg; g += kBlock) {
        const int d = g / r.n_sources + 1;
        const int p = g % r.n_sources;
        const int64_t s = (int64_t)step - 1 - d;
        int slot = 0, c = 0;
        if (s >= 1) {
            slot = (int)(s % r.ring);
            c = r.src[p].ring_cnt[slot];
        }
        s_slot[g] = slot;
        s_pref[g + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_pref[0] = 0;
        for (int g = 0; g < nseg; g++) s_pref[g + 1] += s_pref[g];
    }
    __syncthreads();
    const int64_t n_items = s_pref[nseg];
    const int lane = threadIdx.x % LANES;
    const int64_t group = (int64_t)(bid * kBlock + (int)threadIdx.x) / LANES;
    const int64_t n_groups = (int64_t)nb * kBlock / LANES;
    for (int64_t it = group; it < n_items; it += n_groups) {
        // last segment g with s_pref[g] <= it (segments may be empty)
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pref[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const int g = lo;
        const int d = g / r.n_sources + 1;
        const int p = g % r.n_sources;
        const RingSource &src = r.src[p];
        const int64_t rank = it - s_pref[g];
        const int32_t id = src.ring_idx[(int64_t)s_slot[g] * src.n_src + rank];
        const int64_t b = src.row_ptr[(int64_t)id * r.max_delay + d - 1];
        const int64_t e = src.row_ptr[(int64_t)id * r.max_delay + d];
        uint32_t chunk = 0;
        if (src.train) {
            chunk = (uint32_t)rank;
        } else if (s_pref[g + 1] - s_pref[g] >= 1024) {
            const int slot = s_slot[g];
            if (src.bnd_on && src.bnd_on[slot]) {
                const int32_t *bd = src.bnd + ((int64_t)slot * r.max_delay + (d - 1)) * src.cmax;
                while ((int32_t)chunk < src.cmax && bd[chunk] < rank) chunk++;
            } else {
                chunk = (uint32_t)(rank >> 10);
            }
        }
        const uint64_t ord_hi = ord_key(d, p, chunk, 0);
        if (rank_lo > 0u || r.rank_hi != 0xffffffffu) {
            const uint32_t rk = key_rank(r.rtab, step & 1, ord_hi);
            if (rk < rank_lo || rk >= r.rank_hi) continue;
        }
        double t_arr = 0.0;
        if (PASS != kCount && src.stdp)
            t_arr = (src.poisson ? r.time_ring_p[s_slot[g]] : r.time_ring[s_slot[g]]) + r.delay_s[d];
        for (int64_t q = b + lane; q < e; q += LANES) {
            const uint32_t tt = src.tgt_type[q];
            const uint64_t key =
                (uint64_t)(tt >> 30) * (uint64_t)t.stride + (uint64_t)(tt & 0x3fffffffu);
            fold_event<PASS>(t, par, key, ord_hi | (uint64_t)q, src.w[q]);
            if (PASS != kCount && src.stdp) {
                // STDPWeightChange::ApplyPreSynapticSpike (STDPWeightChange.cpp:43-55),
                // after the event used the weight (PropagatedSpikeGroup.cpp:76-97).
                // A synapse carries at most one event per step: no other lane
                // touches q.
                stdp_decay(src.syn, r.exp_lut, q, t_arr);
                src.syn.pre[q] = src.syn.pre[q] + 1.0f;
                stdp_increment(src.syn, src.w, q, -src.syn.a_minus * src.syn.post[q]);
            }
        }
    }
}

template <int PASS, int LANES>
__global__ __launch_bounds__(kBlock) void k_fold_ring(FoldTarget t, RingEvents r) {
    fold_ring_body<PASS, LANES>(t, r, blockIdx.x, gridDim.x);
}

// Compaction of update k's spikes and the fold of the events update k + 1
// consumes, in one launch (network mode without STDP): the fold reads the
// spike lists of updates <= k - 1 only, and the fold buffers it fills were
// drained by update k.  Workgroups [0, nb_compact) compact, the rest fold.
template <int LANES>
__global__ __launch_bounds__(kBlock) void k_compact_fold(CompactArgs c, FoldTarget t, RingEvents r,
                                                         int nb_compact) {
    clock_next_kernel(c);
    if ((int)blockIdx.x < nb_compact) compact_body(c, blockIdx.x, nb_compact);
    else fold_ring_body<kStatic, LANES>(t, r, blockIdx.x - nb_compact, gridDim.x - nb_compact);
}

// The same with the event-queue replay of update k's window in the first
// workgroup (its chain of dependent steps is the longest): it reads update
// k's spikes from the update kernel's tiles, so it depends on nothing the
// other workgroups write and runs beside them.
template <int LANES>
__global__ __launch_bounds__(kBlock) void k_compact_fold_eq(CompactArgs c, FoldTarget t, RingEvents r,
                                                            int nb_compact, EqArgs q) {
    clock_next_kernel(c);
    const int b = (int)blockIdx.x - 1;
    if (b < 0) eq_body<kEqFusedItems, kEqFusedDelays>(q);
    else if (b < nb_compact) compact_body(c, b, nb_compact);
    else fold_ring_body<kStatic, LANES>(t, r, b - nb_compact, (int)gridDim.x - 1 - nb_compact);
}

// Synapses per (type, target) key of one source, for the fixed segments of
// the network-mode fold.
__global__ __launch_bounds__(kBlock) void k_key_degree(const uint32_t *tgt_type, int64_t nsyn, int64_t stride,
                                                       uint32_t *deg) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < nsyn; q += (int64_t)gridDim.x * kBlock) {
        const uint32_t tt = tgt_type[q];
        atomicAdd(&deg[(uint64_t)(tt >> 30) * (uint64_t)stride + (tt & 0x3fffffffu)], 1u);
    }
}

// ---------------------------------------------------------------------------
// Post-synaptic STDP (STDPWeightChange::ApplyPostSynapticSpike,
// STDPWeightChange.cpp:57-76, called for every neuron that fired at update k,
// TimeDrivenInternalSpike.cpp:94-103): every plastic input synapse of a
// spiking neuron decays to the update time, gains a postsynaptic trace unit
// and potentiates by max_LTP times its presynaptic trace (in_syn lists
// synapse positions by target).
// ---------------------------------------------------------------------------
struct StdpPostArgs {
    const int32_t *spk_idx;    // own spike ring, slot k % ring
    const int32_t *spk_cnt;
    int64_t slot_len;
    int ring;
    StepRef step;
    const int64_t *in_ptr;     // [n_tgt + 1]
    const int64_t *in_syn;     // synapse positions grouped by target
    int chunks;                // kWave-synapse chunks of the longest list
    float *w;
    StdpSyn syn;
    const double *time_ring;
    const float *exp_lut;
};

// Post-synaptic learning of the neurons that fired at update k: every
// incoming synapse of each of them.  A work
// item is one wave-sized chunk of one neuron's list, so a neuron with 2,000
// inputs is spread over 32 waves rather than walked by one; the synapses are
// distinct, so the order of their updates does not matter.
__global__ __launch_bounds__(kBlock) void k_stdp_post(StdpPostArgs a) {
    const int32_t step = a.step.get();
    const int slot = step % a.ring;
    const int64_t n = a.spk_cnt[slot];
    const int32_t *ids = a.spk_idx + (int64_t)slot * a.slot_len;
    const double tk = a.time_ring[slot];
    const int lane = threadIdx.x % kWave;
    const int64_t wave = (int64_t)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int64_t n_waves = (int64_t)gridDim.x * kBlock / kWave;
    for (int64_t it = wave; it < n * a.chunks; it += n_waves) {
        const int32_t i = ids[it / a.chunks];
        const int64_t x = a.in_ptr[i] + (it % a.chunks) * kWave + lane;
        if (x < a.in_ptr[i + 1]) {
            const int64_t q = a.in_syn[x];
            stdp_decay(a.syn, a.exp_lut, q, tk);
            a.syn.post[q] = a.syn.post[q] + 1.0f;
            stdp_increment(a.syn, a.w, q, a.syn.a_plus * a.syn.pre[q]);
        }
    }
}

// Update time of step k: t_k = t_{k-1} + double(float dt), kept per ring slot
// (TimeEventUpdateNeurons.cpp:78).  First node of a step of a plastic network.
// The Poisson devices' clock advances by their own step in a second ring.
__global__ void k_tick(StepRef step, double *time_ring, double *time_ring_p, int ring, double dt, double dt_p) {
    if (threadIdx.x == 0) {
        const int32_t k = step.get();
        time_ring[k % ring] = time_ring[(k - 1) % ring] + dt;
        time_ring_p[k % ring] = time_ring_p[(k - 1) % ring] + dt_p;
    }
}

// Adds every queued event of the population's keys to its conductances, in
// order (the groups delivered before a run's stop event), and clears the
// keys and fold flags.
__global__ __launch_bounds__(kBlock) void k_apply_fold(FoldTarget t, int32_t n, int par) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        for (int ty = 0; ty < 3; ty++) {
            const uint64_t key = (uint64_t)ty * (uint64_t)t.stride + (uint64_t)i;
            const uint32_t m = t.fill[key];
            if (m == 0u) continue;
            float *g = g_ptr(t, key);
            *g = fold_static(t.fill, t.seg_off, t.b_ent, t.rtab, par, key, m, *g);
        }
    }
}

}  // namespace edl
