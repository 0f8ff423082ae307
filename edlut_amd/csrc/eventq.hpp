// eventq.hpp -- the order in which the reference's event queue delivers a
// step's spike groups, replayed on the device.
//
// The reference runs one binary heap of events (EventQueue.cpp:121-198)
// ordered by time, then EventPriority (Event.h:37); equal keys leave in the
// order the heap's layout gives.  Every time-driven model has a stream of
// update events at its own step; an update gathers the model's spiking
// neurons, in ascending index, into one PropagatedSpikeGroup per delay of the
// model's delay structure, queueing a group as soon as it holds 1,024 sources
// and the rest afterwards in ascending delay (TimeDrivenInternalSpike.cpp
// 66-104,163-178, PropagatedSpikeGroup.h:54).  Input neurons' spikes are
// InputSpike events that chain one PropagatedSpike per delay of the neuron
// (InputSpike.cpp, PropagatedSpike.cpp:119-160).  A run is RunSimulationSlot
// (Simulation.cpp:322-373): a stop event at the stop time, then events until
// it leaves.
//
// k_eventq replays exactly these heap operations, once per update k, after
// the spikes of update k are compacted: it pops the events that reach the
// populations between update k and update k + 1 and gives each group (or
// propagated input spike) its rank in that window.  The fold of update k + 1
// sorts each target's events by (rank, synapse), which is the reference's
// delivery order.  The heap holds only keys and small payloads; one wave
// runs its operations lane-parallel, after the workgroup has counted the
// step's spikes per delay.
#pragma once
#include "device_common.hpp"

namespace edl {

constexpr uint32_t kEqPrioStop = 1, kEqPrioSpike = 8, kEqPrioTime = 10;
enum EqKind : uint32_t { kEqTime = 0, kEqGroup = 1, kEqInput = 2, kEqProp = 3, kEqStop = 4 };
enum EqEmitterKind : int { kEqPop = 0, kEqPoisson = 1, kEqTrain = 2 };
constexpr int kEqMaxEmitters = 16;
constexpr int kEqMaskWords = 8;          // delays up to 256 steps
constexpr int kEqChunk = 1024;           // PropagatedSpikeGroup::MaxSize
constexpr int kEqLdsItems = 2048;        // heap entries staged in LDS (larger heaps stay in memory)

// Heap entry: time, then a = priority<<28 | kind<<24 | emitter<<16 | d (or j),
// b = chunk (groups) or input record (input and propagated spikes).
struct EqItem {
    double t;
    uint32_t a;
    uint32_t b;
};

__host__ __device__ inline EqItem eq_item(double t, uint32_t prio, uint32_t kind, uint32_t em, uint32_t d, uint32_t b) {
    return EqItem{t, (prio << 28) | (kind << 24) | (em << 16) | (d & 0xffffu), b};
}
__host__ __device__ inline uint32_t eq_prio(const EqItem &x) { return x.a >> 28; }
__host__ __device__ inline uint32_t eq_kind(const EqItem &x) { return (x.a >> 24) & 0xfu; }
__host__ __device__ inline uint32_t eq_em(const EqItem &x) { return (x.a >> 16) & 0xffu; }
__host__ __device__ inline uint32_t eq_d(const EqItem &x) { return x.a & 0xffffu; }

// EventQueue::InsertEvent (EventQueue.cpp:121-141) on a 1-based array of
// `*n` entries: the reference swaps the new element up while its parent is
// later (or as late and of lower priority); the same moves, with the new
// element held aside until its place is known.
__host__ __device__ inline bool eq_later(const EqItem &p, const EqItem &c) {
    return p.t > c.t || (p.t == c.t && eq_prio(p) < eq_prio(c));
}
__host__ __device__ __forceinline__ void eq_insert(EqItem *h, int32_t *n, const EqItem &x) {
    int32_t c = ++*n;
    while (c > 1) {
        const EqItem p = h[c / 2];
        if (!eq_later(p, x)) break;
        h[c] = p;
        c /= 2;
    }
    h[c] = x;
}

// EventQueue::RemoveEvent (EventQueue.cpp:162-198): the last element moves
// to the top and is swapped down with the earlier child (ties: the left one
// unless the right has higher priority) while that child comes before it;
// again with the moving element held aside.
__host__ __device__ __forceinline__ bool eq_remove(EqItem *h, int32_t *n, EqItem *out) {
    if (*n >= 2) {
        *out = h[1];
        const EqItem x = h[*n];
        --*n;
        const int32_t size = *n;
        int32_t p = 1, c;
        for (c = 2; c < size; p = c, c = p * 2) {
            EqItem a = h[c];
            const EqItem b = h[c + 1];
            if (eq_later(a, b)) {
                c++;
                a = b;
            }
            if (a.t < x.t || (a.t == x.t && eq_prio(a) > eq_prio(x))) h[p] = a;
            else break;
        }
        if (c == size) {
            const EqItem a = h[c];
            if (a.t < x.t || (a.t == x.t && eq_prio(a) > eq_prio(x))) {
                h[p] = a;
                p = c;
            }
        }
        h[p] = x;
        return true;
    }
    if (*n == 1) {
        *out = h[1];
        *n = 0;
        return true;
    }
    return false;
}

// One model whose events the replay follows.
struct EqEmitter {
    int kind;                  // EqEmitterKind
    int n;                     // neurons
    double h;                  // update step (s): populations double(float dt), Poisson 0.1f/1000.0f
    const int32_t *ring_idx;   // spike lists by update slot (populations, Poisson devices)
    const int32_t *ring_cnt;
    int64_t slot_len;
    // the update's spikes still in the update kernel's tiles (a population
    // replayed inside its own compaction kernel): tile counts and staged
    // indices, ascending within a tile (tile_count null: the ring lists)
    const int32_t *tile_count;
    const int32_t *staging;
    int32_t n_tiles, tile_size;
    const uint32_t *dmask;     // [n][words]: bit d-1 set when the neuron has synapses of delay d
    uint32_t full[kEqMaskWords];   // the model's delay structure (union of the neurons')
    int all_full;              // every neuron has every delay of `full` (counts per delay = spikes)
    int cmax;                  // rank-table entries per delay (chunks, or input spikes per update)
    uint32_t *rank;            // [2][max_delay][cmax]: rank in its window, by consuming-update parity
    int32_t *bnd;              // [ring][max_delay][cmax]: list position of the last member of chunk c
    uint8_t *bnd_on;           // [ring]: the slot's chunks are in bnd (else every 1,024 list entries)
    // input spikes (trains): records in load order
    const double *in_t;
    const int32_t *in_neuron;
    const int32_t *in_pos;     // position in its update's list
    const int32_t *nd_ptr;     // per neuron: its delays (ascending), CSR into nd_list
    const uint16_t *nd_list;
    const int64_t *in_seq;     // load sequence number of each record (queue insertion order)
    int32_t in_lo, in_hi;      // records inserted by a resuming call
};

// Replay state in device memory: [0] heap size, [1] rank counter of the
// window, [2] the rank counter when the run's stop event left (-1: not yet),
// [3] the update whose window clears its rank table first, [4..4+n_em) the
// next update of each model.
constexpr int kEqSt = 4;

struct EqArgs {
    const EqEmitter *em;       // [n_em] in device memory
    int n_em;
    int words;                 // mask words per neuron
    int max_delay;
    int ring;
    const double *delay_s;     // [max_delay + 1] RoundPropagationDelay of d steps
    EqItem *heap;              // [cap + 1], 1-based
    int32_t *st;
    int32_t cap;
    double stop_t;             // resuming call: the run's stop time
    int resume;                // 1: insert the pending input spikes and the stop event first
    int with_stop;             // resuming call: 0 when the run continues the previous one (no stop event)
    int32_t clear_step;        // resuming call: the run's last update when its stop time is explicit, else -1
    StepRef step;              // update k just run (resuming call: the update before the run)
    const int32_t *pf_tiles;   // tile counts of model pf_em, loaded at the start (null: none)
    int pf_em;
    uint32_t *err;             // bit 6: heap capacity, bit 7: event of an unknown kind
};

// Pointers read from LDS (the staged model descriptions) are generic; the
// replay casts them to global memory so that their loads and stores are not
// flat accesses, which the LDS waits of the heap operations would also wait on.
#define EQ_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ EQ_GLOBAL T *eq_g(T *p) {
    return (EQ_GLOBAL T *)p;
}

__device__ inline bool eq_has(const uint32_t *m, int d) {
    return (eq_g(m)[(d - 1) >> 5] >> ((d - 1) & 31)) & 1u;
}

constexpr int kMaxDelayEq = 32 * kEqMaskWords;

// The heap operations of one window, run by one wave with wave-uniform
// control flow (always inlined, so that a heap staged in LDS is addressed as
// LDS).  Each heap operation is the reference's, lane-parallel:
//  - removal: every lane takes one node p of the top 64 and picks its child
//    as the reference's sift-down would (the earlier child, the left one on
//    ties), and whether that child moves up past the displaced last element;
//    the path is then a chain of readlanes from the root, and the nodes on
//    it take their chosen child's entry in one write.  Nodes below the top
//    64 (heaps of more than 129 entries) continue one level at a time.
//  - insertion: lane j reads the j-th ancestor of the new position; the
//    ancestors later than the new entry up to the first that is not (the
//    reference's sift-up stops there) move down one level in one write.
// Children of node p are read before any entry on the path is written, as in
// the reference, where the entries written lie above them.
struct EqWin {
    int32_t n;          // heap size
    uint32_t r;         // rank counter of the window
    int32_t stop_r;     // rank counter when the stop event left (-1: it did not)
    bool overflow, unknown;
};

__device__ __forceinline__ bool eq_before(const EqItem &a, const EqItem &x) {
    return a.t < x.t || (a.t == x.t && eq_prio(a) > eq_prio(x));
}

// The same comparisons without short-circuit branches (lane-parallel code).
__device__ __forceinline__ bool eq_later_v(const EqItem &p, const EqItem &c) {
    return (p.t > c.t) | ((p.t == c.t) & (eq_prio(p) < eq_prio(c)));
}
__device__ __forceinline__ bool eq_before_v(const EqItem &a, const EqItem &x) {
    return (a.t < x.t) | ((a.t == x.t) & (eq_prio(a) > eq_prio(x)));
}

__device__ __forceinline__ int eq_lane() { return (int)(threadIdx.x & 63u); }

// Lane 0's entry, in scalar registers.
__device__ __forceinline__ EqItem eq_lane0(const EqItem &v) {
    const uint64_t tb = (uint64_t)__double_as_longlong(v.t);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)tb);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(tb >> 32));
    EqItem r;
    r.t = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    r.a = (uint32_t)__builtin_amdgcn_readfirstlane((int)v.a);
    r.b = (uint32_t)__builtin_amdgcn_readfirstlane((int)v.b);
    return r;
}

// Inserts x; `top` follows the heap's first entry.
__device__ __forceinline__ void eq_insert_wave(EqItem *__restrict__ h, int32_t &n, const EqItem &x, EqItem &top) {
    const int lane = eq_lane();
    const int32_t c = ++n;
    const int depth = 31 - __builtin_clz((uint32_t)c);     // ancestors of c
    // lanes past the root read entry 0 (unused) and take no part
    const EqItem anc = h[lane < depth ? c >> (lane + 1) : 0];
    const bool later = (lane < depth) & eq_later_v(anc, x);
    const uint64_t lm = __ballot(later);
    const int m = (int)__builtin_ctzll(~lm);                // ancestors that move down
    if (lane < m) h[c >> lane] = anc;
    if (lane == 0) h[c >> m] = x;
    if ((c >> m) == 1) top = x;
}

// The entries a removal reads (the last entry and the children of the top
// 64 nodes), loaded before the caller decides to remove.
struct EqRemoveLoads {
    EqItem x, ch;
    int32_t c;
    bool has_l;
};

__device__ __forceinline__ EqRemoveLoads eq_remove_loads(const EqItem *__restrict__ h, int32_t n) {
    const int32_t p0 = eq_lane() + 1;
    const int32_t size = n - 1 > 0 ? n - 1 : 1;
    // both children of node p0, read unconditionally (indices clamped into
    // the heap), then chosen with selects
    const int32_t cl = 2 * p0, cr = 2 * p0 + 1;
    const EqItem L = h[cl <= size ? cl : size];
    const EqItem R = h[cr <= size ? cr : size];
    const bool has_r = cr <= n - 1;
    const bool pick_r = has_r & eq_later_v(L, R);
    EqRemoveLoads r;
    r.x = h[n];
    r.ch.t = pick_r ? R.t : L.t;
    r.ch.a = pick_r ? R.a : L.a;
    r.ch.b = pick_r ? R.b : L.b;
    r.c = pick_r ? cr : cl;
    r.has_l = cl <= n - 1;
    return r;
}

// Removes the top (n >= 1) with the entries `eq_remove_loads` read; `top`
// becomes the new first entry.  Two ballots give, for each of the top 64
// nodes, whether its chosen child moves up and whether that child is the
// right one; every lane then tests its node's ancestors against those bits,
// so the path comes out without a chain of dependent steps.
__device__ __forceinline__ void eq_remove_wave(EqItem *__restrict__ h, int32_t &n, const EqRemoveLoads &r,
                                               EqItem &top) {
    const int lane = eq_lane();
    if (n == 1) {
        n = 0;
        return;
    }
    const int32_t size = --n;
    const bool ok = r.has_l & eq_before_v(r.ch, r.x);
    const uint64_t moves = __ballot(ok);
    const uint64_t right = __ballot((r.c & 1) != 0);
    // node p0 (lane + 1) is on the path when every ancestor moves its child
    // up and chose the child toward p0
    const uint32_t p0 = (uint32_t)lane + 1u;
    bool on = true;
#pragma unroll
    for (int j = 1; j <= 6; j++) {
        const uint32_t q = p0 >> j;          // ancestor j levels up (0: none)
        const uint32_t sh = (q - 1u) & 63u;
        const bool step = (q == 0u) | ((((moves >> sh) & 1ull) != 0ull) &
                                       ((uint32_t)((right >> sh) & 1ull) == ((p0 >> (j - 1)) & 1u)));
        on = on & step;
    }
    const uint64_t path = __ballot(on & ok);
    const uint64_t ends = __ballot(on & !ok);
    int32_t p;
    if (ends) {
        p = (int32_t)__builtin_ctzll(ends) + 1;
    } else {
        // the path leaves the top 64 nodes below its deepest node there
        const uint64_t out = __ballot(on & ok & (r.c > 64));
        p = __builtin_amdgcn_readlane(r.c, (int)__builtin_ctzll(out));
    }
    if ((path >> lane) & 1ull) h[lane + 1] = r.ch;
    top = (moves & 1ull) ? eq_lane0(r.ch) : r.x;
    // below the top 64 nodes: one level at a time
    while (p > 64 && 2 * p <= size) {
        int32_t cc = 2 * p;
        EqItem a = h[cc];
        if (cc + 1 <= size) {
            const EqItem b = h[cc + 1];
            if (eq_later(a, b)) {
                a = b;
                cc++;
            }
        }
        if (!eq_before(a, r.x)) break;
        if (lane == 0) h[p] = a;
        p = cc;
    }
    if (lane == 0) h[p] = r.x;
}

// Ranks given in a window, listed in LDS and written to the models' rank
// tables by the whole wave (entries: model << 16 | delay, chunk or input
// record, rank, 1 for a propagated input spike).
constexpr int kEqRankList = 256;

template <int MAXD>
__device__ __forceinline__ void eq_window(EqItem *__restrict__ h, const EqArgs &a, const EqEmitter *em,
                                          const uint32_t (*cnt)[MAXD], const uint32_t (*pmask)[MAXD / 32 + 1],
                                          const uint32_t *anyfull, uint32_t *cur, int32_t *next_k,
                                          const double *delay_s, uint4 *rk, int32_t k, EqWin &w) {
    const int lane = eq_lane();
    const int par = (k + 1) & 1;          // parity of the update that consumes this window
    EqItem top = h[1];                    // the heap's first entry, kept in registers
    int32_t nrk = 0;
    auto flush = [&]() {
        for (int i = lane; i < nrk; i += 64) {
            const uint4 q = rk[i];
            const EqEmitter &F = em[q.x >> 16];
            const int d = (int)(q.x & 0xffffu);
            const int32_t c = q.w ? eq_g(F.in_pos)[q.y] : (int32_t)q.y;
            eq_g(F.rank)[((int64_t)par * a.max_delay + (d - 1)) * F.cmax + c] = q.z;
        }
        nrk = 0;
    };
    auto give_rank = [&](uint32_t em_d, uint32_t b, uint32_t prop) {
        if (lane == 0) rk[nrk] = make_uint4(em_d, b, w.r, prop);
        w.r++;
        if (++nrk == kEqRankList) flush();
    };
    auto push = [&](const EqItem &x) {
        if (w.n >= a.cap) {
            w.overflow = true;
            return;
        }
        eq_insert_wave(h, w.n, x, top);
    };
    if (a.resume) {
        // input spikes added since the last run (Simulation_API::
        // AddExternalSpikeActivity) are queued in load order, then the run's
        // stop event
        int32_t pos[kEqMaxEmitters];
        for (int e = 0; e < a.n_em; e++) pos[e] = em[e].in_lo;
        for (;;) {
            int be = -1;
            int64_t bs = 0;
            for (int e = 0; e < a.n_em; e++) {
                if (em[e].kind != kEqTrain || pos[e] >= em[e].in_hi) continue;
                const int64_t sq = eq_g(em[e].in_seq)[pos[e]];
                if (be < 0 || sq < bs) {
                    be = e;
                    bs = sq;
                }
            }
            if (be < 0) break;
            const EqEmitter &E = em[be];
            const int32_t q = pos[be]++;
            if (eq_g(E.nd_ptr)[eq_g(E.in_neuron)[q] + 1] > eq_g(E.nd_ptr)[eq_g(E.in_neuron)[q]])
                push(eq_item(eq_g(E.in_t)[q], kEqPrioSpike, kEqInput, (uint32_t)be, 0, (uint32_t)q));
        }
        if (a.with_stop) push(eq_item(a.stop_t, kEqPrioStop, kEqStop, 0, 0, 0));
    }
    uint32_t pops = 0;                    // models that are populations
    for (int e = 0; e < a.n_em; e++) pops |= (em[e].kind == kEqPop ? 1u : 0u) << e;
    while (w.n > 0 && !w.overflow) {
        const EqItem it = top;
        const EqRemoveLoads rl = eq_remove_loads(h, w.n);
        const uint32_t top_a = __builtin_amdgcn_readfirstlane(it.a);
        const uint32_t kind = (top_a >> 24) & 0xfu, e = (top_a >> 16) & 0xffu;
        // the window ends where the populations' next update starts
        if (kind == kEqTime && ((pops >> e) & 1u) && next_k[e] > k) break;
        eq_remove_wave(h, w.n, rl, top);
        if (kind == kEqStop) {
            w.stop_r = (int32_t)w.r;
            break;
        }
        const EqEmitter &E = em[e];
        if (kind == kEqTime) {
            // TimeDrivenInternalSpike::ProcessEvent: full groups as they fill
            // (ascending list position, then ascending delay), then the rest
            // in ascending delay
            const int32_t ke = next_k[e];
            if ((pops >> e) & 1u) w.r = 0;   // a window starts with the populations' updates
            const int slot = ke % a.ring;
            const uint32_t *c_e = cnt[e];
            const bool any_full = anyfull[e] != 0u;
            if (any_full) {
                if (lane == 0)
                    for (int d = 1; d <= a.max_delay; d++) cur[d - 1] = 0u;
            }
            const bool use_bnd = any_full && eq_g(E.bnd_on)[slot] != 0;
            while (any_full) {
                int32_t best = 0x7fffffff;
                int bd = 0;
                for (int d = 1; d <= a.max_delay; d++) {
                    const uint32_t c = cur[d - 1];
                    if (c >= c_e[d - 1] / (uint32_t)kEqChunk) continue;
                    const int32_t end = use_bnd ? eq_g(E.bnd)[((int64_t)slot * a.max_delay + (d - 1)) * E.cmax + c]
                                                : (int32_t)(c * kEqChunk + kEqChunk - 1);
                    if (end < best) {
                        best = end;
                        bd = d;
                    }
                }
                if (!bd) break;
                const uint32_t cb = cur[bd - 1];
                push(eq_item(it.t + delay_s[bd], kEqPrioSpike, kEqGroup, e, (uint32_t)bd, cb));
                if (lane == 0) cur[bd - 1] = cb + 1u;
            }
            // the partial groups: delays whose count is not a multiple of
            // 1,024, 64 delays at a time (lane L holds the time of delay L + 1)
            for (int j = 0; j * 64 < a.max_delay; j++) {
                uint64_t bits = (uint64_t)pmask[e][2 * j];
                if (2 * j + 1 < MAXD / 32 + 1) bits |= (uint64_t)pmask[e][2 * j + 1] << 32;
                if (!bits) continue;
                const int dl = j * 64 + lane + 1;
                const uint64_t tb = (uint64_t)__double_as_longlong(it.t + delay_s[dl <= a.max_delay ? dl : 0]);
                while (bits) {
                    const int b = __builtin_ctzll(bits);
                    bits &= bits - 1ull;
                    const int d = j * 64 + b + 1;
                    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)tb, b);
                    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(tb >> 32), b);
                    const double td = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                    uint32_t ch = 0u;
                    if (any_full) ch = cur[d - 1];
                    push(eq_item(td, kEqPrioSpike, kEqGroup, e, (uint32_t)d, ch));
                }
            }
            if (lane == 0) next_k[e] = ke + 1;
            push(eq_item(it.t + E.h, kEqPrioTime, kEqTime, e, 0, 0));
        } else if (kind == kEqGroup) {
            give_rank((e << 16) | eq_d(it), it.b, 0u);
        } else if (kind == kEqInput) {
            // InputSpike::ProcessEvent: the neuron's first delay
            const int32_t nr = eq_g(E.in_neuron)[it.b];
            const int d0 = eq_g(E.nd_list)[eq_g(E.nd_ptr)[nr]];
            push(eq_item(it.t + delay_s[d0], kEqPrioSpike, kEqProp, e, 0, it.b));
        } else if (kind == kEqProp) {
            // PropagatedSpike::ProcessEvent: this delay, then the next one at
            // time - delay_j + delay_{j+1}
            const int32_t nr = eq_g(E.in_neuron)[it.b];
            const uint32_t j = eq_d(it);
            const int32_t np0 = eq_g(E.nd_ptr)[nr];
            const int d = eq_g(E.nd_list)[np0 + (int32_t)j];
            give_rank((e << 16) | (uint32_t)d, it.b, 1u);
            if (np0 + (int32_t)j + 1 < eq_g(E.nd_ptr)[nr + 1]) {
                const int d2 = eq_g(E.nd_list)[np0 + (int32_t)j + 1];
                push(eq_item(it.t - delay_s[d] + delay_s[d2], kEqPrioSpike, kEqProp, e, j + 1, it.b));
            }
        } else {
            w.unknown = true;
            break;
        }
    }
    flush();
}

// The replay of one window by one workgroup of kBlock threads: the spikes
// per delay are counted by all lanes, the heap operations run on the first wave.
// LDS_ITEMS heap entries are staged in LDS when the heap cannot outgrow
// them; MAXD bounds the delays (host-checked).
template <int LDS_ITEMS, int MAXD>
__device__ __forceinline__ void eq_body(const EqArgs &a) {
    __shared__ EqItem s_heap[LDS_ITEMS + 1];
    __shared__ uint32_t s_cnt[kEqMaxEmitters][MAXD];
    __shared__ uint32_t s_partial[kEqMaxEmitters];
    __shared__ uint32_t s_cur[MAXD];
    __shared__ int32_t s_next[kEqMaxEmitters];
    __shared__ EqEmitter s_em[kEqMaxEmitters];
    __shared__ int32_t s_scan[kBlock], s_scan2[kBlock];
    __shared__ int32_t s_run, s_base;
    __shared__ int32_t s_st[kEqSt];
    __shared__ int32_t s_in_lds;
    __shared__ double s_delay[MAXD + 1];
    __shared__ uint32_t s_pmask[kEqMaxEmitters][MAXD / 32 + 1];
    __shared__ uint32_t s_anyfull[kEqMaxEmitters];
    __shared__ uint4 s_rk[kEqRankList];
    const int32_t k = a.step.get();
    // every independent load at once: state, models, delays, and the first
    // tile counts of the population replayed from its tiles (the array is
    // padded to a multiple of 8 * kBlock entries)
    int32_t pf[8];
    if (a.pf_tiles) {
#pragma unroll
        for (int u = 0; u < 8; u++) pf[u] = eq_g(a.pf_tiles)[u * kBlock + (int32_t)threadIdx.x];
    }
    if (threadIdx.x < kEqSt) s_st[threadIdx.x] = a.st[threadIdx.x];
    if (threadIdx.x < a.n_em) {
        s_em[threadIdx.x] = a.em[threadIdx.x];
        s_next[threadIdx.x] = a.st[kEqSt + threadIdx.x];
    }
    for (int d = threadIdx.x; d <= a.max_delay; d += kBlock) s_delay[d] = a.delay_s[d];
    for (int i = threadIdx.x; i < kEqMaxEmitters * MAXD; i += kBlock) (&s_cnt[0][0])[i] = 0u;
    if (threadIdx.x < kEqMaxEmitters) s_partial[threadIdx.x] = 0u;
    __syncthreads();
    const int32_t size0 = s_st[0];
    // stage the heap in LDS speculatively (the check below may discard it)
    if (size0 + 1 <= LDS_ITEMS)
        for (int i = threadIdx.x; i <= size0; i += kBlock) s_heap[i] = a.heap[i];
    // 1. spikes per delay of every model whose update at k (or before, when
    // resuming) is still to leave the heap; s_partial: some spiking neuron
    // lacks a delay of its model's structure.
    for (int e = 0; e < a.n_em; e++) {
        const EqEmitter &E = s_em[e];
        if (E.kind == kEqTrain) continue;
        const int32_t ke = s_next[e];
        if (ke > k || ke < 1) continue;
        const int slot = ke % a.ring;
        if (E.all_full) {
            // every spike joins a group of every delay: only the spike count
            int32_t tot = 0;
            if (E.tile_count && ke == k) {
                for (int32_t t0 = 0; t0 < E.n_tiles; t0 += 8 * kBlock) {
                    const bool pre = t0 == 0 && e == a.pf_em && a.pf_tiles == E.tile_count;
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int32_t t = t0 + u * kBlock + (int32_t)threadIdx.x;
                        const int32_t c = pre ? pf[u] : eq_g(E.tile_count)[t];
                        tot += t < E.n_tiles ? c : 0;
                    }
                }
            } else if (threadIdx.x == 0) {
                tot = eq_g(E.ring_cnt)[slot];
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
            if ((threadIdx.x & 63u) == 0 && tot > 0)
                for (int w = 0; w < a.words; w++) {
                    uint32_t bits = E.full[w];
                    while (bits) {
                        const int b = __builtin_ctz(bits);
                        bits &= bits - 1u;
                        atomicAdd(&s_cnt[e][(w * 32 + b) % MAXD], (uint32_t)tot);
                    }
                }
            continue;
        }
        auto count = [&](int32_t id) {
            const EQ_GLOBAL uint32_t *m = eq_g(E.dmask) + (int64_t)id * a.words;
            bool part = false;
            for (int w = 0; w < a.words; w++) {
                uint32_t bits = m[w];
                part = part || bits != E.full[w];
                while (bits) {
                    const int b = __builtin_ctz(bits);
                    bits &= bits - 1u;
                    atomicAdd(&s_cnt[e][w * 32 + b], 1u);
                }
            }
            if (part) s_partial[e] = 1u;
        };
        if (E.tile_count && ke == k) {
            // tile counts eight at a time (the array is padded to a multiple
            // of 8 * kBlock entries), so their loads are in flight together
            for (int32_t t0 = 0; t0 < E.n_tiles; t0 += 8 * kBlock) {
                int32_t c8[8];
                const bool pre = t0 == 0 && e == a.pf_em && a.pf_tiles == E.tile_count;
#pragma unroll
                for (int u = 0; u < 8; u++)
                    c8[u] = pre ? pf[u] : eq_g(E.tile_count)[t0 + u * kBlock + (int32_t)threadIdx.x];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int32_t t = t0 + u * kBlock + (int32_t)threadIdx.x;
                    if (t >= E.n_tiles) continue;
                    for (int32_t j = 0; j < c8[u]; j++) count(eq_g(E.staging)[(int64_t)t * E.tile_size + j]);
                }
            }
        } else {
            const int32_t ns = eq_g(E.ring_cnt)[slot];
            for (int32_t r = threadIdx.x; r < ns; r += kBlock) count(eq_g(E.ring_idx)[(int64_t)slot * E.slot_len + r]);
        }
    }
    __syncthreads();
    // per model (one wave each): delays with a partial group, and whether
    // any group is full, from ballots over 64 delays at a time
    for (int e = (int)(threadIdx.x >> 6); e < kEqMaxEmitters; e += kBlock / 64) {
        const int lane = eq_lane();
        bool full = false;
#pragma unroll
        for (int j = 0; j < (MAXD + 63) / 64; j++) {
            const int d = j * 64 + lane + 1;
            const uint32_t c = d <= MAXD && d <= a.max_delay ? s_cnt[e][(d - 1) % MAXD] : 0u;
            const uint64_t pm = __ballot((c % (uint32_t)kEqChunk) != 0u);
            full = full || __ballot(c >= (uint32_t)kEqChunk) != 0ull;
            if (lane == 0) {
                if (2 * j < MAXD / 32 + 1) s_pmask[e][2 * j] = (uint32_t)pm;
                if (2 * j + 1 < MAXD / 32 + 1) s_pmask[e][2 * j + 1] = (uint32_t)(pm >> 32);
            }
        }
        if (lane == 0) s_anyfull[e] = full ? 1u : 0u;
    }
    // 2. last member of every full chunk where a delay has 1,024 spikes or
    // more and not every spiking neuron has every delay (else chunk c ends at
    // list entry 1024c + 1023): a block scan per such delay over the list
    // (ring list, or tiles in ascending order).  The entry past the last
    // full chunk holds INT32_MAX.
    for (int e = 0; e < a.n_em; e++) {
        const EqEmitter &E = s_em[e];
        if (E.kind == kEqTrain) continue;
        const int32_t ke = s_next[e];
        if (ke > k || ke < 1) continue;
        const int slot = ke % a.ring;
        if (threadIdx.x == 0) eq_g(E.bnd_on)[slot] = (uint8_t)s_partial[e];
        if (!s_partial[e]) continue;
        const bool tiles = E.tile_count && ke == k;
        const int32_t ns = tiles ? E.n_tiles : eq_g(E.ring_cnt)[slot];   // list entries, or tiles
        for (int d = 1; d <= a.max_delay; d++) {
            const uint32_t full_c = s_cnt[e][d - 1] / (uint32_t)kEqChunk;
            if (full_c == 0) continue;
            EQ_GLOBAL int32_t *bd = eq_g(E.bnd) + ((int64_t)slot * a.max_delay + (d - 1)) * E.cmax;
            if (threadIdx.x == 0) {
                s_run = 0;
                s_base = 0;
                if ((int)full_c < E.cmax) bd[full_c] = 0x7fffffff;
            }
            __syncthreads();
            for (int32_t r0 = 0; r0 < ns; r0 += kBlock) {
                const int32_t r = r0 + (int32_t)threadIdx.x;
                // members (f) and list entries (g) of this thread's entry or tile
                int32_t f = 0, g = 0;
                if (r < ns) {
                    if (tiles) {
                        g = eq_g(E.tile_count)[r];
                        for (int32_t j = 0; j < g; j++)
                            f += eq_has(E.dmask + (int64_t)eq_g(E.staging)[(int64_t)r * E.tile_size + j] * a.words, d) ? 1 : 0;
                    } else {
                        g = 1;
                        f = eq_has(E.dmask + (int64_t)eq_g(E.ring_idx)[(int64_t)slot * E.slot_len + r] * a.words, d) ? 1 : 0;
                    }
                }
                s_scan[threadIdx.x] = f;
                s_scan2[threadIdx.x] = g;
                __syncthreads();
                for (int off = 1; off < kBlock; off <<= 1) {
                    const int32_t v = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0;
                    const int32_t v2 = threadIdx.x >= (unsigned)off ? s_scan2[threadIdx.x - off] : 0;
                    __syncthreads();
                    s_scan[threadIdx.x] += v;
                    s_scan2[threadIdx.x] += v2;
                    __syncthreads();
                }
                int32_t upto = s_run + s_scan[threadIdx.x] - f;          // members before this entry or tile
                const int32_t pos0 = s_base + s_scan2[threadIdx.x] - g;  // its first list position
                if (r < ns) {
                    if (tiles) {
                        for (int32_t j = 0; j < g; j++) {
                            if (!eq_has(E.dmask + (int64_t)eq_g(E.staging)[(int64_t)r * E.tile_size + j] * a.words, d))
                                continue;
                            upto++;
                            if (upto % kEqChunk == 0 && upto / kEqChunk <= (int32_t)full_c) bd[upto / kEqChunk - 1] = pos0 + j;
                        }
                    } else if (f) {
                        upto++;
                        if (upto % kEqChunk == 0 && upto / kEqChunk <= (int32_t)full_c) bd[upto / kEqChunk - 1] = pos0;
                    }
                }
                __syncthreads();
                if (threadIdx.x == kBlock - 1) {
                    s_run += s_scan[kBlock - 1];
                    s_base += s_scan2[kBlock - 1];
                }
                __syncthreads();
            }
        }
    }
    // 3. the window's rank table is cleared first when the run stops inside
    // it (explicit stop time): groups still queued at the stop hold no rank
    if (threadIdx.x == 0 && a.resume) s_st[3] = a.clear_step;
    __syncthreads();
    if (!a.resume && s_st[3] == k) {
        const int par = (k + 1) & 1;
        for (int e = 0; e < a.n_em; e++) {
            const EqEmitter &E = s_em[e];
            const int64_t half = (int64_t)a.max_delay * E.cmax;
            for (int64_t i = threadIdx.x; i < half; i += kBlock) eq_g(E.rank)[par * half + i] = 0xffffffffu;
        }
    }
    // 4. the heap stays in LDS when it cannot outgrow it in this call
    if (threadIdx.x == 0) {
        int64_t bound = (int64_t)size0 + a.n_em + 2;
        for (int e = 0; e < a.n_em; e++) {
            const EqEmitter &E = s_em[e];
            if (E.kind == kEqTrain) {
                if (a.resume) bound += E.in_hi - E.in_lo;
                continue;
            }
            for (int d = 1; d <= a.max_delay; d++) bound += s_cnt[e][d - 1] / kEqChunk + 1;
        }
        s_in_lds = bound <= LDS_ITEMS ? 1 : 0;
    }
    __syncthreads();
    const bool in_lds = s_in_lds != 0;
    // 5. the heap operations, by the first wave
    if (threadIdx.x < 64) {
        EqWin w{size0, (uint32_t)s_st[1], -1, false, false};
        if (in_lds) eq_window<MAXD>(s_heap, a, s_em, s_cnt, s_pmask, s_anyfull, s_cur, s_next, s_delay, s_rk, k, w);
        else eq_window<MAXD>(a.heap, a, s_em, s_cnt, s_pmask, s_anyfull, s_cur, s_next, s_delay, s_rk, k, w);
        if (threadIdx.x == 0) {
            s_st[0] = w.n;
            a.st[0] = w.n;
            a.st[1] = (int32_t)w.r;
            if (a.resume || w.stop_r >= 0) a.st[2] = w.stop_r;
            if (a.resume) a.st[3] = a.clear_step;
            for (int e = 0; e < a.n_em; e++) a.st[kEqSt + e] = s_next[e];
            if (w.overflow) atomicOr(a.err, 64u);
            if (w.unknown) atomicOr(a.err, 128u);
        }
    }
    __syncthreads();
    if (in_lds)
        for (int i = threadIdx.x; i <= s_st[0]; i += kBlock) a.heap[i] = s_heap[i];
}

// The replay as a kernel of its own (circuits, plastic networks, resuming calls).
__global__ __launch_bounds__(kBlock) void k_eventq(EqArgs a) { eq_body<2048, kMaxDelayEq>(a); }

// Sizes of the replay fused into a network's compaction kernel.
constexpr int kEqFusedItems = 1024;
constexpr int kEqFusedDelays = 64;

}  // namespace edl

namespace edl {
// Delay bitmask of every neuron of a source: bit d-1 when its row of delay d
// is not empty (OR-ed into the emitter's masks, one source at a time).
__global__ __launch_bounds__(kBlock) void k_delay_mask(const int64_t *row_ptr, int64_t n, int max_delay, int words,
                                                       uint32_t *mask) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        for (int w = 0; w < words; w++) {
            uint32_t bits = 0;
            for (int b = 0; b < 32; b++) {
                const int d = w * 32 + b + 1;
                if (d > max_delay) break;
                if (row_ptr[i * max_delay + d] > row_ptr[i * max_delay + d - 1]) bits |= 1u << b;
            }
            mask[i * words + w] |= bits;
        }
    }
}
}  // namespace edl
