// edlut_amd.hip -- C ABI (include/edlut_amd.h) over the gfx950 kernels.
//
// Host-side objects own every device allocation (the reference builds device
// objects with device-side `new`, AdExTimeDrivenModel_GPU_C_INTERFACE.cu:
// 171-208; here they are plain host structs passed to kernels by value).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <limits>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/edlut_amd.h"
#include "adjoint.hpp"
#include "generate.hpp"
#include "kernels.hpp"

using namespace edl;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(EDLUT_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

template <typename T> int dalloc(T **p, size_t count) {
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void **)p, count * sizeof(T)));
    return 0;
}

uint64_t fnv1a(const void *data, size_t bytes) {
    const unsigned char *b = (const unsigned char *)data;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < bytes; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// ExponentialTable contents (ExponentialTable.h:60-66): double-precision exp
// of each float grid point, stored as float.
std::vector<float> build_exp_table() {
    std::vector<float> t(kExpSize);
    for (int i = 0; i < kExpSize; i++) {
        float exponent = kExpMin + ((kExpMax - kExpMin) * i) / (kExpSize - 1);
        t[i] = (float)std::exp((double)exponent);
    }
    return t;
}

// g_nmda_inf table (AdExTimeDrivenModel.cpp:27-35); exp(float) is the float
// overload in the reference's model translation units.
// k: -0.062 per mV, or -62 per V for the SI-unit LIF (LIFTimeDrivenModel_IS.cpp:25-33).
std::vector<float> build_nmda_table(float e_exc, float e_inh, float *aux, float k = -0.062f) {
    std::vector<float> t(kNmdaSize);
    *aux = (kNmdaSize - 1) / (e_exc - e_inh);
    for (int i = 0; i < kNmdaSize; i++) {
        float V = e_inh + ((e_exc - e_inh) * i) / (kNmdaSize - 1);
        t[i] = 1.0f / (1.0f + expf(k * V) * (1.2f / 3.57f));
    }
    return t;
}

// LIF refractory fractions (LIFTimeDrivenModel.cpp:189-201).  After a spike
// at update s the clock holds 0 and gains (double)dt_s per update, so the
// derivative of update s+1+j sees S_j = j-fold double sum of dt_s.  Entry j
// is the fraction for S_j (negative: derivative zero); the table stops at the
// first j with S_j*time_scale > tau_ref, after which the full derivative
// applies.  dt_ms is the step in the model's time scale (ms, or s for the
// SI-unit LIF, LIFTimeDrivenModel_IS.cpp:183-192, time_scale 1).
std::vector<float> build_refractory(float dt_s, float dt_ms, float tau_ref, float time_scale = 1000.0f) {
    std::vector<float> t;
    double S = 0.0;
    for (int j = 0; j < (1 << 24); j++) {
        double lst = S * time_scale;
        if (lst > tau_ref) break;
        if ((lst + dt_ms) > tau_ref) t.push_back((float)((lst + dt_ms - tau_ref) / dt_ms));
        else t.push_back(-1.0f);
        S += (double)dt_s;
    }
    return t;
}

inline float inv_of(float x) { return (float)(1.0 / (double)x); }

// HH channel-rate table (HHTimeDrivenModel.h:372-421): 262144 entries of six
// rates on [-90, 50] mV, evaluated in float with the float exp.
std::vector<float> build_hh_table(float v_thr) {
    std::vector<float> t((size_t)kHHTable * 6);
    for (int i = 0; i < kHHTable; i++) {
        const float V = kHHMinV + ((kHHMaxV - kHHMinV) * i) / (kHHTable - 1);
        float am, bm, an;
        if ((13.0f - V + v_thr) == 0.0) am = 1.279999894092334f;
        else am = 0.32f * (13.0f - V + v_thr) / (expf((13.0f - V + v_thr) * 0.25f) - 1.0f);
        if ((V - v_thr - 40.0f) == 0.0) bm = 1.400030971093304f;
        else bm = 0.28f * (V - v_thr - 40.0f) / (expf((V - v_thr - 40.0f) * 0.2f) - 1.0f);
        const float ah = 0.128f * expf((17.0f - V + v_thr) * 0.05555555f);
        const float bh = 4.0f / (1.0f + expf((40.0f - V + v_thr) * 0.2f));
        if ((15.0f - V + v_thr) == 0.0) an = 0.1599999867615418f;
        else an = 0.032f * (15.0f - V + v_thr) / (expf((15.0f - V + v_thr) * 0.2f) - 1.0f);
        const float bn = 0.5f * expf((10.0f - V + v_thr) * 0.025f);
        float *o = &t[(size_t)i * 6];
        o[0] = am; o[1] = bm; o[2] = ah; o[3] = bh; o[4] = an; o[5] = bn;
    }
    return t;
}

inline int hh_index(float V) {
    int pos;
    const float x = (V - kHHMinV) * kHHAux;
    if (!(x < 2147483648.0f) || x < -2147483648.0f) pos = (int)0x80000000u;
    else pos = (int)x;
    return pos < 0 ? 0 : (pos > kHHTable - 1 ? kHHTable - 1 : pos);
}

// Clamped truncating index of a 1024-entry channel table on [lo, hi].
inline int chan_index(float V, float lo, float aux) {
    const float x = (V - lo) * aux;
    int pos = (!(x < 2147483648.0f) || x < -2147483648.0f) ? (int)0x80000000u : (int)x;
    return pos < 0 ? 0 : (pos > 1023 ? 1023 : pos);
}

// x/(exp(x/y)-1), its limit y(1-x/y/2) near x = 0
// (EgidioGranuleCell_TimeDriven.h:170-178).
inline float gr_linoid(float x, float y) {
    if (fabsf(x / y) < 1e-06f) return y * (1 - x / y / 2);
    return x / (expf(x / y) - 1);
}

// Temperature factors 3^((30 - T0)/10) of the granule model (.cpp:54-57).
inline float gr_q10(float t0) { return powf(3.0f, (30.0f - t0) / 10.0f); }

// Granule channel table (EgidioGranuleCell_TimeDriven.h:402-552): per grid
// point, (steady state, 1/tau) of every voltage-gated variable, and the two
// voltage factors of the Ca-dependent K gate, all in float with the float exp.
std::vector<float> build_granule_table() {
    const float q20 = gr_q10(20.0f), q22 = gr_q10(22.0f), q30 = gr_q10(30.0f), q63 = gr_q10(6.3f);
    std::vector<float> t((size_t)kGrTable * kGrVars);
    for (int i = 0; i < kGrTable; i++) {
        const float V = kGrMinV + ((kGrMaxV - kGrMinV) * i) / (kGrTable - 1);
        float *o = &t[(size_t)i * kGrVars];
        auto pair = [&](int k, float a, float b) {   // x_inf = a/(a+b), 1/tau = a+b
            o[k] = a / (a + b);
            o[k + 1] = a + b;
        };
        pair(0, q20 * (-0.3f) * gr_linoid(V + 19.0f, -10.0f), q20 * 12.0f * expf(-(V + 44.0f) / 18.182f));
        pair(2, q20 * 0.105f * expf(-(V + 44.0f) / 3.333f), q20 * 1.5f / (1.0f + expf(-(V + 11.0f) / 5.0f)));
        pair(4, q20 * (0.00008f - 0.00493f * gr_linoid(V - 4.48754f, -6.81881f)),
             q20 * (0.04752f + 0.01558f * gr_linoid(V + 43.97494f, 0.10818f)));
        pair(6, q20 * 0.31836f * expf(-(V + 80.0f) / 62.52621f), q20 * 0.01014f * expf((V + 83.3332f) / 16.05379f));
        o[8] = 1.0f / (1.0f + expf(-(V + 42.0f) / 5.0f));
        o[9] = (q30 * (-0.091f) * gr_linoid(V + 42.0f, -5.0f) + q30 * 0.062f * gr_linoid(V + 42.0f, 5.0f)) * 0.2f;
        pair(10, q63 * (-0.01f) * gr_linoid(V + 25.0f, -10.0f), q63 * 0.125f * expf(-0.0125f * (V + 35.0f)));
        o[12] = 1.0f / (1.0f + expf((V - (-46.7f)) / (-19.8f)));
        o[13] = (q20 * 4.88826f) / (1 + expf(-(V + 9.17203f) / 23.32708f)) +
                (q20 * 0.99285f) / expf((V + 18.27914f) / 19.47175f);
        o[14] = 1.0f / (1.0f + expf((V - (-78.8f)) / 8.4f));
        o[15] = (q20 * 0.11042f) / (1.0f + expf((V + 111.33209f) / 12.8433f)) +
                (q20 * 0.10353f) / (1.0f + expf(-(V + 49.9537f) / 8.90123f));
        pair(16, q20 * 0.13289f * expf(-(V + 83.94f) / 24.3902f), q20 * 0.16994f * expf((V + 83.94f) / 35.714f));
        o[18] = 0.0015f * expf(-V / 11.765f);
        o[19] = 1.0f / (0.00015f * expf(-V / 11.765f));
        pair(20, q20 * 0.04944f * expf((V + 29.06f) / 15.87301587302f), q20 * 0.08298f * expf(-(V + 18.66f) / 25.641f));
        pair(22, q20 * 0.0013f * expf(-(V + 48.0f) / 18.183f), q20 * 0.0013f * expf((V + 48.0f) / 83.33f));
        o[24] = 1.0f / (1.0f + expf(-(V - (-30.0f)) / 6.0f));
        o[25] = q22 * 0.0033f * expf((V + 30.0f) / 40.0f) + q22 * 0.0033f * (-(V + 30.0f) / 20.0f);
    }
    return t;
}

// Purkinje channel table (TimeDrivenPurkinjeCell.h:396-421).
std::vector<float> build_purkinje_table() {
    std::vector<float> t((size_t)kPkTable * 4);
    for (int i = 0; i < kPkTable; i++) {
        const float V = kPkMinV + ((kPkMaxV - kPkMinV) * i) / (kPkTable - 1);
        const float alpha_ca = 1.6f / (1 + expf(-0.072f * (V - 5.0f)));
        const float beta_ca = (0.02f * (V + 8.9f)) / (expf((V + 8.9f) * 0.2f) - 1.0f);
        const float alpha_m = 0.3f / (1 + expf((-V - 2.0f) * 0.2f));
        const float beta_m = 0.001f * expf((-V - 60.0f) * 0.055555555555555f);
        float *o = &t[(size_t)i * 4];
        o[0] = alpha_ca;
        o[1] = alpha_ca + beta_ca;
        o[2] = alpha_m;
        o[3] = alpha_m + beta_m;
    }
    return t;
}

// Purkinje refractory clock (TimeDrivenPurkinjeCell.cpp:172-224).  After a
// spike the clock S holds 0 and gains (double)dt_s per update; entry j holds,
// for S_j, the derivative mode the right-hand side takes (from the float
// clock in ms: 0 full, 1 rising flank, 2 falling flank) and the membrane
// potential the spike condition imposes (from the double clock; NaN: none).
// The table stops where neither applies any more.
std::vector<float> build_purkinje_clock(float dt_s, float dt_ms, float v_thr, float e_leak, float tau_ref) {
    const float spk_peak = 31.0f, time_scale = 1000.0f;
    const float tau_ref_0_5 = tau_ref / 2;
    const float inv_half = (float)(1.0 / (double)tau_ref_0_5);
    std::vector<float> t;
    double S = 0.0;
    for (int j = 0; j < (1 << 24); j++) {
        const float ls_f = (float)(time_scale * S);
        const float mode = ls_f >= tau_ref ? 0.0f : (ls_f <= tau_ref_0_5 ? 1.0f : 2.0f);
        const double ls = S * time_scale;
        float v = std::numeric_limits<float>::quiet_NaN();
        if (ls < tau_ref) {
            if (ls <= tau_ref_0_5) v = (float)(v_thr + (spk_peak - v_thr) * (ls * inv_half));
            else v = (float)(spk_peak - (spk_peak - e_leak) * ((ls - tau_ref_0_5) * inv_half));
        } else if ((ls - tau_ref) < dt_ms) {
            v = e_leak;
        }
        if (mode == 0.0f && v != v) break;
        t.push_back(mode);
        t.push_back(v);
        S += (double)dt_s;
    }
    return t;
}

// BDF coefficients (BDFn.h:604-669), row = method state (0..6)
void bdf_coefficients(float *C) {
    std::memset(C, 0, sizeof(float) * 49);
    auto set = [&](int r, std::initializer_list<float> v) {
        int c = 0;
        for (float x : v) C[r * 7 + c++] = x;
    };
    set(0, {1.0f, 1.0f});
    set(1, {1.0f, 1.0f});
    set(2, {2.0f / 3.0f, 4.0f / 3.0f, -1.0f / 3.0f});
    set(3, {6.0f / 11.0f, 18.0f / 11.0f, -9.0f / 11.0f, 2.0f / 11.0f});
    set(4, {12.0f / 25.0f, 48.0f / 25.0f, -36.0f / 25.0f, 16.0f / 25.0f, -3.0f / 25.0f});
    set(5, {60.0f / 137.0f, 300.0f / 137.0f, -300.0f / 137.0f, 200.0f / 137.0f, -75.0f / 137.0f,
            12.0f / 137.0f});
    set(6, {60.0f / 147.0f, 360.0f / 147.0f, -450.0f / 147.0f, 400.0f / 147.0f, -225.0f / 147.0f,
            72.0f / 147.0f, -10.0f / 147.0f});
}

// tile-count arrays are padded (and zeroed) to whole compaction chunks so the
// compaction kernel reads them without bounds tests
inline size_t padded_tiles(int n_tiles) {
    const size_t chunk = 8 * (size_t)kBlock;
    return ((size_t)n_tiles + chunk - 1) / chunk * chunk;
}

}  // namespace

// ---------------------------------------------------------------------------
struct edlut_population {
    int model = 0, integ = 0, bdf_order = 0;
    int n_sub = 1;                   // Bifixed methods: sub-steps per update
    int n = 0, nv = 0, nd = 0;
    int64_t stride = 0;
    int n_tiles = 0;
    int device = 0;
    uint32_t mask = 0;
    float dt_s = 0.0f, dt_ms = 0.0f;
    std::vector<float> params;
    ModelParams mp{};
    StepConsts k{};
    hipStream_t stream = nullptr;

    float *state = nullptr;
    int32_t *last_spike = nullptr;
    float *chan = nullptr;           // HH channel-rate table
    float *bdf_c = nullptr;          // BDF coefficients [7][7]
    void *bdf_d = nullptr;           // BDF difference table
    void *bdf_prev = nullptr;        // BDF previous states
    int32_t *bdf_st = nullptr;       // BDF method state per neuron
    bool newton_fp64 = false;
    bool started = false;            // an update has run (BDF precision is fixed then)
    float *exp_lut = nullptr;
    float *nmda_lut = nullptr;
    float *refr = nullptr;
    int refr_len = 0;
    float *bdf_ws = nullptr;          // granule BDF: fallback working storage
    uint32_t *bdf_ws_count = nullptr; // [2] slots handed out, by step parity
    uint32_t *bdf_bits = nullptr;     // granule BDF: spike bits of a launch [stride / 32]
    uint32_t *bdf_done = nullptr;     // granule BDF: finished workgroups of a launch
    int n_cu = 0;                     // compute units of the device
    uint64_t lut_sums[3] = {0, 0, 0};

    // spikes of the last update (tile counts + staging, then compacted)
    int32_t *tile_count = nullptr;
    int32_t *staging = nullptr;
    int32_t *spk_idx = nullptr;
    int32_t *spk_cnt = nullptr;
    int32_t *h_spk = nullptr;   // pinned
    int32_t *h_cnt = nullptr;   // pinned
    int last_n_spikes = 0;
    uint32_t *err = nullptr;

    // ordered input fold
    uint32_t *cnt = nullptr, *fill = nullptr, *tile_flag = nullptr;
    uint64_t *seg = nullptr, *bump = nullptr;
    uint64_t upd_seq = 0;   // plug-in mode: parity of the fold counters
    FoldEntry *b_ent = nullptr;
    uint64_t b_cap = 0;
    std::vector<int32_t> q_tgt;
    std::vector<uint8_t> q_type;
    std::vector<float> q_w;
    int32_t *e_tgt = nullptr;
    uint8_t *e_type = nullptr;
    float *e_w = nullptr;
    int e_cap = 0;

    edlut_network *net = nullptr;

    // grouped populations: per-group parameters on the device
    bool grouped = false;
    std::vector<int32_t> g_start, g_size;
    GroupInfo *d_groups = nullptr;
    uint32_t *d_chunk_group = nullptr;
    std::vector<float> init_state;   // initial SoA state (host copy)
    int items = kItems;              // neurons per thread of the update launch

    int update_tiles() const { return (int)(stride / ((int64_t)kBlock * items)); }
    // workgroups of the update launch: one per tile, except the granule's
    // BDF kernel, whose waves take neurons from equal ranges (at most one
    // workgroup per CU fits)
    int update_blocks() const {
        const int t = update_tiles();
        return (model == EDLUT_GRANULE && integ == EDLUT_BDF && n_cu > 0 && n_cu < t) ? n_cu : t;
    }

    FoldTarget fold_target(int par, bool by_step = false) const {
        FoldTarget t;
        t.seg_off = nullptr;
        t.rtab = nullptr;
        t.par_by_step = by_step ? 1 : 0;
        t.state = state;
        t.stride = stride;
        t.ndiff = nd;
        t.cnt = cnt;
        t.fill = fill;
        t.seg = seg;
        t.tile_flag = tile_flag;
        t.bump = bump;
        t.par = par;
        t.b_ent = b_ent;
        t.b_cap = b_cap;
        t.err = err;
        return t;
    }

    // BDF history: bdf_order difference rows and bdf_order-1 previous states
    // of the nd differential variables, in `elem`-byte words.
    int alloc_bdf_history(size_t elem) {
        if (bdf_d) (void)hipFree(bdf_d);
        if (bdf_prev) (void)hipFree(bdf_prev);
        bdf_d = bdf_prev = nullptr;
        const size_t nd_rows = (size_t)bdf_order * nd, np_rows = (size_t)(bdf_order > 1 ? bdf_order - 1 : 1) * nd;
        HIP_TRY(hipMalloc(&bdf_d, nd_rows * stride * elem));
        HIP_TRY(hipMalloc(&bdf_prev, np_rows * stride * elem));
        HIP_TRY(hipMemset(bdf_d, 0, nd_rows * stride * elem));
        HIP_TRY(hipMemset(bdf_prev, 0, np_rows * stride * elem));
        return 0;
    }

    // Grow the fold's event buffer to hold `events` events.
    int reserve_events(uint64_t events) {
        if (events <= b_cap) return 0;
        (void)hipStreamSynchronize(stream);
        if (b_ent) (void)hipFree(b_ent);
        b_ent = nullptr;
        b_cap = 0;
        int rc;
        if ((rc = dalloc(&b_ent, events))) return rc;
        b_cap = events;
        return 0;
    }
};

struct edlut_network;
struct edlut_circuit;

struct Source {
    bool self = false;
    // projection from another network's population (edlut_network_add_projection):
    // its spikes are that network's own ring
    edlut_network *from = nullptr;
    float *rates = nullptr;       // per-generator Poisson frequencies (or null)
    // spike train (edlut_network_add_spike_train): spikes of update k are
    // train_idx[train_ptr[k-1] .. train_ptr[k])
    bool train = false;
    // optional spike log of a Poisson device (edlut_network_set_source_log),
    // counted like the network's own log
    int32_t *log_step = nullptr;
    int32_t *log_idx = nullptr;
    int64_t *log_count = nullptr;
    int64_t log_cap = 0;
    int64_t *train_ptr = nullptr;
    int32_t *train_idx = nullptr;
    int64_t train_steps = 0;
    int train_blocks = 1;
    int64_t n = 0;
    int32_t *ring_idx = nullptr;
    int32_t *ring_cnt = nullptr;
    int64_t *row_ptr = nullptr;
    uint32_t *tgt_type = nullptr;
    float *w = nullptr;
    uint32_t *lcg = nullptr;
    std::vector<uint32_t> lcg0;   // initial generator states (reset)
    float rate = 0.0f;
    // STDP (edlut_network_set_stdp)
    bool stdp = false;
    StdpSyn syn{};
    std::vector<float> w0;        // weights when STDP was switched on (reset)
    int64_t *in_ptr = nullptr;
    int64_t *in_syn = nullptr;
    int64_t in_max = 0;           // longest per-target list
    int64_t nsyn = 0;
    int32_t *tile_count = nullptr;
    int32_t *staging = nullptr;
    // event-queue order: the model this source belongs to (sources of several
    // networks with the same key are one model) and its type rank; a spike
    // train's input spikes as records in load order (time, neuron, position
    // in its update's list, load sequence number)
    int eq_key = -1, eq_rank = -1;
    std::vector<double> tr_t;
    std::vector<int32_t> tr_neuron, tr_pos;
    std::vector<int64_t> tr_seq;
    std::vector<int64_t> tr_load;   // 0: queued when the simulation starts; k: by the run starting at update k
    int32_t tr_longest = 0;
    std::vector<int64_t> h_train_ptr;   // host copy of the schedule
    std::vector<int32_t> h_train_idx;
};

struct EqSim;

struct edlut_network {
    edlut_population *pop = nullptr;
    edlut_circuit *circuit = nullptr;   // lockstep group this network belongs to
    int links_out = 0;                  // projections of this population into other networks
    int max_delay = 1;
    int ring = 3;
    std::vector<Source> sources;
    int self_src = -1;
    uint64_t total_syn = 0;
    bool has_ext = false;
    ExtDrive ext{};
    // population-side ring when there is no recurrent source
    int32_t *own_ring_idx = nullptr;
    int32_t *own_ring_cnt = nullptr;
    int32_t *tile_count = nullptr;
    int32_t *staging = nullptr;
    // spike log
    int32_t *log_step = nullptr;
    int32_t *log_idx = nullptr;
    int64_t *log_count = nullptr;   // [2], by step parity
    int64_t log_cap = 0;
    int64_t last_step = 0;
    // timing of the update kernel (event pool reserved by set_timing)
    int64_t timed = 0;
    // graph replay: a chunk of `chunk` captured steps reading the update
    // index from d_step; rebuilt when the network's configuration changes
    int32_t *d_step = nullptr;
    hipGraphExec_t graph = nullptr;
    void (*graph_fn)(UpdateArgs) = nullptr;
    int graph_chunk = 0;
    bool graph_fused = false;        // captured steps fold the next step's events
    // timed runs: per-workgroup [start, end] wall-clock stamps of the update
    // kernels of steps [clk_first, clk_first + clk_count)
    unsigned long long *clk = nullptr;
    int64_t clk_cap = 0, clk_tiles_max = 0, clk_first = 0, clk_count = 0, clk_blocks = 0;
    unsigned long long *clk_next = nullptr;   // per timed step: the next kernel's first start
    int64_t fold_ready = -1;         // update whose events are already folded
    bool dirty = true;
    bool use_graph = true;
    // launch shape of the ring fold (set by hand or by the autotuner)
    int fold_lanes = 32;
    int fold_blocks = 512;
    // fixed fold segment of every (type, target) key: prefix of the keys'
    // synapse counts over all sources (rebuilt when sources change)
    uint64_t *seg_off = nullptr;
    bool seg_dirty = true;
    // STDP: update time of each ring slot and delays in seconds
    bool plastic = false;
    double *d_time = nullptr;        // [ring]
    double *d_delay_s = nullptr;     // [max_delay + 1]
    int64_t t_cache_k = 0;           // host copy of t_k for k = t_cache_k
    double t_cache = 0.0;
    // delivery in the order of the reference's event queue (eventq.hpp):
    // the replayed queue of a network that runs alone, the emitter of each
    // source and its rank tables (set when the queue is built)
    bool eq_on = true;
    bool eq_stale = true;            // sources changed since the queue was built
    EqSim *eq = nullptr;
    int model_rank = -1;
    double stop_next = -1.0;
    std::vector<int> src_em;
    const uint32_t *eq_rank[kMaxSources] = {};
    int32_t eq_cmax[kMaxSources] = {};
    RankTab *d_rtab = nullptr;       // the rank tables of the sources, in device memory
    double *d_time_p = nullptr;      // [ring] update times of the Poisson devices' clock
    uint32_t *d_skip = nullptr;      // [2]: update whose fold skips the groups ranked below [1] (delivered at a stop)
    const int32_t *eq_bnd[kMaxSources] = {};
    const uint8_t *eq_bnd_on[kMaxSources] = {};
    double time_of(int64_t k) {
        // t_k = t_{k-1} + double(float dt), t_0 = 0 (TimeEventUpdateNeurons.cpp:78)
        if (k < t_cache_k) {
            t_cache_k = 0;
            t_cache = 0.0;
        }
        for (; t_cache_k < k; t_cache_k++) t_cache += (double)pop->dt_s;
        return t_cache;
    }
};

// ---------------------------------------------------------------------------
// The replayed event queue (eventq.hpp) of a network that runs alone or of a
// circuit: one emitter per population, Poisson model and spike-train model.
struct EqHostEm {
    int kind = kEqPop, rank = 0, key = -1;
    edlut_network *net = nullptr;   // a population's network (its update tiles)
    std::vector<std::pair<edlut_network *, int>> copies;   // (network, source) it feeds
    EqEmitter dev{};
    uint32_t *dmask = nullptr;
    uint32_t *rank_tab = nullptr;
    int32_t *bnd = nullptr;
    uint8_t *bnd_on = nullptr;
    double *in_t = nullptr;
    int32_t *in_neuron = nullptr, *in_pos = nullptr, *nd_ptr = nullptr;
    uint16_t *nd_list = nullptr;
    int64_t *in_seq = nullptr;
    int32_t in_loaded = 0;          // records already in the queue
    std::vector<int32_t> h_nd_ptr;  // trains: delays per neuron (host copy of nd_ptr)
};

struct EqSim {
    std::vector<EqHostEm> em;
    EqEmitter *d_em = nullptr;
    int max_delay = 1, ring = 3, words = 1;
    const double *delay_s = nullptr;
    EqItem *heap = nullptr;
    int32_t *st = nullptr;
    int32_t cap = 0;
    uint32_t *err = nullptr;
    int64_t next_first = 1;
    double h_pop = 0.0;
    int64_t t_k = 0;                // t_cache holds the time of update t_k
    double t_cache = 0.0;
    double time_of(int64_t k) {
        if (k < t_k) {
            t_k = 0;
            t_cache = 0.0;
        }
        for (; t_k < k; t_k++) t_cache += h_pop;
        return t_cache;
    }
    void free_all() {
        for (EqHostEm &E : em) {
            void *d[] = {E.dmask, E.rank_tab, E.bnd, E.bnd_on, E.in_t, E.in_neuron, E.in_pos, E.nd_ptr, E.nd_list,
                         E.in_seq};
            for (void *q : d)
                if (q) (void)hipFree(q);
        }
        em.clear();
        void *d[] = {d_em, heap, st};
        for (void *q : d)
            if (q) (void)hipFree(q);
        d_em = nullptr;
        heap = nullptr;
        st = nullptr;
    }
    ~EqSim() { free_all(); }
};

namespace {

// PoissonGeneratorDeviceVector samples every 0.1f ms of its millisecond time
// scale: 0.1f / 1000.0f in float (PoissonGeneratorDeviceVector.cpp:26,137).
inline double poisson_step_s() { return (double)(0.1f / 1000.0f); }

int64_t g_train_seq = 0;   // load order of input spikes across spike trains

// Spike-train records in load order from a step-major schedule: an input
// spike at the time of its update unless `time` gives it (aligned with idx),
// loaded in step-major order unless `order` ranks the entries.
void train_records(Source &s, double dt, int64_t n_steps, const int64_t *step_ptr, const int32_t *idx,
                   const double *time, const int64_t *order, const int64_t *load = nullptr) {
    const int64_t total = n_steps > 0 ? step_ptr[n_steps] : 0;
    std::vector<double> t((size_t)total);
    std::vector<int32_t> nr((size_t)total), pos((size_t)total);
    double tk = 0.0;
    int32_t longest = 0;
    for (int64_t k = 1; k <= n_steps; k++) {
        tk += dt;
        for (int64_t j = step_ptr[k - 1]; j < step_ptr[k]; j++) {
            t[(size_t)j] = time ? time[j] : tk;
            nr[(size_t)j] = idx[j];
            pos[(size_t)j] = (int32_t)(j - step_ptr[k - 1]);
        }
        longest = std::max<int32_t>(longest, (int32_t)(step_ptr[k] - step_ptr[k - 1]));
    }
    std::vector<int64_t> perm((size_t)total);
    for (int64_t j = 0; j < total; j++) perm[(size_t)j] = j;
    if (order) std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return order[a] < order[b]; });
    const size_t kept = std::min(s.tr_t.size(), (size_t)total);   // records queued before stay first
    s.tr_t.resize((size_t)total);
    s.tr_neuron.resize((size_t)total);
    s.tr_pos.resize((size_t)total);
    s.tr_seq.resize((size_t)total);
    s.tr_load.resize((size_t)total);
    for (size_t q = kept; q < (size_t)total; q++) {
        const int64_t j = perm[q];
        s.tr_t[q] = t[(size_t)j];
        s.tr_neuron[q] = nr[(size_t)j];
        s.tr_pos[q] = pos[(size_t)j];
        s.tr_seq[q] = order ? order[j] : g_train_seq++;
        s.tr_load[q] = load ? load[j] : 0;
    }
    // entries kept from an earlier schedule keep their records, but their
    // positions follow the new lists
    for (size_t q = 0; q < kept; q++) {
        for (int64_t j = 0; j < total; j++)
            if (nr[(size_t)j] == s.tr_neuron[q] && t[(size_t)j] == s.tr_t[q]) {
                s.tr_pos[q] = pos[(size_t)j];
                break;
            }
    }
    s.tr_longest = longest;
}

// Builds the replayed queue of `nets` (a network alone, or a circuit's
// members in order): emitters, delay masks, rank tables, the initial heap
// (queued input spikes, then one update event per model in type order,
// Simulation::InitSimulation, Simulation.cpp:228-268).
int eq_build(EqSim *S, const std::vector<edlut_network *> &nets, uint32_t *err, hipStream_t st) {
    S->free_all();
    const int m = (int)nets.size();
    const int D = nets[0]->max_delay;
    S->max_delay = D;
    S->ring = nets[0]->ring;
    S->words = (D + 31) / 32;
    S->delay_s = nets[0]->d_delay_s;
    S->err = err;
    S->h_pop = (double)nets[0]->pop->dt_s;
    S->t_k = 0;
    S->t_cache = 0.0;
    if (D > kMaxDelayEq) return fail(EDLUT_E_ARG, "the event-queue order supports delays up to %d steps", kMaxDelayEq);
    for (int i = 0; i < m; i++) {
        edlut_network *n = nets[i];
        if (n->pop->grouped) return fail(EDLUT_E_STATE, "grouped populations run in the canonical order");
        if (n->max_delay != D || n->pop->dt_s != nets[0]->pop->dt_s)
            return fail(EDLUT_E_ARG, "networks of one event queue share max_delay and step");
        EqHostEm E;
        E.kind = kEqPop;
        E.rank = n->model_rank >= 0 ? n->model_rank : i;
        E.dev.kind = kEqPop;
        E.dev.n = n->pop->n;
        E.dev.h = (double)n->pop->dt_s;
        E.dev.ring_idx = n->own_ring_idx;
        E.dev.ring_cnt = n->own_ring_cnt;
        E.dev.slot_len = n->pop->n;
        E.net = n;
        S->em.push_back(E);
    }
    int next_key = 1 << 24;
    for (int i = 0; i < m; i++) {
        edlut_network *n = nets[i];
        n->src_em.assign(n->sources.size(), -1);
        for (size_t si = 0; si < n->sources.size(); si++) {
            Source &s = n->sources[si];
            int e = -1;
            if (s.self) e = i;
            else if (s.from) {
                for (int j = 0; j < m; j++)
                    if (nets[j] == s.from) e = j;
                if (e < 0) return fail(EDLUT_E_STATE, "a projection's source network is not in the circuit");
            } else {
                const int key = s.eq_key >= 0 ? s.eq_key : next_key++;
                for (size_t q = (size_t)m; q < S->em.size(); q++)
                    if (S->em[q].key == key) e = (int)q;
                if (e >= 0) {
                    EqHostEm &E0 = S->em[(size_t)e];
                    if ((E0.kind == kEqTrain) != s.train || E0.dev.n != s.n)
                        return fail(EDLUT_E_ARG, "sources of one model must be alike");
                } else {
                    if ((int)S->em.size() >= kEqMaxEmitters) return fail(EDLUT_E_CAP, "at most %d models", kEqMaxEmitters);
                    EqHostEm E;
                    E.kind = s.train ? kEqTrain : kEqPoisson;
                    E.key = key;
                    E.rank = s.eq_rank >= 0 ? s.eq_rank : 1000 + (int)S->em.size();
                    E.dev.kind = E.kind;
                    E.dev.n = (int)s.n;
                    E.dev.h = s.train ? 0.0 : poisson_step_s();
                    E.dev.ring_idx = s.ring_idx;
                    E.dev.ring_cnt = s.ring_cnt;
                    E.dev.slot_len = s.n;
                    S->em.push_back(E);
                    e = (int)S->em.size() - 1;
                }
            }
            if ((int)S->em.size() > kEqMaxEmitters) return fail(EDLUT_E_CAP, "at most %d models", kEqMaxEmitters);
            S->em[(size_t)e].copies.push_back({n, (int)si});
            n->src_em[si] = e;
        }
    }
    const int W = S->words;
    int rc;
    for (EqHostEm &E : S->em) {
        const int64_t nn = E.dev.n;
        if ((rc = dalloc(&E.dmask, (size_t)nn * W))) return rc;
        HIP_TRY(hipMemsetAsync(E.dmask, 0, sizeof(uint32_t) * nn * W, st));
        for (auto &cp : E.copies) {
            const Source &s = cp.first->sources[(size_t)cp.second];
            const int nb = (int)std::min<int64_t>(4096, (nn + kBlock - 1) / kBlock);
            hipLaunchKernelGGL(k_delay_mask, dim3(nb), dim3(kBlock), 0, st, s.row_ptr, nn, D, W, E.dmask);
        }
        HIP_TRY(hipGetLastError());
        std::vector<uint32_t> h((size_t)nn * W);
        HIP_TRY(hipMemcpyAsync(h.data(), E.dmask, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int w = 0; w < kEqMaskWords; w++) E.dev.full[w] = 0u;
        for (int64_t i = 0; i < nn; i++)
            for (int w = 0; w < W; w++) E.dev.full[w] |= h[(size_t)i * W + w];
        E.dev.all_full = 1;
        for (int64_t i = 0; i < nn && E.dev.all_full; i++)
            for (int w = 0; w < W; w++)
                if (h[(size_t)i * W + w] != E.dev.full[w]) E.dev.all_full = 0;
        E.dev.dmask = E.dmask;
        if (E.kind == kEqTrain) {
            const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
            std::vector<int32_t> ptr((size_t)nn + 1, 0);
            std::vector<uint16_t> lst;
            for (int64_t i = 0; i < nn; i++) {
                for (int d = 1; d <= D; d++)
                    if ((h[(size_t)i * W + (d - 1) / 32] >> ((d - 1) % 32)) & 1u) lst.push_back((uint16_t)d);
                ptr[(size_t)i + 1] = (int32_t)lst.size();
            }
            if (lst.empty()) lst.push_back(0);
            E.h_nd_ptr = ptr;
            const size_t nr = std::max<size_t>(s0.tr_t.size(), 1);
            if ((rc = dalloc(&E.nd_ptr, ptr.size())) || (rc = dalloc(&E.nd_list, lst.size())) ||
                (rc = dalloc(&E.in_t, nr)) || (rc = dalloc(&E.in_neuron, nr)) || (rc = dalloc(&E.in_pos, nr)) ||
                (rc = dalloc(&E.in_seq, nr)))
                return rc;
            HIP_TRY(hipMemcpy(E.nd_ptr, ptr.data(), sizeof(int32_t) * ptr.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E.nd_list, lst.data(), sizeof(uint16_t) * lst.size(), hipMemcpyHostToDevice));
            if (!s0.tr_t.empty()) {
                HIP_TRY(hipMemcpy(E.in_t, s0.tr_t.data(), sizeof(double) * s0.tr_t.size(), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(E.in_neuron, s0.tr_neuron.data(), sizeof(int32_t) * s0.tr_t.size(),
                                  hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(E.in_pos, s0.tr_pos.data(), sizeof(int32_t) * s0.tr_t.size(), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(E.in_seq, s0.tr_seq.data(), sizeof(int64_t) * s0.tr_t.size(), hipMemcpyHostToDevice));
            }
            E.dev.in_t = E.in_t;
            E.dev.in_neuron = E.in_neuron;
            E.dev.in_pos = E.in_pos;
            E.dev.in_seq = E.in_seq;
            E.dev.nd_ptr = E.nd_ptr;
            E.dev.nd_list = E.nd_list;
            E.dev.cmax = std::max<int32_t>(1, s0.tr_longest);
        } else {
            E.dev.cmax = (int)std::max<int64_t>(1, (nn + kEqChunk - 1) / kEqChunk);
        }
        if (E.dev.cmax >= (1 << kOrdChunkBits)) return fail(EDLUT_E_CAP, "too many spike groups per update");
        const size_t nrank = (size_t)2 * D * E.dev.cmax, nbnd = (size_t)S->ring * D * E.dev.cmax;
        if ((rc = dalloc(&E.rank_tab, nrank)) || (rc = dalloc(&E.bnd, nbnd)) || (rc = dalloc(&E.bnd_on, (size_t)S->ring)))
            return rc;
        HIP_TRY(hipMemsetAsync(E.rank_tab, 0, sizeof(uint32_t) * nrank, st));
        HIP_TRY(hipMemsetAsync(E.bnd_on, 0, (size_t)S->ring, st));
        E.dev.rank = E.rank_tab;
        E.dev.bnd = E.bnd;
        E.dev.bnd_on = E.bnd_on;
    }
    // rank tables and chunk boundaries of every source's model
    for (edlut_network *n : nets) {
        RankTab rt{};
        rt.delays = D;
        for (size_t si = 0; si < n->sources.size(); si++) {
            const EqHostEm &E = S->em[(size_t)n->src_em[si]];
            n->eq_rank[si] = E.rank_tab;
            n->eq_cmax[si] = E.dev.cmax;
            n->eq_bnd[si] = E.bnd;
            n->eq_bnd_on[si] = E.bnd_on;
            rt.rank[si] = E.rank_tab;
            rt.cmax[si] = E.dev.cmax;
        }
        HIP_TRY(hipMemcpy(n->d_rtab, &rt, sizeof(RankTab), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(n->d_skip, 0, 2 * sizeof(uint32_t)));
    }
    // heap capacity: the groups of every model for every update in flight,
    // the input spikes, the update and stop events
    int64_t cap = 256;
    for (EqHostEm &E : S->em) {
        cap += (int64_t)D * (S->ring + 2) * (E.kind == kEqTrain ? 1 : E.dev.cmax + 1);
        if (E.kind == kEqTrain) {
            const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
            cap += (int64_t)s0.tr_t.size();
        }
    }
    if (cap > (1ll << 28)) return fail(EDLUT_E_CAP, "event queue too large");
    S->cap = (int32_t)cap;
    if ((rc = dalloc(&S->heap, (size_t)cap + 1)) || (rc = dalloc(&S->st, kEqSt + S->em.size())) ||
        (rc = dalloc(&S->d_em, S->em.size())))
        return rc;
    // initial heap: the input spikes queued before the first run in load
    // order, then the update events
    std::vector<EqItem> h((size_t)cap + 1);
    int32_t hn = 0;
    {
        std::vector<size_t> cur(S->em.size(), 0);
        for (;;) {
            int be = -1;
            int64_t bs = 0;
            for (size_t e = 0; e < S->em.size(); e++) {
                EqHostEm &E = S->em[e];
                if (E.kind != kEqTrain) continue;
                const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
                if (cur[e] >= s0.tr_t.size()) continue;
                if (s0.tr_load[cur[e]] != 0) continue;
                if (be < 0 || s0.tr_seq[cur[e]] < bs) {
                    be = (int)e;
                    bs = s0.tr_seq[cur[e]];
                }
            }
            if (be < 0) break;
            EqHostEm &E = S->em[(size_t)be];
            const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
            const size_t q = cur[(size_t)be]++;
            const int32_t nr = s0.tr_neuron[q];
            // a neuron without output synapses propagates nothing
            if (E.h_nd_ptr[(size_t)nr + 1] > E.h_nd_ptr[(size_t)nr])
                eq_insert(h.data(), &hn, eq_item(s0.tr_t[q], kEqPrioSpike, kEqInput, (uint32_t)be, 0, (uint32_t)q));
        }
        for (size_t e = 0; e < S->em.size(); e++) {
            EqHostEm &E = S->em[e];
            if (E.kind != kEqTrain) continue;
            const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
            int32_t q = 0;
            while (q < (int32_t)s0.tr_t.size() && s0.tr_load[(size_t)q] == 0) q++;
            E.in_loaded = q;
        }
    }
    std::vector<int> order;
    for (size_t e = 0; e < S->em.size(); e++)
        if (S->em[e].kind != kEqTrain) order.push_back((int)e);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return S->em[(size_t)a].rank < S->em[(size_t)b].rank; });
    for (int e : order) eq_insert(h.data(), &hn, eq_item(S->em[(size_t)e].dev.h, kEqPrioTime, kEqTime, (uint32_t)e, 0, 0));
    HIP_TRY(hipMemcpy(S->heap, h.data(), sizeof(EqItem) * ((size_t)hn + 1), hipMemcpyHostToDevice));
    std::vector<int32_t> stv(kEqSt + S->em.size(), 1);
    stv[0] = hn;
    stv[1] = 0;
    stv[2] = -1;
    stv[3] = -1;
    HIP_TRY(hipMemcpy(S->st, stv.data(), sizeof(int32_t) * stv.size(), hipMemcpyHostToDevice));
    std::vector<EqEmitter> dev;
    for (EqHostEm &E : S->em) {
        E.dev.in_lo = E.dev.in_hi = 0;
        dev.push_back(E.dev);
    }
    HIP_TRY(hipMemcpy(S->d_em, dev.data(), sizeof(EqEmitter) * dev.size(), hipMemcpyHostToDevice));
    S->next_first = 1;
    return 0;
}

EqArgs eq_args(const EqSim *S, StepRef step, int resume, double stop_t) {
    EqArgs a;
    std::memset(&a, 0, sizeof(a));
    a.em = S->d_em;
    a.n_em = (int)S->em.size();
    a.words = S->words;
    a.max_delay = S->max_delay;
    a.ring = S->ring;
    a.delay_s = S->delay_s;
    a.heap = S->heap;
    a.st = S->st;
    a.cap = S->cap;
    a.stop_t = stop_t;
    a.resume = resume;
    a.with_stop = 1;
    a.clear_step = -1;
    a.step = step;
    a.err = S->err;
    a.pf_em = -1;
    for (size_t e = 0; e < S->em.size(); e++)
        if (S->em[e].kind == kEqPop) {
            a.pf_em = (int)e;
            a.pf_tiles = S->em[e].net->tile_count;
            break;
        }
    return a;
}

// Before a run of updates [first, first + n_steps): input spikes added to
// the trains since the queue last saw them are uploaded, then the resuming
// replay queues them and the run's stop event and finishes the window
// before update `first`.
int eq_resume(EqSim *S, int64_t first, int n_steps, double stop, hipStream_t st) {
    if (first != S->next_first) return fail(EDLUT_E_STATE, "the event-queue order needs runs that continue each other");
    const int64_t last = first + n_steps - 1;
    const double t_last = S->time_of(last);
    const double t_next = t_last + S->h_pop;
    const bool with_stop = stop != -2.0;   // -2: the run continues into the next one (no stop event)
    if (stop < 0) stop = t_last;
    else if (!(stop >= t_last && stop < t_next))
        return fail(EDLUT_E_ARG, "stop time must lie between the run's last update and the next one");
    bool changed = false;
    for (EqHostEm &E : S->em) {
        if (E.kind != kEqTrain) continue;
        const Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
        const int32_t total = (int32_t)s0.tr_t.size();
        if (total < E.in_loaded) return fail(EDLUT_E_STATE, "input spikes already queued were removed from a train");
        int32_t hi = E.in_loaded;
        while (hi < total && s0.tr_load[(size_t)hi] <= first) hi++;
        // re-upload the records (and positions, which follow the lists)
        int rc;
        if (total > 0) {
            if (E.in_t) (void)hipFree(E.in_t);
            if (E.in_neuron) (void)hipFree(E.in_neuron);
            if (E.in_pos) (void)hipFree(E.in_pos);
            if (E.in_seq) (void)hipFree(E.in_seq);
            E.in_t = nullptr;
            E.in_neuron = E.in_pos = nullptr;
            E.in_seq = nullptr;
            if ((rc = dalloc(&E.in_t, (size_t)total)) || (rc = dalloc(&E.in_neuron, (size_t)total)) ||
                (rc = dalloc(&E.in_pos, (size_t)total)) || (rc = dalloc(&E.in_seq, (size_t)total)))
                return rc;
            HIP_TRY(hipMemcpy(E.in_t, s0.tr_t.data(), sizeof(double) * total, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E.in_neuron, s0.tr_neuron.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E.in_pos, s0.tr_pos.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(E.in_seq, s0.tr_seq.data(), sizeof(int64_t) * total, hipMemcpyHostToDevice));
        }
        if (s0.tr_longest > E.dev.cmax) return fail(EDLUT_E_CAP, "a train's update lists outgrew its rank table");
        E.dev.in_t = E.in_t;
        E.dev.in_neuron = E.in_neuron;
        E.dev.in_pos = E.in_pos;
        E.dev.in_seq = E.in_seq;
        E.dev.in_lo = E.in_loaded;
        E.dev.in_hi = hi;
        E.in_loaded = hi;
        changed = true;
    }
    // the populations' update tiles (their layout follows the launch shape)
    for (EqHostEm &E : S->em) {
        if (E.kind != kEqPop) continue;
        edlut_population *p = E.net->pop;
        E.dev.tile_count = E.net->tile_count;
        E.dev.staging = E.net->staging;
        E.dev.n_tiles = p->update_tiles();
        E.dev.tile_size = kBlock * p->items;
    }
    std::vector<EqEmitter> dev, dev2;
    for (EqHostEm &E : S->em) dev.push_back(E.dev);
    HIP_TRY(hipMemcpy(S->d_em, dev.data(), sizeof(EqEmitter) * dev.size(), hipMemcpyHostToDevice));
    EqArgs ra = eq_args(S, StepRef{nullptr, (int32_t)(first - 1)}, 1, stop);
    ra.with_stop = with_stop ? 1 : 0;
    // an explicit stop time may fall inside the last window: its groups
    // still queued at the stop must hold no rank
    ra.clear_step = (with_stop && stop > t_last) ? (int32_t)last : -1;
    hipLaunchKernelGGL(k_eventq, dim3(1), dim3(kBlock), 0, st, ra);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    if (changed) {
        // the run's own replays queue no input spikes
        for (EqHostEm &E : S->em) E.dev.in_lo = E.dev.in_hi = 0;
        for (EqHostEm &E : S->em) dev2.push_back(E.dev);
        HIP_TRY(hipMemcpy(S->d_em, dev2.data(), sizeof(EqEmitter) * dev2.size(), hipMemcpyHostToDevice));
    }
    S->next_first = last + 1;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Kernel dispatch tables.
// ---------------------------------------------------------------------------
namespace {

using UpdateFn = void (*)(UpdateArgs);

// Granule BDF: Newton corrections per step that can take the reference's
// row-swapping arrangement (zero pivot or non-finite values; in practice
// none); more in one step is reported as a capacity error.
constexpr int kBdfWsSlots = 4096;

// Dynamic LDS of a population's update kernel: the granule's BDF kernel
// stages the channel table (1024 x 26 floats) there.
size_t update_lds_bytes(const edlut_population *p) {
    return (p->model == EDLUT_GRANULE && p->integ == EDLUT_BDF) ? sizeof(float) * kGrTable * kGrVars : 0;
}

// Float parameters of each model's GetDefaultParameters() vector.
int model_param_count(int model) {
    static const int n[7] = {10, 14, 13, 0, 8, 10, 10};
    return model >= 0 && model < 7 ? n[model] : -1;
}

// Neurons per thread of an ungrouped population's update kernel: the
// granule's kernels hold two (19 state variables each; the BDF step keeps a
// 15 x 15 matrix per neuron), every other kernel four.
int plain_items(int model, int integrator) {
    (void)integrator;
    return model == EDLUT_GRANULE ? 2 : kItems;
}

// Plain AdEx and LIF populations with spike-driven synapse types only
// (mask 0..3) under Euler, RK2 or RK4 also have a two-neurons-per-thread
// update kernel; the autotuner picks between the two tile shapes.
bool plain_two_items(const edlut_population *p) {
    return !p->grouped && (p->model == EDLUT_ADEX || p->model == EDLUT_LIF) && (p->mask & ~3u) == 0 &&
           (p->integ == EDLUT_EULER || p->integ == EDLUT_RK2 || p->integ == EDLUT_RK4);
}

template <class M, int I, unsigned MASK>
void fill_one(UpdateFn *t) {
    t[(I * 16 + (int)MASK) * 2 + 0] = k_update<M, I, MASK, false, kItems, false>;
    t[(I * 16 + (int)MASK) * 2 + 1] = k_update<M, I, MASK, true, kItems, false>;
}

template <class M, int I, unsigned... MS>
void fill_masks(UpdateFn *t, std::integer_sequence<unsigned, MS...>) {
    (fill_one<M, I, MS>(t), ...);
}

template <class M, int I, unsigned MASK>
void fill_two_one(UpdateFn *t) {
    t[(I * 4 + (int)MASK) * 2 + 0] = k_update<M, I, MASK, false, 2, false>;
    t[(I * 4 + (int)MASK) * 2 + 1] = k_update<M, I, MASK, true, 2, false>;
}
template <class M, int I, unsigned... MS>
void fill_two(UpdateFn *t, std::integer_sequence<unsigned, MS...>) {
    (fill_two_one<M, I, MS>(t), ...);
}
template <class M>
struct TwoItemTable {
    UpdateFn fn[3 * 4 * 2];
    TwoItemTable() {
        fill_two<M, kEuler>(fn, std::make_integer_sequence<unsigned, 4>{});
        fill_two<M, kRK2>(fn, std::make_integer_sequence<unsigned, 4>{});
        fill_two<M, kRK4>(fn, std::make_integer_sequence<unsigned, 4>{});
    }
};

template <class M>
struct Table {
    UpdateFn fn[3 * 16 * 2];
    Table() {
        fill_masks<M, kEuler>(fn, std::make_integer_sequence<unsigned, 16>{});
        fill_masks<M, kRK2>(fn, std::make_integer_sequence<unsigned, 16>{});
        fill_masks<M, kRK4>(fn, std::make_integer_sequence<unsigned, 16>{});
    }
};

// Grouped populations: spike-driven synapse types (mask 0..3), with and
// without external drive, two tile shapes (items per thread 2 or 4).
constexpr int kGroupItems[2] = {2, 4};
template <class M, int I, unsigned MASK>
void fill_grouped_one(UpdateFn *t) {
    t[((I * 4 + (int)MASK) * 2 + 0) * 2 + 0] = k_update<M, I, MASK, false, 2, true>;
    t[((I * 4 + (int)MASK) * 2 + 1) * 2 + 0] = k_update<M, I, MASK, true, 2, true>;
    t[((I * 4 + (int)MASK) * 2 + 0) * 2 + 1] = k_update<M, I, MASK, false, 4, true>;
    t[((I * 4 + (int)MASK) * 2 + 1) * 2 + 1] = k_update<M, I, MASK, true, 4, true>;
}
template <class M, int I, unsigned... MS>
void fill_grouped(UpdateFn *t, std::integer_sequence<unsigned, MS...>) {
    (fill_grouped_one<M, I, MS>(t), ...);
}
template <class M>
struct GroupedTable {
    UpdateFn fn[3 * 4 * 2 * 2];
    GroupedTable() {
        fill_grouped<M, kEuler>(fn, std::make_integer_sequence<unsigned, 4>{});
        fill_grouped<M, kRK2>(fn, std::make_integer_sequence<unsigned, 4>{});
        fill_grouped<M, kRK4>(fn, std::make_integer_sequence<unsigned, 4>{});
    }
};

// Granule and Purkinje cells: one kernel per integrator and drive (the
// synapse types are a run-time mask), ITEMS neurons per thread.
template <class M, int ITEMS>
struct MaskFreeTable {
    UpdateFn fn[7 * 2];   // by integrator (edlut_integrator) and drive
    MaskFreeTable() {
        fn[0] = k_update<M, kEuler, 15u, false, ITEMS, false>;
        fn[1] = k_update<M, kEuler, 15u, true, ITEMS, false>;
        fn[2] = k_update<M, kRK2, 15u, false, ITEMS, false>;
        fn[3] = k_update<M, kRK2, 15u, true, ITEMS, false>;
        fn[4] = k_update<M, kRK4, 15u, false, ITEMS, false>;
        fn[5] = k_update<M, kRK4, 15u, true, ITEMS, false>;
        if constexpr (std::is_same<M, GranuleModel>::value) {
            fn[6] = k_update_granule_bdf<false>;
            fn[7] = k_update_granule_bdf<true>;
        } else {
            fn[6] = k_update_bdf<M, 15u, false, float, ITEMS>;
            fn[7] = k_update_bdf<M, 15u, true, float, ITEMS>;
        }
        fn[8] = k_update<M, kBifEuler, 15u, false, ITEMS, false>;
        fn[9] = k_update<M, kBifEuler, 15u, true, ITEMS, false>;
        fn[10] = k_update<M, kBifRK2, 15u, false, ITEMS, false>;
        fn[11] = k_update<M, kBifRK2, 15u, true, ITEMS, false>;
        fn[12] = k_update<M, kBifRK4, 15u, false, ITEMS, false>;
        fn[13] = k_update<M, kBifRK4, 15u, true, ITEMS, false>;
    }
};

// Bifixed kernels of AdEx, LIF and HH: the synapse types are a run-time mask.
template <class M>
struct BifixedTable {
    UpdateFn fn[3 * 2];
    BifixedTable() {
        constexpr unsigned RT = 15u | kRtMask;
        fn[0] = k_update<M, kBifEuler, RT, false, kItems, false>;
        fn[1] = k_update<M, kBifEuler, RT, true, kItems, false>;
        fn[2] = k_update<M, kBifRK2, RT, false, kItems, false>;
        fn[3] = k_update<M, kBifRK2, RT, true, kItems, false>;
        fn[4] = k_update<M, kBifRK4, RT, false, kItems, false>;
        fn[5] = k_update<M, kBifRK4, RT, true, kItems, false>;
    }
};

template <class M, typename T, unsigned... MS>
void fill_bdf(UpdateFn *t, std::integer_sequence<unsigned, MS...>) {
    ((t[MS * 2 + 0] = k_update_bdf<M, MS, false, T>, t[MS * 2 + 1] = k_update_bdf<M, MS, true, T>), ...);
}

template <class M, typename T>
struct BdfTable {
    UpdateFn fn[16 * 2];
    BdfTable() { fill_bdf<M, T>(fn, std::make_integer_sequence<unsigned, 16>{}); }
};

// Update kernel of a population: null when no kernel covers the combination
// (grouped populations with NMDA or external-current synapses).
UpdateFn pick_update(const edlut_population *p, bool ext) {
    const int m = (int)(p->mask & 15u);
    const int integ = p->integ, model = p->model;
    if (p->grouped) {
        if (m > 3 || integ == EDLUT_BDF) return nullptr;
        static GroupedTable<AdExModel> adex_g;
        static GroupedTable<LIFModel> lif_g;
        const int it = p->items == 2 ? 0 : (p->items == 4 ? 1 : -1);
        if (it < 0) return nullptr;
        const int idx = ((integ * 4 + m) * 2 + (ext ? 1 : 0)) * 2 + it;
        if (model == EDLUT_ADEX) return adex_g.fn[idx];
        if (model == EDLUT_LIF) return lif_g.fn[idx];
        return nullptr;
    }
    if (model == EDLUT_IZHIKEVICH || model == EDLUT_LIF_IS) {
        if (p->items != kItems) return nullptr;
        static MaskFreeTable<IzhModel, kItems> iz;
        static MaskFreeTable<LifIsModel, kItems> li;
        return (model == EDLUT_IZHIKEVICH ? iz.fn : li.fn)[integ * 2 + (ext ? 1 : 0)];
    }
    if (model == EDLUT_GRANULE || model == EDLUT_PURKINJE) {
        if (p->items != plain_items(model, integ)) return nullptr;
        static MaskFreeTable<GranuleModel, 2> gr;
        static MaskFreeTable<PurkinjeModel, kItems> pk;
        static const bool lds_ok = [] {
            // more than the default 64 KB of dynamic LDS for the table-staging kernels
            bool ok = true;
            for (int q = 6; q < 8; q++)
                ok = ok && hipFuncSetAttribute((const void *)gr.fn[q], hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)(sizeof(float) * kGrTable * kGrVars)) == hipSuccess;
            return ok;
        }();
        if (model == EDLUT_GRANULE && integ == EDLUT_BDF && !lds_ok) return nullptr;
        const int idx = integ * 2 + (ext ? 1 : 0);
        return model == EDLUT_GRANULE ? gr.fn[idx] : pk.fn[idx];
    }
    if (p->items == 2 && plain_two_items(p)) {
        static TwoItemTable<AdExModel> adex2;
        static TwoItemTable<LIFModel> lif2;
        const int idx = (integ * 4 + m) * 2 + (ext ? 1 : 0);
        return model == EDLUT_ADEX ? adex2.fn[idx] : lif2.fn[idx];
    }
    if (p->items != kItems) return nullptr;
    if (integ >= EDLUT_BIFIXED_EULER) {
        static BifixedTable<AdExModel> adex_s;
        static BifixedTable<LIFModel> lif_s;
        static BifixedTable<HHModel> hh_s;
        const int idx = (integ - EDLUT_BIFIXED_EULER) * 2 + (ext ? 1 : 0);
        if (model == EDLUT_ADEX) return adex_s.fn[idx];
        if (model == EDLUT_LIF) return lif_s.fn[idx];
        if (model == EDLUT_HH) return hh_s.fn[idx];
        return nullptr;
    }
    if (integ == EDLUT_BDF) {
        static BdfTable<AdExModel, float> adex_b;
        static BdfTable<LIFModel, float> lif_b;
        static BdfTable<HHModel, float> hh_b;
        static BdfTable<HHModel, double> hh_b64;
        const int idx = m * 2 + (ext ? 1 : 0);
        if (model == EDLUT_ADEX) return adex_b.fn[idx];
        if (model == EDLUT_LIF) return lif_b.fn[idx];
        if (model == EDLUT_HH) return p->newton_fp64 ? hh_b64.fn[idx] : hh_b.fn[idx];
        return nullptr;
    }
    static Table<AdExModel> adex;
    static Table<LIFModel> lif;
    static Table<HHModel> hh;
    const int idx = (integ * 16 + m) * 2 + (ext ? 1 : 0);
    if (model == EDLUT_ADEX) return adex.fn[idx];
    if (model == EDLUT_LIF) return lif.fn[idx];
    if (model == EDLUT_HH) return hh.fn[idx];
    return nullptr;
}

int check_err(edlut_population *p) {
    uint32_t e = 0;
    HIP_TRY(hipMemcpyAsync(&e, p->err, sizeof(uint32_t), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (e & 1u) return fail(EDLUT_E_STATE, "spike compaction look-back exceeded its spin limit");
    if (e & 2u)
        return fail(EDLUT_E_CAP, "spike log capacity exceeded (read and clear the log between runs, or raise "
                                 "its capacity)");
    if (e & 4u) return fail(EDLUT_E_CAP, "input event buffer capacity exceeded");
    if (e & 8u) return fail(EDLUT_E_STATE, "input fold segment wait exceeded its spin limit");
    if (e & 16u) return fail(EDLUT_E_CAP, "BDF fallback working storage exhausted in one step");
    if (e & 64u) return fail(EDLUT_E_CAP, "event queue capacity exceeded");
    if (e & 128u) return fail(EDLUT_E_STATE, "event queue holds an event of an unknown kind");
    if (e & 256u)
        return fail(EDLUT_E_ARG, "an input event in device memory names a target out of range or a synapse "
                                 "type that is not enabled (dropped)");
    return 0;
}

UpdateArgs make_update_args(edlut_population *p, int32_t step) {
    UpdateArgs a;
    std::memset(&a, 0, sizeof(a));
    a.state = p->state;
    a.stride = p->stride;
    a.n = p->n;
    a.n_tiles = p->update_tiles();
    a.step = StepRef{nullptr, step};
    a.mp = p->mp;
    // the granule and Purkinje kernels read the enabled synapse types at run
    // time: take them from the population (connections may have enabled more)
    a.mp.gran.mask = a.mp.pkj.mask = a.mp.izh.mask = a.mp.lifis.mask = p->mask;
    a.mp.adex.mask = a.mp.lif.mask = a.mp.hh.mask = p->mask;   // run-time-mask kernels
    a.k = p->k;
    a.n_sub = p->n_sub;
    a.exp_lut = p->exp_lut;
    a.nmda_lut = p->nmda_lut;
    a.refr = p->refr;
    a.refr_len = p->refr_len;
    a.last_spike = p->last_spike;
    a.chan = p->chan;
    a.bdf.order = p->bdf_order;
    a.bdf.C = p->bdf_c;
    a.bdf.d = p->bdf_d;
    a.bdf.prev = p->bdf_prev;
    a.bdf.st = p->bdf_st;
    a.bdf.ws = p->bdf_ws;
    a.bdf.ws_count = p->bdf_ws_count;
    a.bdf.ws_cap = p->bdf_ws ? kBdfWsSlots : 0;
    a.bdf.err = p->err;
    a.bdf.spk_bits = p->bdf_bits;
    a.bdf.done = p->bdf_done;
    a.grp.tab = p->d_groups;
    a.grp.chunk_group = p->d_chunk_group;
    return a;
}

CompactArgs compact_args(int32_t *tile_count, int32_t *staging, int n_tiles, int tile_size, int32_t *out_idx,
                         int32_t *out_count, int ring, int64_t slot_len, int32_t *log_step, int32_t *log_idx,
                         int64_t *log_count, int64_t log_cap, uint32_t *err, StepRef step) {
    CompactArgs c;
    c.tile_count = tile_count;
    c.staging = staging;
    c.n_tiles = n_tiles;
    c.tile_size = tile_size;
    c.out_idx = out_idx;
    c.out_count = out_count;
    c.ring = ring;
    c.slot_len = slot_len;
    c.log_step = log_step;
    c.log_idx = log_idx;
    c.log_count = log_count;
    c.log_cap = log_cap;
    c.err = err;
    c.step = step;
    c.clk_next = nullptr;
    c.clk_first = 0;
    c.clk_count = 0;
    return c;
}

inline int compact_blocks(int n_tiles) { return (n_tiles + kCompactTiles - 1) / kCompactTiles; }

void launch_compact(hipStream_t st, int32_t *tile_count, int32_t *staging, int n_tiles,
                    int tile_size, int32_t *out_idx, int32_t *out_count, int ring, int64_t slot_len,
                    int32_t *log_step, int32_t *log_idx, int64_t *log_count, int64_t log_cap,
                    uint32_t *err, StepRef step) {
    CompactArgs c = compact_args(tile_count, staging, n_tiles, tile_size, out_idx, out_count, ring, slot_len,
                                 log_step, log_idx, log_count, log_cap, err, step);
    hipLaunchKernelGGL(k_compact, dim3(compact_blocks(n_tiles)), dim3(kBlock), 0, st, c);
}

using RingFoldFn = void (*)(FoldTarget, RingEvents);
template <int LANES>
RingFoldFn ring_fold_fn(int pass) {
    if (pass == kCount) return k_fold_ring<kCount, LANES>;
    if (pass == kScatter) return k_fold_ring<kScatter, LANES>;
    return k_fold_ring<kStatic, LANES>;
}
RingFoldFn pick_ring_fold(int pass, int lanes) {
    switch (lanes) {
        case 8: return ring_fold_fn<8>(pass);
        case 16: return ring_fold_fn<16>(pass);
        case 32: return ring_fold_fn<32>(pass);
        case 64: return ring_fold_fn<64>(pass);
        default: return nullptr;
    }
}
using CompactFoldEqFn = void (*)(CompactArgs, FoldTarget, RingEvents, int, EqArgs);
CompactFoldEqFn pick_compact_fold_eq(int lanes) {
    switch (lanes) {
        case 8: return k_compact_fold_eq<8>;
        case 16: return k_compact_fold_eq<16>;
        case 32: return k_compact_fold_eq<32>;
        case 64: return k_compact_fold_eq<64>;
        default: return nullptr;
    }
}
using CompactFoldFn = void (*)(CompactArgs, FoldTarget, RingEvents, int);
CompactFoldFn pick_compact_fold(int lanes) {
    switch (lanes) {
        case 8: return k_compact_fold<8>;
        case 16: return k_compact_fold<16>;
        case 32: return k_compact_fold<32>;
        case 64: return k_compact_fold<64>;
        default: return nullptr;
    }
}

// The network-mode fold (one pass into the keys' fixed segments).
int launch_ring_fold(const FoldTarget &t, hipStream_t st, const RingEvents &re, int lanes, int blocks) {
    RingFoldFn f = pick_ring_fold(kStatic, lanes);
    if (!f) return fail(EDLUT_E_ARG, "fold lanes must be 8, 16, 32 or 64");
    hipLaunchKernelGGL(f, dim3(blocks), dim3(kBlock), 0, st, t, re);
    return 0;
}

// Both passes of the fold of host-pushed events (plug-in mode).
int launch_push_fold(edlut_population *p, int par, int n_events, const PushEvents *pe) {
    if (n_events == 0) return 0;
    FoldTarget t = p->fold_target(par, false);
    int nb = (n_events + kBlock - 1) / kBlock;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL((k_fold_push<kCount>), dim3(nb), dim3(kBlock), 0, p->stream, t, *pe);
    hipLaunchKernelGGL((k_fold_push<kScatter>), dim3(nb), dim3(kBlock), 0, p->stream, t, *pe);
    HIP_TRY(hipGetLastError());
    return 0;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char *edlut_last_error(void) { return g_err.c_str(); }

const char *edlut_device_arch(void) { return "gfx950"; }

uint32_t edlut_drive_hash(uint64_t seed, int64_t step, uint32_t neuron) {
    return drive_hash(seed, step, neuron);
}

}  // extern "C"

namespace {

// Derived per-model constants of one parameter vector (the reference's
// InitializeStates / loadNeuronModel, e.g. AdExTimeDrivenModel.cpp:250-330),
// plus the initial state (VectorNeuronState_GPU_C_INTERFACE.cu:66-131).
void model_params(int model, int integrator, const float *x, float dt_ms, ModelParams &mp,
                  std::vector<float> &init, float &e_exc, float &e_inh) {
    const float elapsed = integrator == EDLUT_RK4 ? dt_ms * 0.5f : dt_ms;
    if (model == EDLUT_ADEX) {
        AdExParams &q = mp.adex;
        const float thr = x[2];
        q.a = x[0];
        q.b = x[1];
        q.inv_thr = inv_of(thr);
        q.v_thr = x[3];
        q.v_cap = x[3] + 6 * thr;
        q.inv_tau_w = inv_of(x[4]);
        q.e_exc = e_exc = x[5];
        q.e_inh = e_inh = x[6];
        q.e_reset = x[7];
        q.e_leak = x[8];
        q.g_leak = x[9];
        q.glk_thr = x[9] * thr;
        q.inv_c_m = inv_of(x[10]);
        q.cf[0] = expf(-elapsed * inv_of(x[11]));
        q.cf[1] = expf(-elapsed * inv_of(x[12]));
        q.cf[2] = expf(-elapsed * inv_of(x[13]));
        init = {x[8], 0, 0, 0, 0, 0};
        for (int j = 0; j < 6; j++) q.init[j] = init[j];
    } else if (model == EDLUT_HH) {
        HHParams &q = mp.hh;   // e_exc e_inh e_leak g_leak c_m v_thr tau_exc tau_inh tau_nmda g_na g_kd e_na e_k
        q.e_exc = e_exc = x[0];
        q.e_inh = e_inh = x[1];
        q.e_leak = x[2];
        q.g_leak = x[3];
        q.inv_c_m = inv_of(x[4]);
        q.inv_c_m_d = 1.0 / (double)x[4];
        q.g_na = x[9];
        q.g_kd = x[10];
        q.e_na = x[11];
        q.e_k = x[12];
        const float tau[3] = {x[6], x[7], x[8]};
        for (int j = 0; j < 3; j++) {
            q.cf[j] = expf(-elapsed * inv_of(tau[j]));
            q.cf_d[j] = std::exp(-(double)elapsed * (1.0 / (double)tau[j]));
        }
        // HHTimeDrivenModel.cpp:136-160: gates at their steady state at e_leak
        const std::vector<float> chan = build_hh_table(x[5]);
        const float *c = &chan[(size_t)hh_index(x[2]) * 6];
        const float m_inf = c[0] / (c[0] + c[1]);
        const float h_inf = c[2] / (c[2] + c[3]);
        const float n_inf = c[4] / (c[4] + c[5]);
        init = {x[2], m_inf, h_inf, n_inf, 0, 0, 0, 0};
        for (int j = 0; j < 8; j++) q.init[j] = init[j];
    } else if (model == EDLUT_GRANULE) {
        // EgidioGranuleCell_TimeDriven.cpp:26-58,90-93,144-200,382-389: no free
        // parameters; tau_exc 0.5, tau_inh 10, tau_nmda 15 ms (exp(-t/tau))
        GranuleParams &q = mp.gran;
        q.e_exc = e_exc = 0.0f;
        q.e_inh = e_inh = -80.0f;
        q.a_xkca = gr_q10(30.0f) * 2.5f;
        q.b_xkca = gr_q10(30.0f) * 1.5f;
        q.k_nernst = 1000 * (8.3134f * (30.0f + 273.15f) / 96485.309f) / 2.0f;
        q.cf[0] = expf(-elapsed / 0.5f);
        q.cf[1] = expf(-elapsed / 10.0f);
        q.cf[2] = expf(-elapsed / 15.0f);
        init = {-80.0f,       0.00047309535f, 1.0f,         0.00013423511f, 0.96227829f,
                0.00050020111f, 0.010183001f, 0.15685486f,  0.53565367f,    0.37337035f,
                0.00012384122f, 0.0021951104f, 0.89509747f, 0.00024031171f, 1e-04f,
                0.0f,         0.0f,           0.0f,         0.0f};
        for (int j = 0; j < 19; j++) q.init[j] = init[j];
    } else if (model == EDLUT_PURKINJE) {
        // e_exc e_inh v_thr e_leak tau_exc tau_inh tau_nmda tau_ref
        // (TimeDrivenPurkinjeCell.cpp:342-404; constants :62-68)
        PurkinjeParams &q = mp.pkj;
        q.e_exc = e_exc = x[0];
        q.e_inh = e_inh = x[1];
        q.v_thr = x[2];
        q.e_leak = x[3];
        const float spk_peak = 31.0f;
        const float tau_ref_0_5 = x[7] / 2;
        const float inv_half = inv_of(tau_ref_0_5);
        q.rise = (spk_peak - x[2]) * inv_half;
        q.fall = (x[3] - spk_peak) * inv_half;
        const float inv_area = 1.0f / (3.141592f * 0.0015f * 2.0f * 0.0008f);
        q.cur_scale = 1e-6 * inv_area;
        q.cf[0] = expf(-elapsed * inv_of(x[4]));
        q.cf[1] = expf(-elapsed * inv_of(x[5]));
        q.cf[2] = expf(-elapsed * inv_of(x[6]));
        // :139-153: Ca and M start at their steady state at e_leak
        const std::vector<float> chan = build_purkinje_table();
        const float *c = &chan[(size_t)chan_index(x[3], kPkMinV, kPkAux) * 4];
        init = {x[3], c[0] / c[1], c[2] / c[3], 0, 0, 0, 0};
        for (int j = 0; j < 7; j++) q.init[j] = init[j];
    } else if (model == EDLUT_LIF_IS) {
        // e_exc e_inh e_leak v_thr c_m tau_exc tau_inh tau_ref g_leak tau_nmda in SI
        // units (LIFTimeDrivenModel_IS.cpp:298-375, defaults :392-404)
        LifIsParams &q = mp.lifis;
        q.e_exc = e_exc = x[0];
        q.e_inh = e_inh = x[1];
        q.e_leak = x[2];
        q.v_thr = x[3];
        q.inv_c_m_nF = (float)(1. / ((double)x[4] * 1.e9));
        q.g_leak_nS = (float)((double)x[8] * 1.e9);
        q.cf[0] = expf(-elapsed * inv_of(x[5]));
        q.cf[1] = expf(-elapsed * inv_of(x[6]));
        q.cf[2] = expf(-elapsed * inv_of(x[9]));
        init = {x[2], 0, 0, 0, 0};
        for (int j = 0; j < 5; j++) q.init[j] = init[j];
    } else if (model == EDLUT_IZHIKEVICH) {
        // a b c d e_exc e_inh c_m tau_exc tau_inh tau_nmda
        // (IzhikevichTimeDrivenModel.cpp:298-375, defaults :389-401)
        IzhParams &q = mp.izh;
        q.a = x[0];
        q.b = x[1];
        q.c = x[2];
        q.d = x[3];
        q.e_exc = e_exc = x[4];
        q.e_inh = e_inh = x[5];
        q.inv_c_m = inv_of(x[6]);
        q.cf[0] = expf(-elapsed * inv_of(x[7]));
        q.cf[1] = expf(-elapsed * inv_of(x[8]));
        q.cf[2] = expf(-elapsed * inv_of(x[9]));
        // :131-135: resting equilibrium (lower root of 0.04V^2 + (5-b)V + 140 = 0, u = bV)
        const float b = x[1];
        const float veq = (((b - 5.0f) - sqrtf((5.0f - b) * (5.0f - b) - 22.4f)) / 0.08f);
        init = {veq, veq * b, 0, 0, 0, 0};
        for (int j = 0; j < 6; j++) q.init[j] = init[j];
    } else {
        LIFParams &q = mp.lif;
        q.e_exc = e_exc = x[0];
        q.e_inh = e_inh = x[1];
        q.e_leak = x[2];
        q.v_thr = x[3];
        q.inv_c_m = inv_of(x[4]);
        q.g_leak = x[8];
        q.cf[0] = expf(-elapsed * inv_of(x[5]));
        q.cf[1] = expf(-elapsed * inv_of(x[6]));
        q.cf[2] = expf(-elapsed * inv_of(x[9]));
        init = {x[2], 0, 0, 0, 0};
        for (int j = 0; j < 5; j++) q.init[j] = init[j];
    }
}

// Common constructor of plain (one group, no padding) and grouped populations.
int create_population(int model, int integrator, int bdf_order, const float *params, int n_params,
                      float dt_s, int n_groups, const int32_t *sizes, bool grouped, int device,
                      edlut_population **out) {
    if (!out) return fail(EDLUT_E_ARG, "out is null");
    *out = nullptr;
    if (model < EDLUT_LIF || model > EDLUT_LIF_IS)
        return fail(EDLUT_E_ARG, "neuron model %d is not available in this build", model);
    if (integrator < EDLUT_EULER || integrator > EDLUT_BIFIXED_RK4)
        return fail(EDLUT_E_ARG, "integration method %d is not available for this model", integrator);
    if (integrator == EDLUT_BDF && (bdf_order < 1 || bdf_order > 6))
        return fail(EDLUT_E_ARG, "bdf_order must be in [1, 6] (BDFn_GPU_C_INTERFACE.cuh:275-288)");
    const bool bifixed = integrator >= EDLUT_BIFIXED_EULER;
    if (bifixed && bdf_order < 1)
        return fail(EDLUT_E_ARG, "n_steps must be positive (BifixedStep_GPU_C_INTERFACE.cuh:144-152)");
    // the base method's sub-step: seconds divided in float
    // (BifixedStep_GPU_C_INTERFACE.cuh:157-158)
    const int n_sub = bifixed ? bdf_order : 1;
    const int base_integ = bifixed ? integrator - EDLUT_BIFIXED_EULER : integrator;
    const float step_s = bifixed ? dt_s / (float)n_sub : dt_s;
    const int need = model_param_count(model);
    if (n_params != need || (need > 0 && !params))
        return fail(EDLUT_E_ARG, "model %d takes %d parameters, got %d", model, need, n_params);
    if (!(dt_s > 0.0f)) return fail(EDLUT_E_ARG, "integration step must be positive");
    if (n_groups <= 0 || !sizes) return fail(EDLUT_E_ARG, "no groups");
    if (grouped) {
        if ((model != EDLUT_ADEX && model != EDLUT_LIF) || integrator > EDLUT_RK4)
            return fail(EDLUT_E_ARG, "grouped populations take AdEx or LIF under Euler, RK2 or RK4");
    }
    int64_t total = 0;
    std::vector<int32_t> start(n_groups);
    for (int g = 0; g < n_groups; g++) {
        if (sizes[g] <= 0) return fail(EDLUT_E_ARG, "group %d: size must be positive", g);
        start[g] = (int32_t)total;
        total += grouped ? ((int64_t)sizes[g] + kGroupAlign - 1) / kGroupAlign * kGroupAlign : sizes[g];
        if (total > (1 << 30) - 1) return fail(EDLUT_E_ARG, "population too large");
    }
    for (int i = 0; i < n_params * n_groups; i++)
        if (!std::isfinite(params[i])) return fail(EDLUT_E_ARG, "parameter %d is not finite", i % n_params);
    const int n_neurons = (int)total;

    float e_exc = 0, e_inh = 0;
    std::vector<float> init;
    ModelParams mp0;
    std::memset(&mp0, 0, sizeof(mp0));
    // IntegrationMethodFast.h:191-193: the step in the model's time scale
    // (milliseconds; seconds for the SI-unit LIF)
    const float time_scale = model == EDLUT_LIF_IS ? 1.0f : 1000.0f;
    const float dt_ms = dt_s * time_scale;
    // the integrator's step in the model scale (Bifixed: the sub-step's
    // seconds times the time scale, BifixedStep_GPU2.cuh:81-82)
    const float int_ms = step_s * time_scale;
    model_params(model, base_integ, params, int_ms, mp0, init, e_exc, e_inh);
    std::vector<GroupInfo> gtab;
    std::vector<float> st_init;   // per group initial state
    if (grouped) {
        gtab.resize(n_groups);
        for (int g = 0; g < n_groups; g++) {
            const float *x = params + (size_t)g * n_params;
            float ee, ei;
            std::vector<float> gi;
            std::memset(&gtab[g], 0, sizeof(GroupInfo));
            model_params(model, integrator, x, dt_ms, gtab[g].mp, gi, ee, ei);
            // shared tables: NMDA (reversal potentials), refractory clock
            if (ee != e_exc || ei != e_inh)
                return fail(EDLUT_E_ARG, "group %d: e_exc/e_inh must equal group 0's (shared NMDA table)", g);
            if (model == EDLUT_LIF && x[7] != params[7])
                return fail(EDLUT_E_ARG, "group %d: tau_ref must equal group 0's (shared refractory table)", g);
            gtab[g].start = start[g];
            gtab[g].n = sizes[g];
            st_init.insert(st_init.end(), gi.begin(), gi.end());
        }
    }
    HIP_TRY(hipSetDevice(device));
    edlut_population *p = new edlut_population;
    p->model = model;
    p->integ = integrator;
    p->bdf_order = bifixed ? 0 : bdf_order;
    p->n_sub = n_sub;
    p->n = n_neurons;
    p->device = device;
    p->params.assign(params, params + (size_t)n_params * n_groups);
    p->dt_s = dt_s;
    p->dt_ms = dt_ms;
    p->k.dt = int_ms;
    p->k.h = int_ms * 0.5f;
    p->k.s6 = int_ms * 0.166666666667f;
    p->grouped = grouped;
    p->g_start = start;
    p->g_size.assign(sizes, sizes + n_groups);
    p->mp = mp0;

    float nmda_aux = 0.0f;
    std::vector<float> nmda = build_nmda_table(e_exc, e_inh, &nmda_aux, model == EDLUT_LIF_IS ? -62.0f : -0.062f);
    if (model == EDLUT_ADEX) p->mp.adex.nmda_aux = nmda_aux;
    else if (model == EDLUT_HH) p->mp.hh.nmda_aux = nmda_aux;
    else if (model == EDLUT_GRANULE) p->mp.gran.nmda_aux = nmda_aux;
    else if (model == EDLUT_PURKINJE) p->mp.pkj.nmda_aux = nmda_aux;
    else if (model == EDLUT_IZHIKEVICH) p->mp.izh.nmda_aux = nmda_aux;
    else if (model == EDLUT_LIF_IS) p->mp.lifis.nmda_aux = nmda_aux;
    else p->mp.lif.nmda_aux = nmda_aux;
    for (GroupInfo &g : gtab) {
        if (model == EDLUT_ADEX) g.mp.adex.nmda_aux = nmda_aux;
        else g.mp.lif.nmda_aux = nmda_aux;
    }
    static const int kNv[7] = {5, 6, 8, 19, 7, 6, 5}, kNd[7] = {1, 2, 4, 15, 3, 2, 1};
    p->nv = kNv[model];
    p->nd = kNd[model];
    // the granule's kernels hold two neurons per thread (19 state variables
    // each); every other kernel holds four
    p->items = plain_items(model, integrator);
    std::vector<float> chan;
    if (model == EDLUT_HH) chan = build_hh_table(params[5]);
    if (model == EDLUT_GRANULE) chan = build_granule_table();
    if (model == EDLUT_PURKINJE) chan = build_purkinje_table();
    std::vector<float> expt = build_exp_table();
    std::vector<float> refr;
    // clock tables over the integrator's (sub-)steps since the last spike
    if (model == EDLUT_LIF) refr = build_refractory(step_s, int_ms, params[7]);
    if (model == EDLUT_LIF_IS) refr = build_refractory(step_s, int_ms, params[7], 1.0f);
    if (model == EDLUT_PURKINJE) refr = build_purkinje_clock(step_s, int_ms, params[2], params[3], params[7]);
    // entries: LIF one fraction each, Purkinje a (mode, V) pair each
    p->refr_len = (int)refr.size() / (model == EDLUT_PURKINJE ? 2 : 1);
    p->lut_sums[0] = fnv1a(expt.data(), expt.size() * sizeof(float));
    p->lut_sums[1] = fnv1a(nmda.data(), nmda.size() * sizeof(float));
    p->lut_sums[2] = fnv1a(refr.data(), refr.size() * sizeof(float));

    // rows padded to whole update tiles: every kernel load stays in bounds
    p->stride = ((int64_t)n_neurons + kUpdTile - 1) / kUpdTile * kUpdTile;
    p->n_tiles = (n_neurons + kUpdTile - 1) / kUpdTile;
    const size_t max_tiles = (size_t)(p->stride / kFlagTile);

    int rc = 0;
    auto cleanup = [&](int code) {
        edlut_population_destroy(p);
        return code;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(EDLUT_E_HIP, "hipStreamCreate failed"));
    if ((rc = dalloc(&p->state, (size_t)p->nv * p->stride))) return cleanup(rc);
    if ((rc = dalloc(&p->exp_lut, kExpSize))) return cleanup(rc);
    if ((rc = dalloc(&p->nmda_lut, kNmdaSize))) return cleanup(rc);
    if ((rc = dalloc(&p->refr, refr.size() + 1))) return cleanup(rc);
    if (!chan.empty()) {
        if ((rc = dalloc(&p->chan, chan.size()))) return cleanup(rc);
        if (hipMemcpy(p->chan, chan.data(), chan.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(EDLUT_E_HIP, "hipMemcpy failed"));
    }
    if (grouped) {
        const size_t chunks = (size_t)(p->stride / kGroupAlign);
        std::vector<uint32_t> cg(chunks, (uint32_t)(n_groups - 1));
        for (int g = 0; g < n_groups; g++) {
            const int64_t c0 = start[g] / kGroupAlign;
            const int64_t c1 = g + 1 < n_groups ? start[g + 1] / kGroupAlign : (int64_t)chunks;
            for (int64_t c = c0; c < c1; c++) cg[(size_t)c] = (uint32_t)g;
        }
        if ((rc = dalloc(&p->d_groups, gtab.size())) || (rc = dalloc(&p->d_chunk_group, chunks)))
            return cleanup(rc);
        if (hipMemcpy(p->d_groups, gtab.data(), gtab.size() * sizeof(GroupInfo), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->d_chunk_group, cg.data(), chunks * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(EDLUT_E_HIP, "hipMemcpy failed"));
    }
    if (integrator == EDLUT_BDF) {
        float C[49];
        bdf_coefficients(C);
        if ((rc = dalloc(&p->bdf_c, 49))) return cleanup(rc);
        if ((rc = dalloc(&p->bdf_st, p->stride))) return cleanup(rc);
        if ((rc = p->alloc_bdf_history(sizeof(float)))) return cleanup(rc);
        if (hipMemcpy(p->bdf_c, C, sizeof(C), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(p->bdf_st, 0, sizeof(int32_t) * p->stride) != hipSuccess)
            return cleanup(fail(EDLUT_E_HIP, "BDF initialisation failed"));
        if (model == EDLUT_GRANULE) {
            if ((rc = dalloc(&p->bdf_ws, (size_t)kBdfWsSlots * kGrWsFloats)) || (rc = dalloc(&p->bdf_ws_count, 2)) ||
                (rc = dalloc(&p->bdf_bits, (size_t)(p->stride / 32))) || (rc = dalloc(&p->bdf_done, 1)))
                return cleanup(rc);
            if (hipMemset(p->bdf_ws_count, 0, 2 * sizeof(uint32_t)) != hipSuccess ||
                hipMemset(p->bdf_bits, 0, sizeof(uint32_t) * (size_t)(p->stride / 32)) != hipSuccess ||
                hipMemset(p->bdf_done, 0, sizeof(uint32_t)) != hipSuccess ||
                hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
                return cleanup(fail(EDLUT_E_HIP, "BDF initialisation failed"));
        }
    }
    if ((rc = dalloc(&p->tile_count, padded_tiles((int)max_tiles)))) return cleanup(rc);
    if ((rc = dalloc(&p->staging, (size_t)p->stride))) return cleanup(rc);
    if ((rc = dalloc(&p->spk_idx, n_neurons))) return cleanup(rc);
    if ((rc = dalloc(&p->spk_cnt, 1))) return cleanup(rc);
    if ((rc = dalloc(&p->err, 1))) return cleanup(rc);
    if ((rc = dalloc(&p->cnt, (size_t)3 * p->stride))) return cleanup(rc);
    if ((rc = dalloc(&p->fill, (size_t)3 * p->stride))) return cleanup(rc);
    if ((rc = dalloc(&p->seg, (size_t)3 * p->stride))) return cleanup(rc);
    if ((rc = dalloc(&p->tile_flag, max_tiles))) return cleanup(rc);
    if ((rc = dalloc(&p->bump, 2))) return cleanup(rc);
    if ((rc = p->reserve_events(1u << 16))) return cleanup(rc);
    if (model == EDLUT_LIF || model == EDLUT_PURKINJE || model == EDLUT_LIF_IS) {
        if ((rc = dalloc(&p->last_spike, p->stride))) return cleanup(rc);
        std::vector<int32_t> never(p->stride, -(1 << 30));
        if (hipMemcpy(p->last_spike, never.data(), sizeof(int32_t) * p->stride,
                      hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(EDLUT_E_HIP, "hipMemcpy failed"));
    }
    if (hipHostMalloc((void **)&p->h_spk, sizeof(int32_t) * n_neurons, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void **)&p->h_cnt, sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(EDLUT_E_HIP, "hipHostMalloc failed"));

    // initial state (VectorNeuronState_GPU_C_INTERFACE.cu:66-131), per group
    std::vector<float> st((size_t)p->nv * p->stride, 0.0f);
    for (int g = 0; g < n_groups; g++) {
        const float *gi = grouped ? &st_init[(size_t)g * p->nv] : init.data();
        const int64_t b = start[g];
        const int64_t e = g + 1 < n_groups ? start[g + 1] : n_neurons;
        for (int v = 0; v < p->nv; v++)
            for (int64_t i = b; i < e; i++) st[(size_t)v * p->stride + i] = gi[v];
    }
    p->init_state = st;
    bool ok = hipMemcpy(p->state, st.data(), st.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(p->exp_lut, expt.data(), expt.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(p->nmda_lut, nmda.data(), nmda.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              (refr.empty() || hipMemcpy(p->refr, refr.data(), refr.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemset(p->tile_count, 0, sizeof(int32_t) * padded_tiles((int)max_tiles)) == hipSuccess &&
              hipMemset(p->spk_cnt, 0, sizeof(int32_t)) == hipSuccess &&
              hipMemset(p->err, 0, sizeof(uint32_t)) == hipSuccess &&
              hipMemset(p->cnt, 0, sizeof(uint32_t) * 3 * p->stride) == hipSuccess &&
              hipMemset(p->fill, 0, sizeof(uint32_t) * 3 * p->stride) == hipSuccess &&
              hipMemset(p->tile_flag, 0, sizeof(uint32_t) * max_tiles) == hipSuccess &&
              hipMemset(p->seg, 0xff, sizeof(uint64_t) * 3 * p->stride) == hipSuccess &&
              hipMemset(p->bump, 0, 2 * sizeof(uint64_t)) == hipSuccess;
    if (!ok) return cleanup(fail(EDLUT_E_HIP, "device initialisation failed"));
    *out = p;
    return 0;
}

}  // namespace

extern "C" {

int edlut_population_create(int model, int integrator, int bdf_order, const float *params,
                            int n_params, float dt_s, int n_neurons, int device,
                            edlut_population **out) {
    const int32_t size = n_neurons;
    if (n_neurons <= 0) {
        if (out) *out = nullptr;
        return fail(EDLUT_E_ARG, "population size must be positive");
    }
    return create_population(model, integrator, bdf_order, params, n_params, dt_s, 1, &size, false,
                             device, out);
}

int edlut_population_create_groups(int model, int integrator, const float *params, int n_params,
                                   float dt_s, int n_groups, const int32_t *sizes, int device,
                                   edlut_population **out) {
    return create_population(model, integrator, 0, params, n_params, dt_s, n_groups, sizes, true,
                             device, out);
}

int edlut_population_groups(const edlut_population *p, int *n_groups, int32_t *start, int32_t *size) {
    if (!p || !n_groups) return fail(EDLUT_E_ARG, "null argument");
    *n_groups = (int)p->g_start.size();
    if (start) std::memcpy(start, p->g_start.data(), sizeof(int32_t) * p->g_start.size());
    if (size) std::memcpy(size, p->g_size.data(), sizeof(int32_t) * p->g_size.size());
    return 0;
}

int edlut_population_destroy(edlut_population *p) {
    if (!p) return 0;
    if (p->net) {
        // refused while the network is in a circuit or feeds projections
        int rc = edlut_network_destroy(p->net);
        if (rc) return rc;
    }
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    void *dev[] = {p->chan, p->bdf_c, p->bdf_d, p->bdf_prev, p->bdf_st, p->bdf_ws, p->bdf_ws_count,
                   p->bdf_bits, p->bdf_done,
                   p->state, p->last_spike, p->exp_lut, p->nmda_lut, p->refr, p->tile_count,
                   p->staging, p->spk_idx, p->spk_cnt, p->err, p->cnt, p->fill, p->seg,
                   p->tile_flag, p->bump, p->b_ent, p->e_tgt,
                   p->e_type, p->e_w, p->d_groups, p->d_chunk_group};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    if (p->h_spk) (void)hipHostFree(p->h_spk);
    if (p->h_cnt) (void)hipHostFree(p->h_cnt);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int edlut_population_layout(const edlut_population *p, int *n_vars, int *n_diff, int64_t *stride) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (n_vars) *n_vars = p->nv;
    if (n_diff) *n_diff = p->nd;
    if (stride) *stride = p->stride;
    return 0;
}

int edlut_population_enable_synapses(edlut_population *p, uint32_t mask) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (mask & ~15u) return fail(EDLUT_E_ARG, "unknown synapse type bits 0x%x", mask);
    p->mask |= mask;
    p->mp.gran.mask = p->mp.pkj.mask = p->mp.izh.mask = p->mp.lifis.mask = p->mask;   // models with a run-time mask
    return 0;
}

int edlut_population_push_inputs(edlut_population *p, const int32_t *target, const uint8_t *type,
                                 const float *weight, int n) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (n < 0 || (n > 0 && (!target || !type || !weight))) return fail(EDLUT_E_ARG, "bad event arrays");
    for (int i = 0; i < n; i++) {
        if (target[i] < 0 || target[i] >= p->n)
            return fail(EDLUT_E_ARG, "event %d: target %d out of range", i, target[i]);
        if (type[i] > 2) return fail(EDLUT_E_ARG, "event %d: synapse type %d does not take spikes", i, type[i]);
        if (!(p->mask & (1u << type[i])))
            return fail(EDLUT_E_ARG, "event %d: synapse type %d is not enabled", i, type[i]);
    }
    p->q_tgt.insert(p->q_tgt.end(), target, target + n);
    p->q_type.insert(p->q_type.end(), type, type + n);
    p->q_w.insert(p->q_w.end(), weight, weight + n);
    return 0;
}

int edlut_population_set_ext_current(edlut_population *p, const float *current) {
    if (!p || !current) return fail(EDLUT_E_ARG, "null argument");
    p->mask |= EDLUT_SYN_EXT_I;
    const int var = p->nd + 3;
    HIP_TRY(hipMemcpyAsync(p->state + (int64_t)var * p->stride, current, sizeof(float) * p->n,
                           hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return 0;
}

int edlut_population_update(edlut_population *p, int64_t step, int *n_spikes) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (step < 1 || step > (1ll << 30)) return fail(EDLUT_E_ARG, "update index out of range");
    HIP_TRY(hipSetDevice(p->device));
    // 1. ordered fold of the queued input spikes
    const int ne = (int)p->q_tgt.size();
    const int par = (int)(++p->upd_seq & 1u);
    int rc0 = 0;
    if (ne > 0) {
        if (ne > p->e_cap) {
            (void)hipFree(p->e_tgt);
            (void)hipFree(p->e_type);
            (void)hipFree(p->e_w);
            p->e_tgt = nullptr;
            p->e_type = nullptr;
            p->e_w = nullptr;
            int rc;
            int cap = ne < 4096 ? 4096 : ne * 2;
            if ((rc = dalloc(&p->e_tgt, cap)) || (rc = dalloc(&p->e_type, cap)) ||
                (rc = dalloc(&p->e_w, cap)))
                return rc;
            p->e_cap = cap;
        }
        HIP_TRY(hipMemcpyAsync(p->e_tgt, p->q_tgt.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipMemcpyAsync(p->e_type, p->q_type.data(), ne, hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipMemcpyAsync(p->e_w, p->q_w.data(), sizeof(float) * ne, hipMemcpyHostToDevice, p->stream));
        if ((rc0 = p->reserve_events((uint64_t)ne))) return rc0;
        PushEvents pe{p->e_tgt, p->e_type, p->e_w, ne, (int32_t)p->n, p->mask & 7u, p->err};
        int rc = launch_push_fold(p, par, ne, &pe);
        if (rc) return rc;
    }
    // 2. integrate, detect, compact
    UpdateArgs a = make_update_args(p, (int32_t)step);
    a.out = TileOut{p->tile_count, p->staging};
    a.fold = p->fold_target(par);
    UpdateFn fn = pick_update(p, false);
    if (!fn) return fail(EDLUT_E_ARG, "no update kernel for this model, integrator and synapse set");
    p->started = true;
    const int tiles = p->update_tiles();
    hipLaunchKernelGGL(fn, dim3(p->update_blocks()), dim3(kBlock), update_lds_bytes(p), p->stream, a);
    launch_compact(p->stream, p->tile_count, p->staging, tiles, kBlock * p->items, p->spk_idx, p->spk_cnt,
                   1, 0, nullptr, nullptr, nullptr, 0, p->err, StepRef{nullptr, (int32_t)step});
    HIP_TRY(hipGetLastError());
    // 3. hand the compacted spikes back (synchronous, like the reference)
    HIP_TRY(hipMemcpyAsync(p->h_cnt, p->spk_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    const int ns = *p->h_cnt;
    if (ns > 0) {
        HIP_TRY(hipMemcpyAsync(p->h_spk, p->spk_idx, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
    }
    p->q_tgt.clear();
    p->q_type.clear();
    p->q_w.clear();
    int rc = check_err(p);
    if (rc) return rc;
    p->last_n_spikes = ns;
    if (n_spikes) *n_spikes = ns;
    return 0;
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Allocation-free update: the input fold's event buffer (order keys and
// weights of the step's events, grouped by target) lives in this file-scope
// device array instead of the population's growable buffer.  Each device
// holds its own copy at its own address, looked up once per device.
constexpr int64_t kNoAllocMaxEvents = 1 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocMaxEvents * sizeof(FoldEntry)];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on `device` (the current device).
int scratch_ptr(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kScratchDevices) return fail(EDLUT_E_ARG, "device %d out of range", device);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(EDLUT_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

}  // namespace

extern "C" {

int edlut_population_update_noalloc(edlut_population *p, int64_t step, const int32_t *target, const uint8_t *type,
                                    const float *weight, int n_events, int32_t *spk_idx, int32_t *spk_cnt,
                                    void *stream) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (step < 1 || step > (1ll << 30)) return fail(EDLUT_E_ARG, "update index out of range");
    if (n_events < 0 || (n_events > 0 && (!target || !type || !weight))) return fail(EDLUT_E_ARG, "bad event arrays");
    if (n_events > kNoAllocMaxEvents)
        return fail(EDLUT_E_CAP, "%d input events exceed the allocation-free update's maximum of %lld", n_events,
                    (long long)kNoAllocMaxEvents);
    if (!spk_idx || !spk_cnt) return fail(EDLUT_E_ARG, "null spike output");
    if (!p->q_tgt.empty())
        return fail(EDLUT_E_STATE, "input events queued by edlut_population_push_inputs: use edlut_population_update");
    UpdateFn fn = pick_update(p, false);
    if (!fn) return fail(EDLUT_E_ARG, "no update kernel for this model, integrator and synapse set");
    HIP_TRY(hipSetDevice(p->device));
    void *base = nullptr;
    int rc = scratch_ptr(p->device, &base);
    if (rc) return rc;
    const hipStream_t st = stream ? (hipStream_t)stream : p->stream;
    // 1. ordered fold of the events (position in the arrays = delivery order)
    const int par = (int)(++p->upd_seq & 1u);
    FoldTarget t = p->fold_target(par);
    t.b_ent = (FoldEntry *)base;
    t.b_cap = (uint64_t)kNoAllocMaxEvents;
    if (n_events > 0) {
        const PushEvents pe{target, type, weight, n_events, (int32_t)p->n, p->mask & 7u, p->err};
        int nb = (n_events + kBlock - 1) / kBlock;
        if (nb > 4096) nb = 4096;
        hipLaunchKernelGGL((k_fold_push<kCount>), dim3(nb), dim3(kBlock), 0, st, t, pe);
        hipLaunchKernelGGL((k_fold_push<kScatter>), dim3(nb), dim3(kBlock), 0, st, t, pe);
    }
    // 2. integrate, detect, compact into the caller's spike list
    UpdateArgs a = make_update_args(p, (int32_t)step);
    a.out = TileOut{p->tile_count, p->staging};
    a.fold = t;
    p->started = true;
    const int tiles = p->update_tiles();
    hipLaunchKernelGGL(fn, dim3(p->update_blocks()), dim3(kBlock), update_lds_bytes(p), st, a);
    launch_compact(st, p->tile_count, p->staging, tiles, kBlock * p->items, spk_idx, spk_cnt, 1, 0, nullptr,
                   nullptr, nullptr, 0, p->err, StepRef{nullptr, (int32_t)step});
    HIP_TRY(hipGetLastError());
    return 0;
}

int edlut_population_get_spikes(const edlut_population *p, int32_t *idx, int cap) {
    if (!p || (!idx && p->last_n_spikes > 0)) return fail(EDLUT_E_ARG, "null argument");
    if (cap < p->last_n_spikes) return fail(EDLUT_E_CAP, "spike buffer too small (%d < %d)", cap, p->last_n_spikes);
    std::memcpy(idx, p->h_spk, sizeof(int32_t) * p->last_n_spikes);
    return 0;
}

int edlut_population_get_state(const edlut_population *p, int var, float *out) {
    if (!p || !out) return fail(EDLUT_E_ARG, "null argument");
    if (var < 0 || var >= p->nv) return fail(EDLUT_E_ARG, "state variable %d out of range", var);
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipMemcpyAsync(out, p->state + (int64_t)var * p->stride, sizeof(float) * p->n,
                           hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return 0;
}

int edlut_population_set_state(edlut_population *p, int var, const float *in) {
    if (!p || !in) return fail(EDLUT_E_ARG, "null argument");
    if (var < 0 || var >= p->nv) return fail(EDLUT_E_ARG, "state variable %d out of range", var);
    if (var == p->nd + 3) p->mask |= EDLUT_SYN_EXT_I;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipMemcpyAsync(p->state + (int64_t)var * p->stride, in, sizeof(float) * p->n,
                           hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return 0;
}

int edlut_population_set_newton_fp64(edlut_population *p, int on) {
    if (!p) return fail(EDLUT_E_ARG, "population is null");
    if (p->integ != EDLUT_BDF || p->model != EDLUT_HH)
        return fail(EDLUT_E_ARG, "the FP64 Newton step applies to HH populations under BDF");
    if (p->started) return fail(EDLUT_E_STATE, "the Newton precision is fixed once updates have run");
    if ((bool)on == p->newton_fp64) return 0;
    HIP_TRY(hipSetDevice(p->device));
    int rc = p->alloc_bdf_history(on ? sizeof(double) : sizeof(float));
    if (rc) return rc;
    p->newton_fp64 = on != 0;
    return 0;
}

int edlut_population_device_state(const edlut_population *p, float **state) {
    if (!p || !state) return fail(EDLUT_E_ARG, "null argument");
    *state = p->state;
    return 0;
}

int edlut_population_lut_checksums(const edlut_population *p, uint64_t *out3) {
    if (!p || !out3) return fail(EDLUT_E_ARG, "null argument");
    for (int i = 0; i < 3; i++) out3[i] = p->lut_sums[i];
    return 0;
}

// ---------------------------------------------------------------------------
int edlut_network_create(edlut_population *pop, int max_delay, edlut_network **out) {
    if (!pop || !out) return fail(EDLUT_E_ARG, "null argument");
    *out = nullptr;
    if (pop->net) return fail(EDLUT_E_STATE, "population already has a network");
    if (max_delay < 1 || max_delay > kMaxDelay)
        return fail(EDLUT_E_ARG, "max_delay must be in [1, %d]", kMaxDelay);
    HIP_TRY(hipSetDevice(pop->device));
    edlut_network *n = new edlut_network;
    n->pop = pop;
    n->max_delay = max_delay;
    n->ring = max_delay + 2;
    n->log_cap = 1ll << 26;
    int rc;
    if ((rc = dalloc(&n->own_ring_idx, (size_t)n->ring * pop->n)) ||
        (rc = dalloc(&n->own_ring_cnt, n->ring)) ||
        (rc = dalloc(&n->tile_count, padded_tiles((int)(pop->stride / kFlagTile)))) ||
        (rc = dalloc(&n->staging, (size_t)pop->stride)) ||
        (rc = dalloc(&n->log_step, n->log_cap)) || (rc = dalloc(&n->log_idx, n->log_cap)) ||
        (rc = dalloc(&n->log_count, 2)) || (rc = dalloc(&n->d_step, 1)) ||
        (rc = dalloc(&n->d_time, n->ring)) || (rc = dalloc(&n->d_time_p, n->ring)) ||
        (rc = dalloc(&n->d_rtab, 1)) || (rc = dalloc(&n->d_skip, 2)) || (rc = dalloc(&n->d_delay_s, max_delay + 1))) {
        delete n;
        return rc;
    }
    {
        // Network::RoundPropagationDelay (Network.cpp:1311-1317) of d steps
        std::vector<double> ds(max_delay + 1, 0.0);
        for (int d = 1; d <= max_delay; d++) {
            const double m = 0.0001;
            const double r = std::floor((double)d * (double)pop->dt_s * (1.0 / 0.0001) + m * 0.5) * m;
            ds[d] = r <= 0 ? m : r;
        }
        HIP_TRY(hipMemcpy(n->d_delay_s, ds.data(), sizeof(double) * ds.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(n->d_time, 0, sizeof(double) * n->ring));
        HIP_TRY(hipMemset(n->d_time_p, 0, sizeof(double) * n->ring));
        HIP_TRY(hipMemset(n->d_rtab, 0, sizeof(RankTab)));
        HIP_TRY(hipMemset(n->d_skip, 0, 2 * sizeof(uint32_t)));
    }
    HIP_TRY(hipMemset(n->own_ring_cnt, 0, sizeof(int32_t) * n->ring));
    HIP_TRY(hipMemset(n->log_count, 0, 2 * sizeof(int64_t)));
    HIP_TRY(hipMemset(n->tile_count, 0, sizeof(int32_t) * padded_tiles((int)(pop->stride / kFlagTile))));
    pop->net = n;
    *out = n;
    return 0;
}

int edlut_network_destroy(edlut_network *n) {
    if (!n) return 0;
    if (n->circuit) return fail(EDLUT_E_STATE, "the network belongs to a circuit: destroy the circuit first");
    if (n->links_out > 0)
        return fail(EDLUT_E_STATE, "%d projections read this network's spikes: destroy their networks first",
                    n->links_out);
    (void)hipSetDevice(n->pop->device);
    (void)hipStreamSynchronize(n->pop->stream);
    for (Source &s : n->sources) {
        // the rings of recurrent and projected sources belong to their populations' networks
        const bool own_ring = !s.self && !s.from;
        void *dev[] = {own_ring ? s.ring_idx : nullptr, own_ring ? s.ring_cnt : nullptr, s.row_ptr,
                       s.tgt_type, s.w, s.lcg, s.tile_count, s.staging, s.syn.pre, s.syn.post,
                       s.syn.last, s.in_ptr, s.in_syn, s.rates, s.train_ptr, s.train_idx,
                       s.log_step, s.log_idx, s.log_count};
        for (void *q : dev)
            if (q) (void)hipFree(q);
        if (s.from) s.from->links_out--;
    }
    if (n->graph) (void)hipGraphExecDestroy(n->graph);
    void *dev[] = {n->own_ring_idx, n->own_ring_cnt, n->tile_count, n->staging, n->log_step,
                   n->log_idx, n->log_count, n->d_step, n->d_time, n->d_time_p, n->d_rtab, n->d_skip, n->d_delay_s,
                   n->seg_off, n->clk, n->clk_next};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    delete n->eq;
    n->pop->net = nullptr;
    delete n;
    return 0;
}

static int add_synapses(edlut_network *n, Source &s, const int64_t *row_ptr,
                        const uint32_t *tgt_type, const float *w) {
    const size_t nrow = (size_t)s.n * n->max_delay + 1;
    if (!row_ptr) return fail(EDLUT_E_ARG, "row_ptr is null");
    if (row_ptr[0] != 0) return fail(EDLUT_E_ARG, "row_ptr[0] must be 0");
    for (size_t i = 1; i < nrow; i++)
        if (row_ptr[i] < row_ptr[i - 1]) return fail(EDLUT_E_ARG, "row_ptr must be non-decreasing");
    const int64_t nsyn = row_ptr[nrow - 1];
    if (nsyn >= (1ll << 40)) return fail(EDLUT_E_ARG, "too many synapses");
    if (nsyn > 0 && (!tgt_type || !w)) return fail(EDLUT_E_ARG, "synapse arrays are null");
    uint32_t types = 0;
    for (int64_t q = 0; q < nsyn; q++) {
        const uint32_t t = tgt_type[q] & 0x3fffffffu, ty = tgt_type[q] >> 30;
        if ((int64_t)t >= n->pop->n) return fail(EDLUT_E_ARG, "synapse %lld: target out of range", (long long)q);
        if (ty > 2) return fail(EDLUT_E_ARG, "synapse %lld: type %u does not take spikes", (long long)q, ty);
        types |= 1u << ty;
    }
    // CheckSynapseType (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:239-262)
    n->pop->mask |= types;
    if (nsyn >= (int64_t)(1ull << kOrdSynBits)) return fail(EDLUT_E_ARG, "too many synapses in one source");
    n->eq_stale = true;
    int rc;
    if ((rc = dalloc(&s.row_ptr, nrow)) || (rc = dalloc(&s.tgt_type, nsyn)) || (rc = dalloc(&s.w, nsyn)))
        return rc;
    HIP_TRY(hipMemcpy(s.row_ptr, row_ptr, sizeof(int64_t) * nrow, hipMemcpyHostToDevice));
    s.nsyn = nsyn;
    if (nsyn > 0) {
        HIP_TRY(hipMemcpy(s.tgt_type, tgt_type, sizeof(uint32_t) * nsyn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s.w, w, sizeof(float) * nsyn, hipMemcpyHostToDevice));
    }
    return 0;
}

int edlut_network_add_recurrent(edlut_network *n, const int64_t *row_ptr, const uint32_t *tgt_type,
                                const float *w) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (n->self_src >= 0) return fail(EDLUT_E_STATE, "recurrent synapses already added");
    if ((int)n->sources.size() >= kMaxSources) return fail(EDLUT_E_CAP, "too many sources");
    HIP_TRY(hipSetDevice(n->pop->device));
    Source s;
    s.self = true;
    s.n = n->pop->n;
    s.ring_idx = n->own_ring_idx;
    s.ring_cnt = n->own_ring_cnt;
    int rc = add_synapses(n, s, row_ptr, tgt_type, w);
    if (rc) return rc;
    n->sources.push_back(s);
    n->dirty = true;
    n->seg_dirty = true;
    n->self_src = (int)n->sources.size() - 1;
    n->total_syn += (uint64_t)row_ptr[(size_t)s.n * n->max_delay];
    return n->pop->reserve_events(n->total_syn + 1);
}

int edlut_network_add_poisson(edlut_network *n, int n_src, const uint32_t *lcg_state, float rate_hz,
                              const int64_t *row_ptr, const uint32_t *tgt_type, const float *w) {
    if (!n || !lcg_state) return fail(EDLUT_E_ARG, "null argument");
    if (n_src <= 0) return fail(EDLUT_E_ARG, "n_src must be positive");
    if ((int)n->sources.size() >= kMaxSources) return fail(EDLUT_E_CAP, "too many sources");
    if (!std::isfinite(rate_hz)) return fail(EDLUT_E_ARG, "rate must be finite");
    // PoissonGeneratorDeviceVector samples every 0.1 ms: exactly one of its
    // updates falls between two population updates only at that step
    if (n->pop->dt_s != 1e-4f)
        return fail(EDLUT_E_ARG, "Poisson devices sample every 0.1 ms: the population's step must be 0.1 ms");
    HIP_TRY(hipSetDevice(n->pop->device));
    Source s;
    s.self = false;
    s.n = n_src;
    s.rate = rate_hz;
    int rc = add_synapses(n, s, row_ptr, tgt_type, w);
    if (rc) return rc;
    const int tiles = (n_src + kBlock - 1) / kBlock;
    if ((rc = dalloc(&s.lcg, n_src)) || (rc = dalloc(&s.ring_idx, (size_t)n->ring * n_src)) ||
        (rc = dalloc(&s.ring_cnt, n->ring)) || (rc = dalloc(&s.tile_count, padded_tiles(tiles))) ||
        (rc = dalloc(&s.staging, (size_t)tiles * kBlock)))
        return rc;
    HIP_TRY(hipMemcpy(s.lcg, lcg_state, sizeof(uint32_t) * n_src, hipMemcpyHostToDevice));
    s.lcg0.assign(lcg_state, lcg_state + n_src);
    HIP_TRY(hipMemset(s.ring_cnt, 0, sizeof(int32_t) * n->ring));
    HIP_TRY(hipMemset(s.tile_count, 0, sizeof(int32_t) * padded_tiles(tiles)));
    n->sources.push_back(s);
    n->dirty = true;
    n->seg_dirty = true;
    n->total_syn += (uint64_t)row_ptr[(size_t)s.n * n->max_delay];
    return n->pop->reserve_events(n->total_syn + 1);
}

// Validates a spike schedule and uploads it into the train arrays of `s`
// (replacing earlier ones).
static int load_train(edlut_network *n, Source &s, int64_t n_steps, const int64_t *step_ptr,
                      const int32_t *idx) {
    if (n_steps > 0 && !step_ptr) return fail(EDLUT_E_ARG, "step_ptr is null");
    if (n_steps < 0 || n_steps >= (int64_t)INT32_MAX) return fail(EDLUT_E_ARG, "n_steps out of range");
    int64_t longest = 0;
    if (n_steps > 0) {
        if (step_ptr[0] != 0) return fail(EDLUT_E_ARG, "step_ptr[0] must be 0");
        for (int64_t k = 1; k <= n_steps; k++) {
            const int64_t b = step_ptr[k - 1], e = step_ptr[k];
            if (e < b) return fail(EDLUT_E_ARG, "step_ptr must be non-decreasing");
            if (e - b > s.n) return fail(EDLUT_E_ARG, "update %lld: more spikes than sources", (long long)k);
            if (e > b && !idx) return fail(EDLUT_E_ARG, "idx is null");
            for (int64_t j = b; j < e; j++) {
                if (idx[j] < 0 || idx[j] >= s.n)
                    return fail(EDLUT_E_ARG, "update %lld: source %d out of range", (long long)k, idx[j]);
                // one spike per source and update, in ascending order (the
                // delivery order and the fold's fixed segments rely on it)
                if (j > b && idx[j] <= idx[j - 1])
                    return fail(EDLUT_E_ARG, "update %lld: spike sources must ascend without repeats",
                                (long long)k);
            }
            longest = std::max(longest, e - b);
        }
    }
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    if (s.train_ptr) (void)hipFree(s.train_ptr);
    if (s.train_idx) (void)hipFree(s.train_idx);
    s.train_ptr = nullptr;
    s.train_idx = nullptr;
    s.train_steps = 0;
    const int64_t total = n_steps > 0 ? step_ptr[n_steps] : 0;
    int rc;
    if ((rc = dalloc(&s.train_ptr, (size_t)n_steps + 1)) ||
        (rc = dalloc(&s.train_idx, (size_t)std::max<int64_t>(total, 1))))
        return rc;
    if (n_steps > 0) {
        HIP_TRY(hipMemcpy(s.train_ptr, step_ptr, sizeof(int64_t) * (size_t)(n_steps + 1), hipMemcpyHostToDevice));
        if (total > 0) HIP_TRY(hipMemcpy(s.train_idx, idx, sizeof(int32_t) * (size_t)total, hipMemcpyHostToDevice));
    }
    s.train_steps = n_steps;
    s.train_blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (longest + kBlock - 1) / kBlock));
    s.h_train_ptr.assign(step_ptr, step_ptr + (n_steps > 0 ? n_steps + 1 : 0));
    s.h_train_idx.assign(idx, idx + total);
    if (s.h_train_ptr.empty()) s.h_train_ptr.push_back(0);
    train_records(s, (double)n->pop->dt_s, n_steps, s.h_train_ptr.data(), s.h_train_idx.data(), nullptr, nullptr);
    n->dirty = true;   // captured kernels carry the train's arrays
    return 0;
}

int edlut_network_add_spike_train(edlut_network *n, int n_src, int64_t n_steps, const int64_t *step_ptr,
                                  const int32_t *idx, const int64_t *row_ptr, const uint32_t *tgt_type,
                                  const float *w) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (n_src <= 0) return fail(EDLUT_E_ARG, "n_src must be positive");
    if ((int)n->sources.size() >= kMaxSources) return fail(EDLUT_E_CAP, "too many sources");
    HIP_TRY(hipSetDevice(n->pop->device));
    Source s;
    s.self = false;
    s.train = true;
    s.n = n_src;
    int rc = add_synapses(n, s, row_ptr, tgt_type, w);
    if (rc) return rc;
    if ((rc = load_train(n, s, n_steps, step_ptr, idx))) return rc;
    if ((rc = dalloc(&s.ring_idx, (size_t)n->ring * n_src)) || (rc = dalloc(&s.ring_cnt, n->ring))) return rc;
    HIP_TRY(hipMemset(s.ring_cnt, 0, sizeof(int32_t) * n->ring));
    n->sources.push_back(s);
    n->dirty = true;
    n->seg_dirty = true;
    n->total_syn += (uint64_t)row_ptr[(size_t)s.n * n->max_delay];
    return n->pop->reserve_events(n->total_syn + 1);
}

int edlut_network_set_spike_train(edlut_network *n, int source, int64_t n_steps, const int64_t *step_ptr,
                                  const int32_t *idx) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    Source &s = n->sources[(size_t)source];
    if (!s.train) return fail(EDLUT_E_ARG, "source %d is not a spike train", source);
    return load_train(n, s, n_steps, step_ptr, idx);
}

int edlut_network_add_projection(edlut_network *n, edlut_network *from, const int64_t *row_ptr,
                                 const uint32_t *tgt_type, const float *w) {
    if (!n || !from) return fail(EDLUT_E_ARG, "network is null");
    if (from == n) return fail(EDLUT_E_ARG, "a population's own synapses are added with add_recurrent");
    if ((int)n->sources.size() >= kMaxSources) return fail(EDLUT_E_CAP, "too many sources");
    if (from->pop->device != n->pop->device) return fail(EDLUT_E_ARG, "networks on different devices");
    if (from->max_delay != n->max_delay) return fail(EDLUT_E_ARG, "networks of a projection must share max_delay");
    if (from->pop->dt_s != n->pop->dt_s) return fail(EDLUT_E_ARG, "networks of a projection must share the step");
    HIP_TRY(hipSetDevice(n->pop->device));
    Source s;
    s.self = false;
    s.from = from;
    s.n = from->pop->n;
    s.ring_idx = from->own_ring_idx;
    s.ring_cnt = from->own_ring_cnt;
    int rc = add_synapses(n, s, row_ptr, tgt_type, w);
    if (rc) return rc;
    n->sources.push_back(s);
    from->links_out++;
    n->dirty = true;
    n->seg_dirty = true;
    n->total_syn += (uint64_t)row_ptr[(size_t)s.n * n->max_delay];
    return n->pop->reserve_events(n->total_syn + 1);
}

int edlut_network_set_poisson_rates(edlut_network *n, int source, const float *rates_hz) {
    if (!n || !rates_hz) return fail(EDLUT_E_ARG, "null argument");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    Source &s = n->sources[(size_t)source];
    if (s.self || s.from || s.train) return fail(EDLUT_E_ARG, "source %d is not a Poisson device", source);
    for (int64_t i = 0; i < s.n; i++)
        if (!std::isfinite(rates_hz[i])) return fail(EDLUT_E_ARG, "rate %lld is not finite", (long long)i);
    HIP_TRY(hipSetDevice(n->pop->device));
    int rc;
    if (!s.rates && (rc = dalloc(&s.rates, (size_t)s.n))) return rc;
    HIP_TRY(hipMemcpy(s.rates, rates_hz, sizeof(float) * s.n, hipMemcpyHostToDevice));
    n->dirty = true;
    return 0;
}

int edlut_network_set_source_model(edlut_network *n, int source, int key, int rank) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    Source &s = n->sources[(size_t)source];
    if (s.self || s.from) return fail(EDLUT_E_ARG, "source %d is a population's own spikes", source);
    if (key < 0 || key >= (1 << 24)) return fail(EDLUT_E_ARG, "model key out of range");
    s.eq_key = key;
    s.eq_rank = rank;
    n->eq_stale = true;
    return 0;
}

int edlut_network_set_model_rank(edlut_network *n, int rank) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    n->model_rank = rank;
    n->eq_stale = true;
    return 0;
}

int edlut_network_set_event_order(edlut_network *n, int on) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if ((n->eq_on != (on != 0))) {
        n->eq_on = on != 0;
        n->eq_stale = true;
        n->dirty = true;
    }
    return 0;
}

// The replayed event queue's heap as stored (1-based array order): time and
// the packed key (priority << 28 | kind << 24 | model << 16 | delay).
static int eq_dump(EqSim *S, hipStream_t st, double *t, uint32_t *key, int64_t cap, int64_t *size) {
    *size = 0;
    if (!S || !S->st) return 0;
    HIP_TRY(hipStreamSynchronize(st));
    int32_t n = 0;
    HIP_TRY(hipMemcpy(&n, S->st, sizeof(int32_t), hipMemcpyDeviceToHost));
    *size = n;
    if (cap < n) return fail(EDLUT_E_CAP, "the queue holds %d events", n);
    std::vector<EqItem> h((size_t)n + 1);
    HIP_TRY(hipMemcpy(h.data(), S->heap, sizeof(EqItem) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < n; i++) {
        t[i] = h[(size_t)i + 1].t;
        key[i] = h[(size_t)i + 1].a;
    }
    return 0;
}

int edlut_network_event_queue(edlut_network *n, double *t, uint32_t *key, int64_t cap, int64_t *size) {
    if (!n || !size) return fail(EDLUT_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(n->pop->device));
    return eq_dump(n->circuit ? nullptr : n->eq, n->pop->stream, t, key, cap, size);
}

int edlut_network_set_stop_time(edlut_network *n, double t) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    n->stop_next = t;
    return 0;
}

int edlut_network_set_spike_train_times(edlut_network *n, int source, int64_t n_spikes, const double *time,
                                        const int64_t *order, const int64_t *load_first) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    Source &s = n->sources[(size_t)source];
    if (!s.train) return fail(EDLUT_E_ARG, "source %d is not a spike train", source);
    if (n_spikes != (int64_t)s.h_train_idx.size())
        return fail(EDLUT_E_ARG, "the schedule has %lld spikes", (long long)s.h_train_idx.size());
    if (n_spikes > 0 && !time) return fail(EDLUT_E_ARG, "time is null");
    for (int64_t j = 0; j < n_spikes; j++)
        if (!std::isfinite(time[j])) return fail(EDLUT_E_ARG, "spike %lld: time is not finite", (long long)j);
    if (order) {
        std::vector<int64_t> o(order, order + n_spikes);
        std::sort(o.begin(), o.end());
        for (int64_t j = 1; j < n_spikes; j++)
            if (o[(size_t)j] == o[(size_t)j - 1]) return fail(EDLUT_E_ARG, "load order has repeats");
    }
    s.tr_t.clear();   // every record is rebuilt (those already queued keep their place)
    const size_t queued = s.tr_seq.size();
    std::vector<int64_t> seq = s.tr_seq;
    if (load_first)
        for (int64_t j = 0; j < n_spikes; j++)
            if (load_first[j] < 0) return fail(EDLUT_E_ARG, "spike %lld: load update is negative", (long long)j);
    train_records(s, (double)n->pop->dt_s, s.train_steps, s.h_train_ptr.data(), s.h_train_idx.data(), time, order,
                  load_first);
    for (size_t q = 1; q < s.tr_load.size(); q++)
        if (s.tr_load[q] < s.tr_load[q - 1])
            return fail(EDLUT_E_ARG, "input spikes must be loaded in the order of their load updates");
    // records already queued keep their sequence numbers
    for (size_t q = 0; q < queued && q < s.tr_seq.size(); q++) s.tr_seq[q] = seq[q];
    return 0;
}

int edlut_network_set_external(edlut_network *n, int n_thr, const uint32_t *thresholds, float w_ext,
                               uint64_t seed) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (n_thr < 0 || n_thr > 16) return fail(EDLUT_E_ARG, "at most 16 thresholds");
    if (n_thr > 0 && !thresholds) return fail(EDLUT_E_ARG, "thresholds are null");
    for (int j = 1; j < n_thr; j++)
        if (thresholds[j] > thresholds[j - 1]) return fail(EDLUT_E_ARG, "thresholds must not increase");
    n->has_ext = n_thr > 0;
    std::memset(&n->ext, 0, sizeof(n->ext));
    n->ext.seed = seed;
    n->ext.n_thr = n_thr;
    n->ext.w = w_ext;
    for (int j = 0; j < n_thr; j++) n->ext.thr[j] = thresholds[j];
    if (n->has_ext) n->pop->mask |= EDLUT_SYN_EXC;
    n->dirty = true;
    return 0;
}

// Enqueues every kernel of one network step on the population's stream.
// fold_done: the fold of update k's events already ran (fused into the
// previous step's compaction).  Returns through *fused whether this step's
// compaction also folded the events of update k + 1.
// The fold target of a network, with the rank tables of the event-queue order
// when the network runs in that order.
static FoldTarget net_fold_target(const edlut_network *n) {
    FoldTarget ft = n->pop->fold_target(0, true);
    ft.seg_off = n->seg_off;
    ft.rtab = (n->eq_on && !n->pop->grouped && !n->sources.empty() && n->eq_rank[0]) ? n->d_rtab : nullptr;
    return ft;
}

// eq: the replayed event queue to advance after this step (a network that
// runs alone), or null.
static int enqueue_step(edlut_network *n, StepRef step, UpdateFn fn, const RingEvents &re0,
                        bool fold_done, bool *fused, const EqSim *eq = nullptr) {
    edlut_population *p = n->pop;
    if (n->plastic)
        hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, p->stream, step, n->d_time, n->d_time_p, n->ring,
                           (double)p->dt_s, poisson_step_s());
    // Poisson devices fire at update k (their spikes feed later updates).
    for (Source &s : n->sources) {
        if (s.self || s.from) continue;
        if (s.train) {
            TrainArgs ta;
            ta.ptr = s.train_ptr;
            ta.idx = s.train_idx;
            ta.n_steps = s.train_steps;
            ta.ring_idx = s.ring_idx;
            ta.ring_cnt = s.ring_cnt;
            ta.n_src = s.n;
            ta.ring = n->ring;
            ta.step = step;
            hipLaunchKernelGGL(k_train, dim3(s.train_blocks), dim3(kBlock), 0, p->stream, ta);
            continue;
        }
        PoissonArgs pa;
        std::memset(&pa, 0, sizeof(pa));
        pa.lcg = s.lcg;
        pa.n = (int)s.n;
        pa.n_tiles = (int)((s.n + kBlock - 1) / kBlock);
        pa.rate = s.rate;
        pa.rates = s.rates;
        pa.step_s = poisson_step_s();
        pa.out = TileOut{s.tile_count, s.staging};
        hipLaunchKernelGGL(k_poisson, dim3(pa.n_tiles), dim3(kBlock), 0, p->stream, pa);
        launch_compact(p->stream, s.tile_count, s.staging, pa.n_tiles, kBlock, s.ring_idx, s.ring_cnt,
                       n->ring, s.n, s.log_step, s.log_idx, s.log_count, s.log_cap, p->err, step);
    }
    // Deliver the events consumed by update k into their keys' segments.
    FoldTarget ft = net_fold_target(n);
    const bool ring = re0.n_sources > 0;
    if (ring && !fold_done) {
        RingEvents re = re0;
        re.step = step;
        int rc = launch_ring_fold(ft, p->stream, re, n->fold_lanes, n->fold_blocks);
        if (rc) return rc;
    }
    UpdateArgs a = make_update_args(p, 0);
    a.step = step;
    a.ext = n->ext;
    a.out = TileOut{n->tile_count, n->staging};
    a.fold = ft;
    const int tiles = p->update_tiles();
    if (n->clk) {
        // timed run: every workgroup stamps its start and end
        a.clk = n->clk;
        a.clk_first = (int32_t)n->clk_first;
        a.clk_count = (int32_t)n->clk_count;
    }
    hipLaunchKernelGGL(fn, dim3(p->update_blocks()), dim3(kBlock), update_lds_bytes(p), p->stream, a);
    CompactArgs c = compact_args(n->tile_count, n->staging, tiles, kBlock * p->items, n->own_ring_idx,
                                 n->own_ring_cnt, n->ring, p->n, n->log_cap ? n->log_step : nullptr,
                                 n->log_cap ? n->log_idx : nullptr, n->log_count, n->log_cap, p->err, step);
    if (n->clk) {
        c.clk_next = n->clk_next;
        c.clk_first = (int32_t)n->clk_first;
        c.clk_count = (int32_t)n->clk_count;
    }
    const int nbc = compact_blocks(tiles);
    // STDP orders the post-synaptic learning of update k before the arrivals
    // of update k + 1, so plastic networks fold in their own step.
    *fused = ring && !n->plastic;
    // the event-queue replay of this update's window runs in the compaction
    // kernel's last workgroup when its sizes fit there
    const bool eq_fused = *fused && eq && eq->max_delay <= kEqFusedDelays;
    if (*fused) {
        RingEvents rn = re0;
        rn.step = StepRef{step.ptr, step.imm + 1};
        if (eq_fused) {
            CompactFoldEqFn f = pick_compact_fold_eq(n->fold_lanes);
            if (!f) return fail(EDLUT_E_ARG, "fold lanes must be 8, 16, 32 or 64");
            hipLaunchKernelGGL(f, dim3(nbc + n->fold_blocks + 1), dim3(kBlock), 0, p->stream, c, ft, rn, nbc,
                               eq_args(eq, step, 0, 0.0));
        } else {
            CompactFoldFn f = pick_compact_fold(n->fold_lanes);
            if (!f) return fail(EDLUT_E_ARG, "fold lanes must be 8, 16, 32 or 64");
            hipLaunchKernelGGL(f, dim3(nbc + n->fold_blocks), dim3(kBlock), 0, p->stream, c, ft, rn, nbc);
        }
    } else {
        hipLaunchKernelGGL(k_compact, dim3(nbc), dim3(kBlock), 0, p->stream, c);
    }
    // post-synaptic learning of the neurons that fired at this update
    for (Source &s : n->sources) {
        if (!s.stdp) continue;
        StdpPostArgs pa;
        pa.spk_idx = n->own_ring_idx;
        pa.spk_cnt = n->own_ring_cnt;
        pa.slot_len = p->n;
        pa.ring = n->ring;
        pa.step = step;
        pa.in_ptr = s.in_ptr;
        pa.in_syn = s.in_syn;
        pa.chunks = (int)((s.in_max + kWave - 1) / kWave);
        if (pa.chunks < 1) pa.chunks = 1;
        pa.w = s.w;
        pa.syn = s.syn;
        pa.time_ring = n->d_time;
        pa.exp_lut = p->exp_lut;
        hipLaunchKernelGGL(k_stdp_post, dim3(64), dim3(kBlock), 0, p->stream, pa);
    }
    if (eq && !eq_fused) hipLaunchKernelGGL(k_eventq, dim3(1), dim3(kBlock), 0, p->stream, eq_args(eq, step, 0, 0.0));
    return 0;
}

static RingEvents ring_events(edlut_network *n) {
    RingEvents re;
    std::memset(&re, 0, sizeof(re));
    re.n_sources = (int)n->sources.size();
    re.max_delay = n->max_delay;
    re.ring = n->ring;
    re.time_ring = n->d_time;
    re.time_ring_p = n->d_time_p;
    re.delay_s = n->d_delay_s;
    re.rtab = (n->eq_on && !n->pop->grouped && !n->sources.empty() && n->eq_rank[0]) ? n->d_rtab : nullptr;
    re.skip = n->d_skip;
    re.rank_hi = 0xffffffffu;
    re.exp_lut = n->pop->exp_lut;
    for (int q = 0; q < re.n_sources; q++) {
        const Source &s = n->sources[q];
        re.src[q] = RingSource{s.ring_idx, s.ring_cnt, s.row_ptr, s.tgt_type, s.w, s.n,
                               s.stdp ? 1 : 0, s.syn, s.train ? 1 : 0, nullptr, nullptr, 1,
                               (!s.self && !s.from && !s.train) ? 1 : 0};
        if (n->eq_on && !n->pop->grouped && n->eq_rank[q]) {
            re.src[q].bnd = n->eq_bnd[q];
            re.src[q].bnd_on = n->eq_bnd_on[q];
            re.src[q].cmax = n->eq_cmax[q];
        }
    }
    return re;
}

static constexpr int kGraphChunk = 16;

// The replayed event queue of a network that runs alone, built at update 1
// (and whenever a run restarts there).
static int eq_prepare_net(edlut_network *n, int64_t first) {
    if (!n->eq) n->eq = new EqSim;
    if (first == 1 || n->eq->em.empty() || n->eq_stale) {
        if (first != 1) return fail(EDLUT_E_STATE, "sources changed: the event queue restarts at update 1 (reset)");
        int rc = eq_build(n->eq, {n}, n->pop->err, n->pop->stream);
        if (rc) return rc;
        n->eq_stale = false;
        n->dirty = true;
    }
    return 0;
}

static EqSim *net_eq(edlut_network *n) { return (n->eq_on && !n->pop->grouped && !n->circuit) ? n->eq : nullptr; }

// Captures kGraphChunk steps; step i of the chunk reads its update index as
// d_step + i, and the chunk's last node advances d_step by kGraphChunk.
static int build_graph(edlut_network *n, UpdateFn fn) {
    edlut_population *p = n->pop;
    if (n->graph) {
        (void)hipGraphExecDestroy(n->graph);
        n->graph = nullptr;
    }
    const RingEvents re = ring_events(n);
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    // the fold of each captured step ran in the step before it when fused
    // (the run launches the fold of the chunk's first step when needed)
    const bool will_fuse = re.n_sources > 0 && !n->plastic;
    bool fused = will_fuse;
    for (int i = 0; i < kGraphChunk && rc == 0; i++)
        rc = enqueue_step(n, StepRef{n->d_step, i}, fn, re, fused, &fused, net_eq(n));
    n->graph_fused = will_fuse;
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, p->stream, n->d_step, (int32_t)kGraphChunk);
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(EDLUT_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&n->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(EDLUT_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    n->graph_chunk = kGraphChunk;
    n->graph_fn = fn;
    n->dirty = false;
    return 0;
}

// Fixed fold segments: per (type, target) key, the count of synapses of all
// sources ending there, as an exclusive prefix (events of one step never
// exceed it).
static int build_segments(edlut_network *n) {
    edlut_population *p = n->pop;
    const int64_t keys = 3 * p->stride;
    const int64_t nch = (keys + kScanChunk - 1) / kScanChunk;
    uint32_t *deg = nullptr;
    int64_t *sums = nullptr;
    int rc;
    if (!n->seg_off && (rc = dalloc(&n->seg_off, keys + 1))) return rc;
    if ((rc = dalloc(&deg, keys)) || (rc = dalloc(&sums, nch + 1))) {
        if (deg) (void)hipFree(deg);
        return rc;
    }
    hipError_t e = hipMemsetAsync(deg, 0, sizeof(uint32_t) * keys, p->stream);
    for (const Source &s : n->sources) {
        if (s.nsyn == 0) continue;
        int64_t nb = (s.nsyn + kBlock - 1) / kBlock;
        if (nb > 65536) nb = 65536;
        hipLaunchKernelGGL(k_key_degree, dim3((unsigned)nb), dim3(kBlock), 0, p->stream, s.tgt_type, s.nsyn,
                           p->stride, deg);
    }
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nch), dim3(kBlock), 0, p->stream, deg, keys, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, p->stream, sums, nch);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nch), dim3(kBlock), 0, p->stream, deg, keys,
                       (const int64_t *)sums, (int64_t *)n->seg_off);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(deg);
    (void)hipFree(sums);
    if (e != hipSuccess) return fail(EDLUT_E_HIP, "fold segment set-up failed: %s", hipGetErrorString(e));
    n->seg_dirty = false;
    return 0;
}

// Everything a run needs before its first step: stale folds dropped, rings
// cleared when the run does not continue the previous one, segments built,
// the timing window set.  Returns the update kernel through *fn.
static int run_prepare(edlut_network *n, int64_t first_step, int n_steps, UpdateFn *fn) {
    edlut_population *p = n->pop;
    if (!p->q_tgt.empty()) return fail(EDLUT_E_STATE, "queued host inputs are not used in network mode");
    if (n->fold_ready >= 0 && (first_step != n->fold_ready || n->seg_dirty)) {
        // events folded ahead for an update that will not run next (or into
        // segments about to change): drop them, the fold is redone
        HIP_TRY(hipMemsetAsync(p->fill, 0, sizeof(uint32_t) * 3 * p->stride, p->stream));
        HIP_TRY(hipMemsetAsync(p->tile_flag, 0, sizeof(uint32_t) * (p->stride / kFlagTile), p->stream));
        n->fold_ready = -1;
    }
    if (first_step != n->last_step + 1) {
        // Rings hold the spikes of the steps just before `first_step` only
        // when the run continues the previous one.
        for (Source &s : n->sources)
            if (!s.from) HIP_TRY(hipMemsetAsync(s.ring_cnt, 0, sizeof(int32_t) * n->ring, p->stream));
        HIP_TRY(hipMemsetAsync(n->own_ring_cnt, 0, sizeof(int32_t) * n->ring, p->stream));
        int64_t lc[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(lc, n->log_count, sizeof(lc), hipMemcpyDeviceToHost, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        const int64_t total = lc[(n->last_step + 1) & 1];
        lc[0] = lc[1] = total;
        HIP_TRY(hipMemcpyAsync(n->log_count, lc, sizeof(lc), hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        for (Source &s : n->sources) {
            if (!s.log_count) continue;
            int64_t sc[2] = {0, 0};
            HIP_TRY(hipMemcpy(sc, s.log_count, sizeof(sc), hipMemcpyDeviceToHost));
            sc[0] = sc[1] = sc[(n->last_step + 1) & 1];
            HIP_TRY(hipMemcpy(s.log_count, sc, sizeof(sc), hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(hipMemsetAsync(p->bump, 0, 2 * sizeof(uint64_t), p->stream));
    if (n->plastic) {
        // update time of the step before the run, in its ring slot
        const double t0 = n->time_of(first_step - 1);
        double tp = 0.0;
        for (int64_t k = 1; k < first_step; k++) tp += poisson_step_s();
        HIP_TRY(hipMemcpyAsync(n->d_time + (first_step - 1) % n->ring, &t0, sizeof(double),
                               hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipMemcpyAsync(n->d_time_p + (first_step - 1) % n->ring, &tp, sizeof(double),
                               hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
    }
    *fn = pick_update(p, n->has_ext);
    if (!*fn) return fail(EDLUT_E_ARG, "no update kernel for this model, integrator and synapse set");
    if (n->seg_dirty) {
        int rc = build_segments(n);
        if (rc) return rc;
        n->dirty = true;
    }
    p->started = true;
    // timed runs: the update kernels of the first clk_cap steps stamp their
    // workgroups' start and end (captured chunks included)
    n->timed = 0;
    if (n->clk) {
        n->clk_first = first_step;
        n->clk_count = n_steps < n->clk_cap ? n_steps : n->clk_cap;
        n->clk_blocks = p->update_blocks();
        n->dirty = true;   // captured kernel arguments carry the timing window
        HIP_TRY(hipMemsetAsync(n->clk, 0, sizeof(unsigned long long) * 2 * n->clk_cap * n->clk_tiles_max,
                               p->stream));
        HIP_TRY(hipMemsetAsync(n->clk_next, 0xff, sizeof(unsigned long long) * n->clk_cap, p->stream));
    }
    return 0;
}

// Before a chunk of captured steps starting at update k: when the captured
// steps expect their events folded already, fold update k's.
static int fold_ahead(edlut_network *n, const RingEvents &re, int64_t k) {
    if (!n->graph_fused || n->fold_ready == k) return 0;
    RingEvents rk = re;
    rk.step = StepRef{nullptr, (int32_t)k};
    return launch_ring_fold(net_fold_target(n), n->pop->stream, rk, n->fold_lanes, n->fold_blocks);
}

// A run whose stop event leaves inside the window after its last update
// (Simulation_API::RunSimulation with an end time past that update) delivers,
// in the reference, the window's groups that come before the stop event.
// Here: the fold of the next update restricted to the groups ranked below
// r_stop, added to the conductances now; the next update's fold skips them
// (d_skip).
static int stop_delivery(edlut_network *n, int64_t next, uint32_t r_stop) {
    edlut_population *p = n->pop;
    if (n->sources.empty()) return 0;
    if (n->fold_ready >= 0) {
        // events folded ahead for the next update: folded again, in two parts
        HIP_TRY(hipMemsetAsync(p->fill, 0, sizeof(uint32_t) * 3 * p->stride, p->stream));
        HIP_TRY(hipMemsetAsync(p->tile_flag, 0, sizeof(uint32_t) * (p->stride / kFlagTile), p->stream));
        n->fold_ready = -1;
    }
    const uint32_t skip[2] = {(uint32_t)next, r_stop};
    RingEvents re = ring_events(n);
    re.step = StepRef{nullptr, (int32_t)next};
    re.skip = nullptr;
    re.rank_hi = r_stop;
    FoldTarget ft = net_fold_target(n);
    int rc = launch_ring_fold(ft, p->stream, re, n->fold_lanes, n->fold_blocks);
    if (rc) return rc;
    const int nb = (int)std::min<int64_t>(4096, ((int64_t)p->n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_apply_fold, dim3(nb), dim3(kBlock), 0, p->stream, ft, (int32_t)p->n, (int)(next & 1));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(n->d_skip, skip, sizeof(skip), hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return 0;
}

// After a run with an explicit stop time: the rank counter when its stop
// event left (0 or less: no group came before it).
static int stop_rank(EqSim *S, hipStream_t st, uint32_t *r_stop) {
    int32_t v = -1;
    HIP_TRY(hipMemcpyAsync(&v, S->st + 2, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *r_stop = v > 0 ? (uint32_t)v : 0u;
    return 0;
}

static void run_finish(edlut_network *n, int64_t first_step, int n_steps) {
    if (n->clk) n->timed = n->clk_count;
    n->last_step = first_step + n_steps - 1;
}

int edlut_network_run(edlut_network *n, int64_t first_step, int n_steps) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (n->circuit) return fail(EDLUT_E_STATE, "the network belongs to a circuit: run the circuit");
    if (first_step < 1 || n_steps < 0 || first_step + n_steps > (1ll << 30))
        return fail(EDLUT_E_ARG, "step range out of bounds");
    for (const Source &s : n->sources)
        if (s.from) return fail(EDLUT_E_STATE, "a network with projections from other networks runs in a circuit");
    edlut_population *p = n->pop;
    HIP_TRY(hipSetDevice(p->device));
    UpdateFn fn = nullptr;
    int rc = run_prepare(n, first_step, n_steps, &fn);
    if (rc) return rc;
    const double stop = n->stop_next;
    n->stop_next = -1.0;
    if (n->eq_on && !p->grouped) {
        if ((rc = eq_prepare_net(n, first_step)) || (rc = eq_resume(n->eq, first_step, n_steps, stop, p->stream)))
            return rc;
    }
    const RingEvents re = ring_events(n);
    int64_t k = first_step;
    if (n->use_graph && n_steps >= kGraphChunk) {
        if (n->dirty || !n->graph || n->graph_fn != fn) {
            if ((rc = build_graph(n, fn))) return rc;
        }
        if ((rc = fold_ahead(n, re, k))) return rc;
        int32_t k32 = (int32_t)k;
        HIP_TRY(hipMemcpyAsync(n->d_step, &k32, sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        while (k + kGraphChunk <= first_step + n_steps) {
            HIP_TRY(hipGraphLaunch(n->graph, p->stream));
            k += kGraphChunk;
        }
        n->fold_ready = n->graph_fused ? k : -1;
    }
    for (; k < first_step + n_steps; k++) {
        bool fused = false;
        if ((rc = enqueue_step(n, StepRef{nullptr, (int32_t)k}, fn, re, n->fold_ready == k, &fused, net_eq(n))))
            return rc;
        n->fold_ready = fused ? k + 1 : -1;
    }
    HIP_TRY(hipGetLastError());
    run_finish(n, first_step, n_steps);
    if (n->eq_on && !p->grouped && stop >= 0) {
        uint32_t r_stop = 0;
        if ((rc = stop_rank(n->eq, p->stream, &r_stop))) return rc;
        if (r_stop > 0 && (rc = stop_delivery(n, first_step + n_steps, r_stop))) return rc;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Circuits: several networks advanced in lockstep on one stream (update k of
// every network, in the order given, before update k + 1 of any), so that
// projections between them deliver in the integer order of a single network.
struct edlut_circuit {
    std::vector<edlut_network *> nets;
    std::vector<hipStream_t> own_streams;   // the populations' streams, restored on destroy
    hipStream_t stream = nullptr;
    int device = 0;
    int32_t *d_step = nullptr;
    hipGraphExec_t graph = nullptr;
    std::vector<UpdateFn> graph_fns;
    // delivery in the event queue's order: one replayed queue for all members
    EqSim *eq = nullptr;
    bool eq_on = true;
    bool eq_stale = true;
    double stop_next = -1.0;
};

int edlut_circuit_create(edlut_network *const *nets, int n_nets, edlut_circuit **out) {
    if (!nets || !out || n_nets <= 0) return fail(EDLUT_E_ARG, "no networks");
    *out = nullptr;
    for (int i = 0; i < n_nets; i++) {
        if (!nets[i]) return fail(EDLUT_E_ARG, "network %d is null", i);
        if (nets[i]->circuit) return fail(EDLUT_E_STATE, "network %d already belongs to a circuit", i);
        if (nets[i]->pop->device != nets[0]->pop->device) return fail(EDLUT_E_ARG, "networks on different devices");
        for (int j = 0; j < i; j++)
            if (nets[j] == nets[i]) return fail(EDLUT_E_ARG, "network %d appears twice", i);
    }
    // every projection must come from a member of the circuit
    for (int i = 0; i < n_nets; i++)
        for (const Source &s : nets[i]->sources) {
            if (!s.from) continue;
            bool member = false;
            for (int j = 0; j < n_nets; j++) member = member || nets[j] == s.from;
            if (!member) return fail(EDLUT_E_ARG, "network %d has a projection from a network outside the circuit", i);
        }
    edlut_circuit *c = new edlut_circuit;
    c->device = nets[0]->pop->device;
    HIP_TRY(hipSetDevice(c->device));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(EDLUT_E_HIP, "hipStreamCreate failed");
    }
    int rc = dalloc(&c->d_step, 1);
    if (rc) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    for (int i = 0; i < n_nets; i++) {
        edlut_network *n = nets[i];
        (void)hipStreamSynchronize(n->pop->stream);
        c->nets.push_back(n);
        c->own_streams.push_back(n->pop->stream);
        n->pop->stream = c->stream;
        n->circuit = c;
        n->dirty = true;
    }
    *out = c;
    return 0;
}

int edlut_circuit_destroy(edlut_circuit *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->nets.size(); i++) {
        c->nets[i]->pop->stream = c->own_streams[i];
        c->nets[i]->circuit = nullptr;
        c->nets[i]->dirty = true;
    }
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->d_step) (void)hipFree(c->d_step);
    delete c->eq;
    (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

// Captures kGraphChunk lockstep steps of every member network.
static int build_circuit_graph(edlut_circuit *c, const std::vector<UpdateFn> &fns,
                               const std::vector<RingEvents> &res, const EqSim *eq) {
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    const size_t m = c->nets.size();
    std::vector<bool> fused(m);
    for (size_t j = 0; j < m; j++) fused[j] = res[j].n_sources > 0 && !c->nets[j]->plastic;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < kGraphChunk && rc == 0; i++) {
        for (size_t j = 0; j < m && rc == 0; j++) {
            bool f = fused[j];
            rc = enqueue_step(c->nets[j], StepRef{c->d_step, i}, fns[j], res[j], f, &f);
            fused[j] = f;
        }
        if (eq) hipLaunchKernelGGL(k_eventq, dim3(1), dim3(kBlock), 0, c->stream, eq_args(eq, StepRef{c->d_step, i}, 0, 0.0));
    }
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, c->stream, c->d_step, (int32_t)kGraphChunk);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(EDLUT_E_HIP, "graph capture failed");
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(EDLUT_E_HIP, "graph instantiation failed");
    for (size_t j = 0; j < m; j++) {
        edlut_network *n = c->nets[j];
        n->graph_fused = res[j].n_sources > 0 && !n->plastic;
        // a network's own graph is stale once it has run inside the circuit
        if (n->graph) (void)hipGraphExecDestroy(n->graph);
        n->graph = nullptr;
        n->dirty = false;
    }
    c->graph_fns = fns;
    return 0;
}

int edlut_circuit_run(edlut_circuit *c, int64_t first_step, int n_steps) {
    if (!c) return fail(EDLUT_E_ARG, "circuit is null");
    if (first_step < 1 || n_steps < 0 || first_step + n_steps > (1ll << 30))
        return fail(EDLUT_E_ARG, "step range out of bounds");
    HIP_TRY(hipSetDevice(c->device));
    const size_t m = c->nets.size();
    std::vector<UpdateFn> fns(m);
    std::vector<RingEvents> res(m);
    bool use_graph = true, rebuild = !c->graph;
    for (size_t j = 0; j < m; j++) {
        edlut_network *n = c->nets[j];
        int rc = run_prepare(n, first_step, n_steps, &fns[j]);
        if (rc) return rc;
        res[j] = ring_events(n);
        use_graph = use_graph && n->use_graph;
        rebuild = rebuild || n->dirty;
    }
    rebuild = rebuild || fns != c->graph_fns;
    int64_t k = first_step;
    int rc = 0;
    bool grouped = false;
    for (edlut_network *n : c->nets) grouped = grouped || n->pop->grouped;
    bool all_on = true;
    for (edlut_network *n : c->nets) all_on = all_on && n->eq_on;
    if (all_on != c->eq_on) {
        c->eq_on = all_on;
        c->eq_stale = true;
        rebuild = true;
    }
    const bool eq_on = c->eq_on && !grouped;
    if (eq_on) {
        if (!c->eq) c->eq = new EqSim;
        bool stale = c->eq_stale || c->eq->em.empty();
        for (edlut_network *n : c->nets) stale = stale || n->eq_stale;
        if (first_step == 1 || stale) {
            if (first_step != 1) return fail(EDLUT_E_STATE, "sources changed: the event queue restarts at update 1 (reset)");
            if ((rc = eq_build(c->eq, c->nets, c->nets[0]->pop->err, c->stream))) return rc;
            c->eq_stale = false;
            for (edlut_network *n : c->nets) n->eq_stale = false;
            rebuild = true;
            for (size_t j = 0; j < m; j++) res[j] = ring_events(c->nets[j]);
        }
        if ((rc = eq_resume(c->eq, first_step, n_steps, c->stop_next, c->stream))) return rc;
    }
    const double stop = c->stop_next;
    c->stop_next = -1.0;
    const EqSim *eq = eq_on ? c->eq : nullptr;
    if (use_graph && n_steps >= kGraphChunk) {
        if (rebuild && (rc = build_circuit_graph(c, fns, res, eq))) return rc;
        for (size_t j = 0; j < m; j++)
            if ((rc = fold_ahead(c->nets[j], res[j], k))) return rc;
        int32_t k32 = (int32_t)k;
        HIP_TRY(hipMemcpyAsync(c->d_step, &k32, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        while (k + kGraphChunk <= first_step + n_steps) {
            HIP_TRY(hipGraphLaunch(c->graph, c->stream));
            k += kGraphChunk;
        }
        for (size_t j = 0; j < m; j++) c->nets[j]->fold_ready = c->nets[j]->graph_fused ? k : -1;
    }
    for (; k < first_step + n_steps; k++) {
        for (size_t j = 0; j < m; j++) {
            edlut_network *n = c->nets[j];
            bool fused = false;
            if ((rc = enqueue_step(n, StepRef{nullptr, (int32_t)k}, fns[j], res[j], n->fold_ready == k, &fused)))
                return rc;
            n->fold_ready = fused ? k + 1 : -1;
        }
        if (eq) hipLaunchKernelGGL(k_eventq, dim3(1), dim3(kBlock), 0, c->stream, eq_args(eq, StepRef{nullptr, (int32_t)k}, 0, 0.0));
    }
    HIP_TRY(hipGetLastError());
    for (size_t j = 0; j < m; j++) run_finish(c->nets[j], first_step, n_steps);
    if (eq && stop >= 0) {
        uint32_t r_stop = 0;
        if ((rc = stop_rank(c->eq, c->stream, &r_stop))) return rc;
        if (r_stop > 0)
            for (edlut_network *n : c->nets)
                if ((rc = stop_delivery(n, first_step + n_steps, r_stop))) return rc;
    }
    return 0;
}

int edlut_circuit_event_queue(edlut_circuit *c, double *t, uint32_t *key, int64_t cap, int64_t *size) {
    if (!c || !size) return fail(EDLUT_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    return eq_dump(c->eq, c->stream, t, key, cap, size);
}

int edlut_circuit_set_stop_time(edlut_circuit *c, double t) {
    if (!c) return fail(EDLUT_E_ARG, "circuit is null");
    c->stop_next = t;
    return 0;
}

int edlut_circuit_synchronize(edlut_circuit *c) {
    if (!c) return fail(EDLUT_E_ARG, "circuit is null");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (edlut_network *n : c->nets) {
        int rc = check_err(n->pop);
        if (rc) return rc;
    }
    return 0;
}

int edlut_network_set_graph(edlut_network *n, int on) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    n->use_graph = on != 0;
    return 0;
}

int edlut_network_synchronize(edlut_network *n) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    return check_err(n->pop);
}

int edlut_network_spike_log(edlut_network *n, int32_t *step, int32_t *idx, int64_t cap, int64_t *n_out) {
    if (!n || !n_out) return fail(EDLUT_E_ARG, "null argument");
    int rc = edlut_network_synchronize(n);
    if (rc) return rc;
    int64_t lc[2] = {0, 0};
    HIP_TRY(hipMemcpy(lc, n->log_count, sizeof(lc), hipMemcpyDeviceToHost));
    int64_t cnt = lc[(n->last_step + 1) & 1];
    if (cnt > n->log_cap) cnt = n->log_cap;   // 0 when only the count is kept
    *n_out = cnt;
    if (cnt > cap) return fail(EDLUT_E_CAP, "spike log has %lld entries, buffer holds %lld", (long long)cnt, (long long)cap);
    if (cnt > 0) {
        HIP_TRY(hipMemcpy(step, n->log_step, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(idx, n->log_idx, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
    }
    return 0;
}

int edlut_network_set_source_log(edlut_network *n, int source, int64_t cap) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    Source &s = n->sources[(size_t)source];
    if (s.self || s.from || s.train) return fail(EDLUT_E_ARG, "source %d is not a Poisson device", source);
    if (cap < 0 || cap > (1ll << 31)) return fail(EDLUT_E_ARG, "log capacity out of range");
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    void *old[] = {s.log_step, s.log_idx, s.log_count};
    for (void *q : old)
        if (q) (void)hipFree(q);
    s.log_step = s.log_idx = nullptr;
    s.log_count = nullptr;
    s.log_cap = 0;
    if (cap > 0) {
        int rc;
        if ((rc = dalloc(&s.log_step, (size_t)cap)) || (rc = dalloc(&s.log_idx, (size_t)cap)) ||
            (rc = dalloc(&s.log_count, 2)))
            return rc;
        HIP_TRY(hipMemset(s.log_count, 0, 2 * sizeof(int64_t)));
        s.log_cap = cap;
    }
    n->dirty = true;   // captured kernels carry the log
    return 0;
}

int edlut_network_source_spike_log(edlut_network *n, int source, int32_t *step, int32_t *idx, int64_t cap,
                                   int64_t *n_out) {
    if (!n || !n_out) return fail(EDLUT_E_ARG, "null argument");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    const Source &s = n->sources[(size_t)source];
    if (!s.log_count) return fail(EDLUT_E_STATE, "source %d keeps no spike log", source);
    int rc = edlut_network_synchronize(n);
    if (rc) return rc;
    int64_t lc[2] = {0, 0};
    HIP_TRY(hipMemcpy(lc, s.log_count, sizeof(lc), hipMemcpyDeviceToHost));
    int64_t cnt = lc[(n->last_step + 1) & 1];
    if (cnt > s.log_cap) return fail(EDLUT_E_CAP, "source %d produced %lld spikes, its log holds %lld", source,
                                     (long long)cnt, (long long)s.log_cap);
    *n_out = cnt;
    if (cnt > cap) return fail(EDLUT_E_CAP, "spike log has %lld entries, buffer holds %lld", (long long)cnt, (long long)cap);
    if (cnt > 0) {
        HIP_TRY(hipMemcpy(step, s.log_step, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(idx, s.log_idx, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
    }
    return 0;
}

int edlut_network_clear_log(edlut_network *n) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipMemsetAsync(n->log_count, 0, 2 * sizeof(int64_t), n->pop->stream));
    for (Source &s : n->sources)
        if (s.log_count) HIP_TRY(hipMemsetAsync(s.log_count, 0, 2 * sizeof(int64_t), n->pop->stream));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    return 0;
}

// Network::FindOutConnections (Network.cpp:70-123) sorts synapses with
// qsort_inters (Network.cpp:45-68): target queue, source, delay, target.  With
// one queue this is (source, delay, target); two stable counting passes
// (target, then source*max_delay+delay) give that order in O(E + N), keeping
// the input order of exact duplicates.
int edlut_build_out_csr(int64_t n_src, int64_t n_tgt, int max_delay, int64_t n_edges,
                        const int32_t *src, const int32_t *tgt, const int32_t *delay,
                        const uint8_t *type, const float *w, int64_t *row_ptr,
                        uint32_t *tgt_type, float *w_out) {
    if (n_src <= 0 || n_tgt <= 0 || n_tgt >= (1ll << 30) || max_delay < 1 || n_edges < 0)
        return fail(EDLUT_E_ARG, "bad sizes");
    if (n_edges > 0 && (!src || !tgt || !delay || !type || !w || !tgt_type || !w_out))
        return fail(EDLUT_E_ARG, "null edge array");
    if (!row_ptr) return fail(EDLUT_E_ARG, "row_ptr is null");
    for (int64_t e = 0; e < n_edges; e++) {
        if (src[e] < 0 || src[e] >= n_src) return fail(EDLUT_E_ARG, "edge %lld: source out of range", (long long)e);
        if (tgt[e] < 0 || tgt[e] >= n_tgt) return fail(EDLUT_E_ARG, "edge %lld: target out of range", (long long)e);
        if (delay[e] < 1 || delay[e] > max_delay) return fail(EDLUT_E_ARG, "edge %lld: delay out of range", (long long)e);
        if (type[e] > 2) return fail(EDLUT_E_ARG, "edge %lld: type must be 0, 1 or 2", (long long)e);
    }
    // pass 1: stable by target
    std::vector<int64_t> off((size_t)n_tgt + 1, 0);
    for (int64_t e = 0; e < n_edges; e++) off[(size_t)tgt[e] + 1]++;
    for (int64_t t = 0; t < n_tgt; t++) off[(size_t)t + 1] += off[(size_t)t];
    std::vector<int64_t> by_t((size_t)n_edges);
    for (int64_t e = 0; e < n_edges; e++) by_t[(size_t)off[(size_t)tgt[e]]++] = e;
    // pass 2: stable by (source, delay)
    const int64_t nrow = n_src * (int64_t)max_delay;
    std::fill(row_ptr, row_ptr + nrow + 1, 0);
    for (int64_t e = 0; e < n_edges; e++) row_ptr[(int64_t)src[e] * max_delay + delay[e]]++;
    for (int64_t r = 0; r < nrow; r++) row_ptr[r + 1] += row_ptr[r];
    std::vector<int64_t> fill(row_ptr, row_ptr + nrow);
    for (int64_t q = 0; q < n_edges; q++) {
        const int64_t e = by_t[(size_t)q];
        const int64_t r = (int64_t)src[e] * max_delay + delay[e] - 1;
        const int64_t pos = fill[(size_t)r]++;
        tgt_type[pos] = (uint32_t)tgt[e] | ((uint32_t)type[e] << 30);
        w_out[pos] = w[e];
    }
    return 0;
}

int edlut_network_set_timing(edlut_network *n, int64_t max_steps) {
    if (!n || max_steps < 0) return fail(EDLUT_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    if (n->clk) (void)hipFree(n->clk);
    if (n->clk_next) (void)hipFree(n->clk_next);
    n->clk = nullptr;
    n->clk_next = nullptr;
    n->timed = 0;
    n->clk_cap = 0;
    n->dirty = true;
    if (max_steps == 0) return 0;
    n->clk_tiles_max = n->pop->stride / kFlagTile;   // the most workgroups of any launch shape
    int rc;
    if ((rc = dalloc(&n->clk, 2 * (size_t)max_steps * (size_t)n->clk_tiles_max)) ||
        (rc = dalloc(&n->clk_next, (size_t)max_steps)))
        return rc;
    n->clk_cap = max_steps;
    return 0;
}

int edlut_network_update_timing(edlut_network *n, double *ms_clock, int64_t *launches) {
    if (!n || !launches) return fail(EDLUT_E_ARG, "null argument");
    int rc = edlut_network_synchronize(n);
    if (rc) return rc;
    *launches = n->timed;
    double t_clock = 0.0;
    if (n->timed > 0) {
        const size_t nb = (size_t)n->clk_blocks;
        std::vector<unsigned long long> c(2 * (size_t)n->timed * nb);
        HIP_TRY(hipMemcpy(c.data(), n->clk, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost));
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, n->pop->device));
        if (khz <= 0) return fail(EDLUT_E_HIP, "device wall-clock rate unavailable");
        for (int64_t i = 0; i < n->timed; i++) {
            unsigned long long lo = ~0ull, hi = 0;
            for (size_t b = 0; b < nb; b++) {
                const unsigned long long s0 = c[(i * nb + b) * 2], s1 = c[(i * nb + b) * 2 + 1];
                if (s0 == 0 || s1 == 0) return fail(EDLUT_E_STATE, "a timed launch left no clock stamp");
                lo = s0 < lo ? s0 : lo;
                hi = s1 > hi ? s1 : hi;
            }
            t_clock += (double)(hi - lo) / (double)khz;   // ms
        }
    }
    if (ms_clock) *ms_clock = t_clock;
    return 0;
}

int edlut_network_update_dispatch_timing(edlut_network *n, double *ms, int64_t *launches) {
    if (!n || !ms || !launches) return fail(EDLUT_E_ARG, "null argument");
    int rc = edlut_network_synchronize(n);
    if (rc) return rc;
    *launches = n->timed;
    double t = 0.0;
    if (n->timed > 0) {
        const size_t nb = (size_t)n->clk_blocks;
        std::vector<unsigned long long> c(2 * (size_t)n->timed * nb), nx((size_t)n->timed);
        HIP_TRY(hipMemcpy(c.data(), n->clk, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(nx.data(), n->clk_next, sizeof(unsigned long long) * nx.size(), hipMemcpyDeviceToHost));
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, n->pop->device));
        if (khz <= 0) return fail(EDLUT_E_HIP, "device wall-clock rate unavailable");
        for (int64_t i = 0; i < n->timed; i++) {
            unsigned long long lo = ~0ull;
            for (size_t b = 0; b < nb; b++) {
                const unsigned long long s0 = c[(i * nb + b) * 2];
                if (s0 == 0) return fail(EDLUT_E_STATE, "a timed launch left no clock stamp");
                lo = s0 < lo ? s0 : lo;
            }
            if (nx[(size_t)i] == ~0ull || nx[(size_t)i] < lo)
                return fail(EDLUT_E_STATE, "timed update %lld has no following compaction stamp", (long long)i);
            t += (double)(nx[(size_t)i] - lo) / (double)khz;   // ms
        }
    }
    *ms = t;
    return 0;
}

// ---------------------------------------------------------------------------
// Synthetic recurrent wiring generated on the device (one source, like
// edlut_network_add_recurrent): see generate.hpp for the distribution.
int edlut_network_generate_recurrent(edlut_network *n, int in_degree, float exc_fraction, float w_mu,
                                     float w_sigma, uint64_t seed) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (n->self_src >= 0) return fail(EDLUT_E_STATE, "recurrent synapses already added");
    if ((int)n->sources.size() >= kMaxSources) return fail(EDLUT_E_CAP, "too many sources");
    if (in_degree < 1 || in_degree > 256) return fail(EDLUT_E_ARG, "in_degree must be in [1, 256]");
    if (!(exc_fraction >= 0.0f && exc_fraction <= 1.0f)) return fail(EDLUT_E_ARG, "exc_fraction must be in [0, 1]");
    if (!std::isfinite(w_mu) || !std::isfinite(w_sigma) || w_sigma < 0.0f)
        return fail(EDLUT_E_ARG, "bad weight distribution");
    edlut_population *p = n->pop;
    HIP_TRY(hipSetDevice(p->device));
    const int ng = (int)p->g_start.size();
    std::vector<int32_t> g_exc(ng);
    int64_t n_syn = 0;
    uint32_t types = 0;
    for (int g = 0; g < ng; g++) {
        if (p->g_size[g] < 2) return fail(EDLUT_E_ARG, "group %d has fewer than 2 neurons", g);
        g_exc[g] = (int32_t)std::lround((double)exc_fraction * p->g_size[g]);
        n_syn += (int64_t)p->g_size[g] * in_degree;
        if (g_exc[g] > 0) types |= EDLUT_SYN_EXC;
        if (g_exc[g] < p->g_size[g]) types |= EDLUT_SYN_INH;
    }
    if (n_syn >= (1ll << 40)) return fail(EDLUT_E_ARG, "too many synapses");
    const int64_t rows = (int64_t)p->n * n->max_delay;
    const int64_t nch = (rows + kScanChunk - 1) / kScanChunk;
    int32_t *d_start = nullptr, *d_size = nullptr, *d_exc = nullptr;
    uint32_t *row_cnt = nullptr;
    int64_t *sums = nullptr;
    uint64_t *key = nullptr;
    Source s;
    s.self = true;
    s.n = p->n;
    s.ring_idx = n->own_ring_idx;
    s.ring_cnt = n->own_ring_cnt;
    auto release = [&](int code) {
        void *tmp[] = {d_start, d_size, d_exc, row_cnt, sums, key};
        for (void *q : tmp)
            if (q) (void)hipFree(q);
        if (code) {
            void *kept[] = {s.row_ptr, s.tgt_type, s.w};
            for (void *q : kept)
                if (q) (void)hipFree(q);
        }
        return code;
    };
    int rc;
    if ((rc = dalloc(&d_start, ng)) || (rc = dalloc(&d_size, ng)) || (rc = dalloc(&d_exc, ng)) ||
        (rc = dalloc(&row_cnt, rows)) || (rc = dalloc(&sums, nch + 1)) || (rc = dalloc(&key, n_syn)) ||
        (rc = dalloc(&s.row_ptr, rows + 1)) || (rc = dalloc(&s.tgt_type, n_syn)) || (rc = dalloc(&s.w, n_syn)))
        return release(rc);
    if (hipMemcpy(d_start, p->g_start.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_size, p->g_size.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_exc, g_exc.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(row_cnt, 0, sizeof(uint32_t) * rows, p->stream) != hipSuccess)
        return release(fail(EDLUT_E_HIP, "generator set-up failed"));
    GenArgs ga;
    ga.g_start = d_start;
    ga.g_size = d_size;
    ga.g_exc = d_exc;
    ga.chunk_group = p->grouped ? p->d_chunk_group : nullptr;
    ga.n = p->n;
    ga.in_degree = in_degree;
    ga.max_delay = n->max_delay;
    ga.w_mu = w_mu;
    ga.w_sigma = w_sigma;
    ga.seed = seed;
    ga.row_cnt = row_cnt;
    ga.row_ptr = s.row_ptr;
    ga.key = key;
    ga.tgt_type = s.tgt_type;
    ga.w = s.w;
    const unsigned tb = (unsigned)((p->n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_gen_edges<false>), dim3(tb), dim3(kBlock), 0, p->stream, ga);
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nch), dim3(kBlock), 0, p->stream, row_cnt, rows, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, p->stream, sums, nch);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nch), dim3(kBlock), 0, p->stream, row_cnt, rows,
                       (const int64_t *)sums, s.row_ptr);
    if (hipMemsetAsync(row_cnt, 0, sizeof(uint32_t) * rows, p->stream) != hipSuccess)
        return release(fail(EDLUT_E_HIP, "hipMemsetAsync failed"));
    hipLaunchKernelGGL((k_gen_edges<true>), dim3(tb), dim3(kBlock), 0, p->stream, ga);
    hipLaunchKernelGGL(k_gen_rows, dim3((unsigned)((rows + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       p->stream, ga);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess)
        return release(fail(EDLUT_E_HIP, "connectivity generation failed"));
    int64_t total = 0;
    if (hipMemcpy(&total, s.row_ptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
        return release(fail(EDLUT_E_HIP, "hipMemcpy failed"));
    if (total != n_syn) return release(fail(EDLUT_E_STATE, "generated %lld synapses, expected %lld",
                                            (long long)total, (long long)n_syn));
    release(0);
    s.nsyn = n_syn;
    p->mask |= types;
    n->sources.push_back(s);
    n->dirty = true;
    n->seg_dirty = true;
    n->self_src = (int)n->sources.size() - 1;
    n->total_syn += (uint64_t)n_syn;
    return p->reserve_events(n->total_syn + 1);
}

int edlut_network_get_synapses(edlut_network *n, int source, int64_t *row_ptr, uint32_t *tgt_type,
                               float *w, int64_t *n_rows, int64_t *n_syn) {
    if (!n || !n_rows || !n_syn) return fail(EDLUT_E_ARG, "null argument");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    const Source &s = n->sources[source];
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    const int64_t rows = s.n * n->max_delay;
    int64_t total = 0;
    HIP_TRY(hipMemcpy(&total, s.row_ptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost));
    *n_rows = rows;
    *n_syn = total;
    if (row_ptr) HIP_TRY(hipMemcpy(row_ptr, s.row_ptr, sizeof(int64_t) * (rows + 1), hipMemcpyDeviceToHost));
    if (tgt_type && total) HIP_TRY(hipMemcpy(tgt_type, s.tgt_type, sizeof(uint32_t) * total, hipMemcpyDeviceToHost));
    if (w && total) HIP_TRY(hipMemcpy(w, s.w, sizeof(float) * total, hipMemcpyDeviceToHost));
    return 0;
}

// Back to the state after construction: initial neuron state, generators at
// their initial states, empty spike rings and log, next run starts at step 1.
int edlut_network_reset(edlut_network *n) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    edlut_population *p = n->pop;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    // every state row except the external current (an input, kept as set)
    HIP_TRY(hipMemcpy(p->state, p->init_state.data(), sizeof(float) * (size_t)(p->nv - 1) * p->stride,
                      hipMemcpyHostToDevice));
    if (p->last_spike) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(kBlock), 0, p->stream, p->last_spike, p->stride,
                           -(1 << 30));
        HIP_TRY(hipGetLastError());
    }
    if (p->bdf_st) {
        HIP_TRY(hipMemsetAsync(p->bdf_st, 0, sizeof(int32_t) * p->stride, p->stream));
        int rc = p->alloc_bdf_history(p->newton_fp64 ? sizeof(double) : sizeof(float));
        if (rc) return rc;
    }
    for (Source &s : n->sources) {
        if (!s.self && !s.lcg0.empty())
            HIP_TRY(hipMemcpy(s.lcg, s.lcg0.data(), sizeof(uint32_t) * s.lcg0.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(s.ring_cnt, 0, sizeof(int32_t) * n->ring, p->stream));
        if (s.log_count) HIP_TRY(hipMemsetAsync(s.log_count, 0, 2 * sizeof(int64_t), p->stream));
    }
    HIP_TRY(hipMemsetAsync(n->own_ring_cnt, 0, sizeof(int32_t) * n->ring, p->stream));
    HIP_TRY(hipMemsetAsync(n->log_count, 0, 2 * sizeof(int64_t), p->stream));
    HIP_TRY(hipMemsetAsync(p->err, 0, sizeof(uint32_t), p->stream));
    HIP_TRY(hipMemsetAsync(n->d_time, 0, sizeof(double) * n->ring, p->stream));
    HIP_TRY(hipMemsetAsync(n->d_time_p, 0, sizeof(double) * n->ring, p->stream));
    HIP_TRY(hipMemsetAsync(n->d_skip, 0, 2 * sizeof(uint32_t), p->stream));
    for (Source &s : n->sources) {
        if (!s.stdp) continue;
        if (s.w0.size() == (size_t)s.nsyn && s.nsyn)
            HIP_TRY(hipMemcpy(s.w, s.w0.data(), sizeof(float) * s.nsyn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(s.syn.pre, 0, sizeof(float) * s.nsyn, p->stream));
        HIP_TRY(hipMemsetAsync(s.syn.post, 0, sizeof(float) * s.nsyn, p->stream));
        HIP_TRY(hipMemsetAsync(s.syn.last, 0, sizeof(double) * s.nsyn, p->stream));
    }
    HIP_TRY(hipMemsetAsync(p->cnt, 0, sizeof(uint32_t) * 3 * p->stride, p->stream));
    HIP_TRY(hipMemsetAsync(p->fill, 0, sizeof(uint32_t) * 3 * p->stride, p->stream));
    HIP_TRY(hipMemsetAsync(p->seg, 0xff, sizeof(uint64_t) * 3 * p->stride, p->stream));
    HIP_TRY(hipMemsetAsync(p->tile_flag, 0, sizeof(uint32_t) * (p->stride / kFlagTile), p->stream));
    n->fold_ready = -1;
    HIP_TRY(hipStreamSynchronize(p->stream));
    n->last_step = 0;
    return 0;
}

int edlut_network_set_launch(edlut_network *n, int items, int fold_lanes, int fold_blocks) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    edlut_population *p = n->pop;
    const int fixed = plain_items(p->model, p->integ);
    const bool two_or_four = p->grouped || plain_two_items(p);
    if (two_or_four ? (items != 2 && items != 4) : items != fixed)
        return fail(EDLUT_E_ARG, "items per thread must be %s", two_or_four ? "2 or 4" : (fixed == 2 ? "2" : "4"));
    if (!pick_ring_fold(kCount, fold_lanes)) return fail(EDLUT_E_ARG, "fold lanes must be 8, 16, 32 or 64");
    if (fold_blocks < 1 || fold_blocks > 65536) return fail(EDLUT_E_ARG, "fold blocks must be in [1, 65536]");
    p->items = items;
    n->fold_lanes = fold_lanes;
    n->fold_blocks = fold_blocks;
    n->dirty = true;
    return 0;
}

int edlut_network_get_launch(const edlut_network *n, int *items, int *fold_lanes, int *fold_blocks) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (items) *items = n->pop->items;
    if (fold_lanes) *fold_lanes = n->fold_lanes;
    if (fold_blocks) *fold_blocks = n->fold_blocks;
    return 0;
}

// Times `trial_steps` steps (after 16 warm-up steps) of every launch shape
// and keeps the fastest; the network is reset before each trial and after
// the search.  The results of a run do not depend on the launch shape.
int edlut_network_autotune(edlut_network *n, int trial_steps, int *items, int *fold_lanes,
                           int *fold_blocks, double *ms_per_step) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (trial_steps < 16 || trial_steps > 100000) return fail(EDLUT_E_ARG, "trial_steps must be in [16, 100000]");
    edlut_population *p = n->pop;
    HIP_TRY(hipSetDevice(p->device));
    std::vector<int> it_c = (p->grouped || plain_two_items(p)) ? std::vector<int>{2, 4}
                                                                : std::vector<int>{plain_items(p->model, p->integ)};
    const int lanes_c[] = {16, 32, 64};
    const int blocks_c[] = {64, 256, 1024, 4096};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e300;
    int b_it = p->items, b_l = n->fold_lanes, b_b = n->fold_blocks;
    int rc = 0;
    for (int it : it_c)
        for (int l : lanes_c)
            for (int b : blocks_c) {
                if ((rc = edlut_network_set_launch(n, it, l, b)) || (rc = edlut_network_reset(n)) ||
                    (rc = edlut_network_run(n, 1, 16)))
                    goto done;
                if (hipEventRecord(e0, p->stream) != hipSuccess) {
                    rc = fail(EDLUT_E_HIP, "hipEventRecord failed");
                    goto done;
                }
                if ((rc = edlut_network_run(n, 17, trial_steps))) goto done;
                if (hipEventRecord(e1, p->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(EDLUT_E_HIP, "timing failed");
                    goto done;
                }
                if ((rc = check_err(p))) goto done;
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms < best) {
                    best = ms;
                    b_it = it;
                    b_l = l;
                    b_b = b;
                }
            }
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = edlut_network_set_launch(n, b_it, b_l, b_b)) || (rc = edlut_network_reset(n))) return rc;
    if (items) *items = b_it;
    if (fold_lanes) *fold_lanes = b_l;
    if (fold_blocks) *fold_blocks = b_b;
    if (ms_per_step) *ms_per_step = best / trial_steps;
    return 0;
}

// ---------------------------------------------------------------------------
// Additive STDP on every synapse of one source (STDPWeightChange): times in
// seconds, weights clamped to [0, max_weight].
int edlut_network_set_stdp(edlut_network *n, int source, float max_ltp, float tau_ltp, float max_ltd,
                           float tau_ltd, float max_weight) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    // STDPWeightChange::SetParameters (STDPWeightChange.cpp:107-160)
    if (!(tau_ltp > 0.0f) || !(tau_ltd > 0.0f)) return fail(EDLUT_E_ARG, "STDP time constants must be positive");
    if (!std::isfinite(max_ltp) || !std::isfinite(max_ltd) || !std::isfinite(max_weight) || max_weight < 0.0f)
        return fail(EDLUT_E_ARG, "bad STDP amplitudes or maximum weight");
    Source &s = n->sources[(size_t)source];
    if (s.stdp) return fail(EDLUT_E_STATE, "source %d already has STDP", source);
    edlut_population *p = n->pop;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (n->fold_ready >= 0) {
        // events folded ahead without the learning rule: fold them again
        HIP_TRY(hipMemset(p->fill, 0, sizeof(uint32_t) * 3 * p->stride));
        HIP_TRY(hipMemset(p->tile_flag, 0, sizeof(uint32_t) * (p->stride / kFlagTile)));
        n->fold_ready = -1;
    }
    // targets of every synapse -> per-target list of synapse positions
    std::vector<uint32_t> tt((size_t)s.nsyn);
    s.w0.resize((size_t)s.nsyn);
    if (s.nsyn) {
        HIP_TRY(hipMemcpy(tt.data(), s.tgt_type, sizeof(uint32_t) * s.nsyn, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(s.w0.data(), s.w, sizeof(float) * s.nsyn, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> ptr((size_t)p->n + 1, 0), syn((size_t)s.nsyn);
    for (int64_t q = 0; q < s.nsyn; q++) ptr[(tt[(size_t)q] & 0x3fffffffu) + 1]++;
    s.in_max = 0;
    for (int i = 0; i < p->n; i++) {
        if (ptr[(size_t)i + 1] > s.in_max) s.in_max = ptr[(size_t)i + 1];
        ptr[(size_t)i + 1] += ptr[(size_t)i];
    }
    std::vector<int64_t> fillp(ptr.begin(), ptr.end() - 1);
    for (int64_t q = 0; q < s.nsyn; q++) syn[(size_t)fillp[tt[(size_t)q] & 0x3fffffffu]++] = q;
    int rc;
    const size_t m = s.nsyn > 0 ? (size_t)s.nsyn : 1;
    if ((rc = dalloc(&s.in_ptr, ptr.size())) || (rc = dalloc(&s.in_syn, m)) || (rc = dalloc(&s.syn.pre, m)) ||
        (rc = dalloc(&s.syn.post, m)) || (rc = dalloc(&s.syn.last, m)))
        return rc;
    HIP_TRY(hipMemcpy(s.in_ptr, ptr.data(), sizeof(int64_t) * ptr.size(), hipMemcpyHostToDevice));
    if (s.nsyn) HIP_TRY(hipMemcpy(s.in_syn, syn.data(), sizeof(int64_t) * s.nsyn, hipMemcpyHostToDevice));
    // ConnectionState.cpp:22: traces and last update times start at zero
    HIP_TRY(hipMemset(s.syn.pre, 0, sizeof(float) * m));
    HIP_TRY(hipMemset(s.syn.post, 0, sizeof(float) * m));
    HIP_TRY(hipMemset(s.syn.last, 0, sizeof(double) * m));
    s.syn.a_plus = max_ltp;
    s.syn.a_minus = max_ltd;
    // STDPState.cpp:24-27
    s.syn.inv_tau_plus = 1.0f / tau_ltp;
    s.syn.inv_tau_minus = 1.0f / tau_ltd;
    s.syn.w_max = max_weight;
    s.stdp = true;
    n->plastic = true;
    n->dirty = true;
    return 0;
}

int edlut_network_get_stdp(edlut_network *n, int source, float *pre, float *post, double *last) {
    if (!n) return fail(EDLUT_E_ARG, "network is null");
    if (source < 0 || source >= (int)n->sources.size()) return fail(EDLUT_E_ARG, "no source %d", source);
    const Source &s = n->sources[(size_t)source];
    if (!s.stdp) return fail(EDLUT_E_STATE, "source %d has no STDP", source);
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    if (s.nsyn == 0) return 0;
    if (pre) HIP_TRY(hipMemcpy(pre, s.syn.pre, sizeof(float) * s.nsyn, hipMemcpyDeviceToHost));
    if (post) HIP_TRY(hipMemcpy(post, s.syn.post, sizeof(float) * s.nsyn, hipMemcpyDeviceToHost));
    if (last) HIP_TRY(hipMemcpy(last, s.syn.last, sizeof(double) * s.nsyn, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// Trace fitting: forward simulation and discrete adjoint (adjoint.hpp).
}  // extern "C"

struct edlut_fit {
    int model = 0, n = 0, steps = 0, n_syn = 0, ckpt = 100, device = 0;
    FitArgs base{};
    float *refr = nullptr;
    float *ck_state = nullptr;
    int32_t *ck_int = nullptr;
    float *scratch = nullptr;
    int32_t *scratch_int = nullptr;
    bool forward_done = false;
};

extern "C" {

int edlut_refractory_table(float dt_s, float tau_ref, float *out, int cap, int *len) {
    if (!len || !(dt_s > 0.0f)) return fail(EDLUT_E_ARG, "bad argument");
    std::vector<float> r = build_refractory(dt_s, dt_s * 1000.0f, tau_ref);
    *len = (int)r.size();
    if (out) {
        if (cap < (int)r.size()) return fail(EDLUT_E_CAP, "refractory table has %d entries", (int)r.size());
        std::memcpy(out, r.data(), sizeof(float) * r.size());
    }
    return 0;
}

int edlut_fit_create(int model, const float *params, int n_params, float dt_s, int n_neurons, int steps,
                     int n_syn, int checkpoint, int device, edlut_fit **out) {
    if (!out) return fail(EDLUT_E_ARG, "out is null");
    *out = nullptr;
    if (model != EDLUT_LIF && model != EDLUT_ADEX)
        return fail(EDLUT_E_ARG, "trace fitting takes LIF (Euler) or AdEx (RK4)");
    const int need = model == EDLUT_ADEX ? 14 : 10;
    if (!params || n_params != need) return fail(EDLUT_E_ARG, "model takes %d parameters", need);
    if (!(dt_s > 0.0f) || n_neurons <= 0 || steps <= 0 || n_syn <= 0 || checkpoint <= 0)
        return fail(EDLUT_E_ARG, "sizes must be positive");
    for (int i = 0; i < n_params; i++)
        if (!std::isfinite(params[i])) return fail(EDLUT_E_ARG, "parameter %d is not finite", i);
    HIP_TRY(hipSetDevice(device));
    edlut_fit *f = new edlut_fit;
    f->model = model;
    f->n = n_neurons;
    f->steps = steps;
    f->n_syn = n_syn;
    f->ckpt = checkpoint < steps ? checkpoint : steps;
    f->device = device;
    FitArgs &a = f->base;
    std::memset(&a, 0, sizeof(a));
    a.model = model == EDLUT_LIF ? 0 : 1;
    a.n = n_neurons;
    a.steps = steps;
    a.n_syn = n_syn;
    a.ckpt = f->ckpt;
    a.dt = dt_s * 1000.0f;
    a.h = model == EDLUT_ADEX ? a.dt * 0.5f : a.dt;
    a.s6 = a.dt * 0.166666666667f;
    std::vector<float> refr;
    if (model == EDLUT_LIF) {
        // e_exc e_inh e_leak v_thr c_m tau_exc tau_inh tau_ref g_leak tau_nmda
        a.e_exc = params[0];
        a.e_inh = params[1];
        a.v_thr = params[3];
        a.inv_c_m = inv_of(params[4]);
        refr = build_refractory(dt_s, dt_s * 1000.0f, params[7]);
    } else {
        // a b thr_slo_fac v_thr tau_w e_exc e_inh e_reset e_leak g_leak c_m tau_exc tau_inh tau_nmda
        a.a = params[0];
        a.b = params[1];
        a.thr = params[2];
        a.inv_thr = inv_of(params[2]);
        a.v_thr = params[3];
        a.inv_tau_w = inv_of(params[4]);
        a.e_exc = params[5];
        a.e_inh = params[6];
        a.e_reset = params[7];
        a.inv_c_m = inv_of(params[10]);
        a.v_cap = params[3] + 6 * params[2];
    }
    a.refr_len = (int)refr.size();
    const int segs = (steps + f->ckpt - 1) / f->ckpt;
    int rc;
    if ((rc = dalloc(&f->refr, refr.size() + 1)) || (rc = dalloc(&f->ck_state, (size_t)(segs + 1) * 4 * n_neurons)) ||
        (rc = dalloc(&f->ck_int, (size_t)(segs + 1) * 2 * n_neurons)) ||
        (rc = dalloc(&f->scratch, (size_t)f->ckpt * kFitStore * n_neurons)) ||
        (rc = dalloc(&f->scratch_int, (size_t)f->ckpt * n_neurons))) {
        edlut_fit_destroy(f);
        return rc;
    }
    if (!refr.empty() &&
        hipMemcpy(f->refr, refr.data(), sizeof(float) * refr.size(), hipMemcpyHostToDevice) != hipSuccess) {
        edlut_fit_destroy(f);
        return fail(EDLUT_E_HIP, "hipMemcpy failed");
    }
    a.refr = f->refr;
    a.ck_state = f->ck_state;
    a.ck_int = f->ck_int;
    a.scratch = f->scratch;
    a.scratch_int = f->scratch_int;
    *out = f;
    return 0;
}

int edlut_fit_destroy(edlut_fit *f) {
    if (!f) return 0;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();
    void *dev[] = {f->refr, f->ck_state, f->ck_int, f->scratch, f->scratch_int};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    delete f;
    return 0;
}

static int fit_args(edlut_fit *f, const edlut_fit_tensors *t, FitArgs &a) {
    if (!f || !t) return fail(EDLUT_E_ARG, "null argument");
    if (!t->g_leak || !t->e_leak || !t->tau_exc || !t->tau_inh || !t->weight || !t->syn_type || !t->ev_ptr ||
        !t->ev_step || !t->ev_syn)
        return fail(EDLUT_E_ARG, "null tensor");
    a = f->base;
    a.gl = t->g_leak;
    a.el = t->e_leak;
    a.te = t->tau_exc;
    a.ti = t->tau_inh;
    a.w = t->weight;
    a.syn_type = t->syn_type;
    a.ev_ptr = t->ev_ptr;
    a.ev_step = t->ev_step;
    a.ev_syn = t->ev_syn;
    a.v_trace = t->v_trace;
    a.spikes = t->spikes;
    a.traj = t->traj;
    a.traj_since = t->traj_since;
    if ((a.traj == nullptr) != (a.traj_since == nullptr))
        return fail(EDLUT_E_ARG, "traj and traj_since go together");
    a.grad_v = t->grad_v;
    a.grad_par = t->grad_par;
    a.grad_w = t->grad_w;
    return 0;
}

int edlut_fit_forward(edlut_fit *f, const edlut_fit_tensors *t, void *stream) {
    FitArgs a;
    int rc = fit_args(f, t, a);
    if (rc) return rc;
    if (!t->v_trace) return fail(EDLUT_E_ARG, "v_trace is null");
    HIP_TRY(hipSetDevice(f->device));
    const int blocks = (f->n + kFitBlock - 1) / kFitBlock;
    hipLaunchKernelGGL(k_fit_forward, dim3(blocks), dim3(kFitBlock), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError());
    f->forward_done = true;
    return 0;
}

int edlut_fit_backward(edlut_fit *f, const edlut_fit_tensors *t, void *stream) {
    FitArgs a;
    int rc = fit_args(f, t, a);
    if (rc) return rc;
    if (!t->grad_v || !t->grad_par || !t->grad_w) return fail(EDLUT_E_ARG, "gradient tensors are null");
    if (!f->forward_done) return fail(EDLUT_E_STATE, "the backward pass needs the forward pass's checkpoints");
    HIP_TRY(hipSetDevice(f->device));
    const int blocks = (f->n + kFitBlock - 1) / kFitBlock;
    hipLaunchKernelGGL(k_fit_backward, dim3(blocks), dim3(kFitBlock), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Spike log capacity (entries); 0 keeps only the spike count.
int edlut_network_set_log_capacity(edlut_network *n, int64_t cap) {
    if (!n || cap < 0) return fail(EDLUT_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(n->pop->device));
    HIP_TRY(hipStreamSynchronize(n->pop->stream));
    if (n->log_step) (void)hipFree(n->log_step);
    if (n->log_idx) (void)hipFree(n->log_idx);
    n->log_step = nullptr;
    n->log_idx = nullptr;
    n->log_cap = 0;
    int rc;
    if (cap > 0 && ((rc = dalloc(&n->log_step, cap)) || (rc = dalloc(&n->log_idx, cap)))) return rc;
    n->log_cap = cap;
    n->dirty = true;
    HIP_TRY(hipMemset(n->log_count, 0, 2 * sizeof(int64_t)));
    return 0;
}

int edlut_network_spike_count(edlut_network *n, int64_t *count) {
    if (!n || !count) return fail(EDLUT_E_ARG, "null argument");
    int rc = edlut_network_synchronize(n);
    if (rc) return rc;
    int64_t lc[2] = {0, 0};
    HIP_TRY(hipMemcpy(lc, n->log_count, sizeof(lc), hipMemcpyDeviceToHost));
    *count = lc[(n->last_step + 1) & 1];
    return 0;
}

}  // extern "C"
