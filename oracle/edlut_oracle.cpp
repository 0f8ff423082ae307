/*
 * edlut_oracle.cpp -- CPU restatement of EDLUT's time-driven neuron update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, the smoke entry point and bench.py's CPU-baseline leg use it.
 *
 * Layout follows the reference's CPU path: array-of-structures state
 * (VectorNeuronState.cpp:244-276, index*NumberOfVariables+var), one loop over
 * neurons per integration step, spikes appended in ascending index order.
 * Float arithmetic keeps the reference's association; build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Citations are relative to the reference repository root (Kernel/...).
 */
#include "edlut_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

/* ------------------------------------------------------------------------ */
/* ExponentialTable: Kernel/include/simulation/ExponentialTable.h:60-103,    */
/* Kernel/src/simulation/ExponentialTable.cpp:23-32.                        */
/* 16384 entries on [-20, 20]; entry i holds exp of the float grid point     */
/* Min + ((Max-Min)*i)/(TableSize-1), evaluated by the double-precision      */
/* exp and stored as float.  GetResult truncates (x-Min)*aux to an int      */
/* index, returns exp(x) above Max and 0 below Min.                          */
/* ------------------------------------------------------------------------ */
const float kExpMin = -20.0f;
const float kExpMax = 20.0f;
const int kExpSize = 1024 * 16;
const float kExpAux = (kExpSize - 1) / (kExpMax - kExpMin);

struct ExpTable {
    float v[kExpSize];
    ExpTable() {
        for (int i = 0; i < kExpSize; i++) {
            float exponent = kExpMin + ((kExpMax - kExpMin) * i) / (kExpSize - 1);
            v[i] = (float)::exp((double)exponent);
        }
    }
};
const ExpTable &exp_table() {
    static ExpTable t;
    return t;
}

inline float exp_lut(float value) {
    if (value >= kExpMin && value <= kExpMax) {
        int position = int((value - kExpMin) * kExpAux);
        return exp_table().v[position];
    } else if (value > kExpMax) {
        /* In the model translation units `using namespace std` is active
         * (Kernel/include/neuron_model/NeuronModel.h), so exp(float) is the
         * float overload. */
        return std::exp(value);
    }
    return 0.0f;
}

/* float -> int conversion as compiled for x86-64 (cvttss2si): NaN and values
 * outside the int32 range give INT_MIN.  Written out explicitly because the
 * C++ conversion is undefined there and the reference's results depend on
 * what the x86 instruction returns. */
inline int trunc_i32_x86(float x) {
    if (!(x < 2147483648.0f) || x < -2147483648.0f) return (int)0x80000000u;
    return (int)x;
}

/* g_nmda_inf table: AdExTimeDrivenModel.cpp:27-47 (same in LIF/HH/granule),
 * TableSizeNMDA = 256 (TimeDrivenNeuronModel.h:103).  exp(float) resolves to
 * the float overload (using namespace std in scope). */
const int kNmdaSize = 256;
struct NmdaTable {
    float v[kNmdaSize];
    float aux;
    /* k = -0.062 per mV; the SI-unit LIF uses -62 per V
     * (LIFTimeDrivenModel_IS.cpp:25-33) */
    void build(float e_exc, float e_inh, float k = -0.062f) {
        aux = (kNmdaSize - 1) / (e_exc - e_inh);
        for (int i = 0; i < kNmdaSize; i++) {
            float V = e_inh + ((e_exc - e_inh) * i) / (kNmdaSize - 1);
            v[i] = 1.0f / (1.0f + std::exp(k * V) * (1.2f / 3.57f));
        }
    }
    float get(float V_m, float e_inh) const {
        int position = trunc_i32_x86((V_m - e_inh) * aux);
        if (position < 0) position = 0;
        else if (position > (kNmdaSize - 1)) position = kNmdaSize - 1;
        return v[position];
    }
};

/* ------------------------------------------------------------------------ */
/* Real-type traits: float = bit-exact restatement, double = FP64 oracle.   */
/* ------------------------------------------------------------------------ */
template <typename R> struct Num;
template <> struct Num<float> {
    static float lut_exp(float x) { return exp_lut(x); }
    static float decay(float elapsed, float inv_tau) { return expf(-elapsed * inv_tau); }
    /* inv_x = 1./x computed in double and stored in a float member
     * (AdExTimeDrivenModel.cpp:347,362,405,413). */
    static float inv(float x) { return (float)(1.0 / (double)x); }
};
template <> struct Num<double> {
    static double lut_exp(double x) { return std::exp(x); }
    static double decay(double elapsed, double inv_tau) { return std::exp(-elapsed * inv_tau); }
    static double inv(double x) { return 1.0 / x; }
};

/* ------------------------------------------------------------------------ */
/* Population: state and bookkeeping shared by every model.                 */
/* ------------------------------------------------------------------------ */
template <typename R> struct PopBase {
    int N = 0;
    int nvars = 0;
    int ndiff = 0;
    std::vector<R> state;          /* AoS: index*nvars + var */
    std::vector<double> last_spike; /* LastSpikeTime, seconds */
    std::vector<R> initial;
    float dt_s = 1e-4f;
    R dt_model = 0;                /* elapsedTimeInNeuronModelScale */
    bool EXC = false, INH = false, NMDA = false, EXT_I = false;
    std::vector<int32_t> spikes;   /* InternalSpikeIndexs of the last update */
    std::vector<uint8_t> fired_tmp; /* per-neuron spike flags of a parallel update */
    std::vector<int> bdf_state;     /* BDFn::state, empty for other integrators */
    /* integration_method->resetState(index) (BDFn.h resetState: state = 0) */
    void reset_integrator(int i) {
        if (!bdf_state.empty()) bdf_state[i] = 0;
    }

    void init_states(const R *init) {
        state.assign((size_t)N * nvars, 0);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < nvars; j++) state[(size_t)i * nvars + j] = init[j];
        initial.assign(init, init + nvars);
        /* VectorNeuronState.cpp:272-274 */
        last_spike.assign(N, 100.0);
    }
    R *at(int i) { return &state[(size_t)i * nvars]; }
    /* NeuronModel time scale: milliseconds (MilisecondScale, NeuronModel.h:71) */
    float time_scale_factor() const { return 1000.0f; }
    /* VectorNeuronState.cpp:221-224 */
    void add_elapsed(int i) { last_spike[i] += (double)dt_s; }
    /* VectorNeuronState.cpp:215-218 */
    void fired(int i) { last_spike[i] = 0; }
    /* VectorNeuronState.cpp:302-306 via TimeDrivenNeuronModel.cpp:75-88 */
    void reset(int i) {
        for (int j = 0; j < nvars; j++) state[(size_t)i * nvars + j] = initial[j];
    }
};

/* ------------------------------------------------------------------------ */
/* AdEx: Kernel/src/neuron_model/AdExTimeDrivenModel.cpp                    */
/* state: V, w | g_exc, g_inh, g_nmda, ext_I                                 */
/* ------------------------------------------------------------------------ */
template <typename R> struct AdEx : PopBase<R> {
    R a, b, thr_slo_fac, inv_thr_slo_fac, v_thr, tau_w, inv_tau_w, e_exc, e_inh, e_reset,
        e_leak, g_leak, c_m, inv_c_m, tau_exc, inv_tau_exc, tau_inh, inv_tau_inh, tau_nmda,
        inv_tau_nmda;
    NmdaTable nmda;
    R cond[3]; /* conductance decay factors, elapsed-time index 0 */

    /* Default parameters: AdExTimeDrivenModel.cpp:449-467 (order of p[]). */
    void set_params(const float *p) {
        a = p[0]; b = p[1]; thr_slo_fac = p[2]; v_thr = p[3]; tau_w = p[4]; e_exc = p[5];
        e_inh = p[6]; e_reset = p[7]; e_leak = p[8]; g_leak = p[9]; c_m = p[10];
        tau_exc = p[11]; tau_inh = p[12]; tau_nmda = p[13];
        inv_thr_slo_fac = Num<R>::inv(thr_slo_fac);
        inv_tau_w = Num<R>::inv(tau_w);
        inv_c_m = Num<R>::inv(c_m);
        inv_tau_exc = Num<R>::inv(tau_exc);
        inv_tau_inh = Num<R>::inv(tau_inh);
        inv_tau_nmda = Num<R>::inv(tau_nmda);
        nmda.build(p[5], p[6]);
        this->nvars = 6;
        this->ndiff = 2;
    }
    R g_nmda_inf(R V) const {
        if (sizeof(R) == sizeof(float)) return (R)nmda.get((float)V, (float)e_inh);
        return 1.0 / (1.0 + std::exp(-0.062 * (double)V) * (1.2 / 3.57));
    }
    /* AdExTimeDrivenModel.cpp:136-148 */
    void init_states() {
        R init[6] = {e_leak, 0, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    /* AdExTimeDrivenModel.cpp:247-254 */
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* AdExTimeDrivenModel.cpp:170-218 */
    void deriv(R *S, R *D, int /*index*/, R /*elapsed*/) const {
        if (S[0] <= v_thr + 6 * thr_slo_fac) {
            R current = 0.0;
            if (this->EXC) current += S[2] * (e_exc - S[0]);
            if (this->INH) current += S[3] * (e_inh - S[0]);
            if (this->NMDA) {
                R ginf = g_nmda_inf(S[0]);
                current += S[4] * ginf * (e_exc - S[0]);
            }
            current += S[5];
            D[0] = (g_leak * (e_leak - S[0]) +
                    g_leak * thr_slo_fac * Num<R>::lut_exp((S[0] - v_thr) * inv_thr_slo_fac) - S[1] +
                    current) * inv_c_m;
            D[1] = (a * (S[0] - e_leak) - S[1]) * inv_tau_w;
        } else if (S[0] <= 0.0f) {
            R Vm = v_thr + 6 * thr_slo_fac;
            R current = 0.0;
            if (this->EXC) current += S[2] * (e_exc - Vm);
            if (this->INH) current += S[3] * (e_inh - Vm);
            if (this->NMDA) {
                R ginf = g_nmda_inf(Vm);
                current += S[4] * ginf * (e_exc - Vm);
            }
            current += S[5];
            D[0] = (g_leak * (e_leak - Vm) +
                    g_leak * thr_slo_fac * Num<R>::lut_exp((Vm - v_thr) * inv_thr_slo_fac) - S[1] +
                    current) * inv_c_m;
            D[1] = (a * (S[0] - e_leak) - S[1]) * inv_tau_w;
        } else {
            D[0] = 0;
            D[1] = 0;
        }
    }
    /* AdExTimeDrivenModel.cpp:220-245 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[2] < limit) S[2] = 0.0f; else S[2] *= cond[0]; }
        if (this->INH) { if (S[3] < limit) S[3] = 0.0f; else S[3] *= cond[1]; }
        if (this->NMDA) { if (S[4] < limit) S[4] = 0.0f; else S[4] *= cond[2]; }
    }
    /* AdExTimeDrivenModel.cpp:159-168 */
    bool spike(R /*prev_V*/, R *S, int index) {
        if (S[0] > 0.0f) {
            S[0] = e_reset;
            S[1] += b;
            this->fired(index);
            this->reset_integrator(index);
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------------ */
/* LIF: Kernel/src/neuron_model/LIFTimeDrivenModel.cpp                        */
/* state: V | g_exc, g_inh, g_nmda, ext_I                                     */
/* ------------------------------------------------------------------------ */
template <typename R> struct LIF : PopBase<R> {
    R e_exc, e_inh, e_leak, v_thr, c_m, inv_c_m, tau_exc, inv_tau_exc, tau_inh, inv_tau_inh,
        tau_ref, g_leak, tau_nmda, inv_tau_nmda;
    NmdaTable nmda;
    R cond[3];
    const float time_scale = 1000.0f; /* MilisecondScale (NeuronModel.h:71) */

    /* Default parameters: LIFTimeDrivenModel.cpp:435-450 (order of p[]). */
    void set_params(const float *p) {
        e_exc = p[0]; e_inh = p[1]; e_leak = p[2]; v_thr = p[3]; c_m = p[4];
        tau_exc = p[5]; tau_inh = p[6]; tau_ref = p[7]; g_leak = p[8]; tau_nmda = p[9];
        inv_c_m = Num<R>::inv(c_m);
        inv_tau_exc = Num<R>::inv(tau_exc);
        inv_tau_inh = Num<R>::inv(tau_inh);
        inv_tau_nmda = Num<R>::inv(tau_nmda);
        nmda.build(p[0], p[1]);
        this->nvars = 5;
        this->ndiff = 1;
    }
    R g_nmda_inf(R V) const {
        if (sizeof(R) == sizeof(float)) return (R)nmda.get((float)V, (float)e_inh);
        return 1.0 / (1.0 + std::exp(-0.062 * (double)V) * (1.2 / 3.57));
    }
    /* LIFTimeDrivenModel.cpp:135-148 */
    void init_states() {
        R init[5] = {e_leak, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* LIFTimeDrivenModel.cpp:170-202.  The refractory test works on the
     * double LastSpikeTime scaled to ms (double * float -> double). */
    void deriv(R *S, R *D, int index, R elapsed_time) const {
        R current = 0.0;
        if (this->EXC) current += S[1] * (e_exc - S[0]);
        if (this->INH) current += S[2] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[3] * ginf * (e_exc - S[0]);
        }
        current += S[4];
        double lst = this->last_spike[index] * time_scale;
        if (lst > tau_ref) {
            D[0] = (current + g_leak * (e_leak - S[0])) * inv_c_m;
        } else if ((lst + elapsed_time) > tau_ref) {
            R fraction = (R)((lst + elapsed_time - tau_ref) / elapsed_time);
            D[0] = fraction * ((current + g_leak * (e_leak - S[0])) * inv_c_m);
        } else {
            D[0] = 0;
        }
    }
    /* LIFTimeDrivenModel.cpp:204-242 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[1] < limit) S[1] = 0.0f; else S[1] *= cond[0]; }
        if (this->INH) { if (S[2] < limit) S[2] = 0.0f; else S[2] *= cond[1]; }
        if (this->NMDA) { if (S[3] < limit) S[3] = 0.0f; else S[3] *= cond[2]; }
    }
    /* LIFTimeDrivenModel.cpp:158-168 */
    bool spike(R /*prev_V*/, R *S, int index) {
        if (S[0] > v_thr) {
            S[0] = e_leak;
            this->fired(index);
            this->reset_integrator(index);
            return true;
        }
        return false;
    }
};


/* ------------------------------------------------------------------------ */
/* HH: Kernel/src/neuron_model/HHTimeDrivenModel.cpp,                         */
/* Kernel/include/neuron_model/HHTimeDrivenModel.h:372-441                    */
/* state: V, m, h, n | g_exc, g_inh, g_nmda, ext_I                           */
/* Channel rates come from a 262144-entry table on [-90, 50] mV built with    */
/* the float exp (the header's inline generator runs in a translation unit    */
/* with `using namespace std`).                                              */
/* ------------------------------------------------------------------------ */
const int kHHTable = 1024 * 256;
const float kHHMinV = -90.0f;
const float kHHMaxV = 50.0f;
const float kHHAux = (kHHTable - 1) / (kHHMaxV - kHHMinV);

struct HHTable {
    std::vector<float> v;
    void build(float v_thr) {
        v.assign((size_t)kHHTable * 6, 0.0f);
        for (int i = 0; i < kHHTable; i++) {
            float V = kHHMinV + ((kHHMaxV - kHHMinV) * i) / (kHHTable - 1);
            float am, bm, an;
            if ((13.0f - V + v_thr) == 0.0) am = 1.279999894092334f;
            else am = 0.32f * (13.0f - V + v_thr) / (std::exp((13.0f - V + v_thr) * 0.25f) - 1.0f);
            if ((V - v_thr - 40.0f) == 0.0) bm = 1.400030971093304f;
            else bm = 0.28f * (V - v_thr - 40.0f) / (std::exp((V - v_thr - 40.0f) * 0.2f) - 1.0f);
            float ah = 0.128f * std::exp((17.0f - V + v_thr) * 0.05555555f);
            float bh = 4.0f / (1.0f + std::exp((40.0f - V + v_thr) * 0.2f));
            if ((15.0f - V + v_thr) == 0.0) an = 0.1599999867615418f;
            else an = 0.032f * (15.0f - V + v_thr) / (std::exp((15.0f - V + v_thr) * 0.2f) - 1.0f);
            float bn = 0.5f * std::exp((10.0f - V + v_thr) * 0.025f);
            float *o = &v[(size_t)i * 6];
            o[0] = am; o[1] = bm; o[2] = ah; o[3] = bh; o[4] = an; o[5] = bn;
        }
    }
    /* Get_channel_values: clamped truncating index */
    const float *at(float V) const {
        int position = trunc_i32_x86((V - kHHMinV) * kHHAux);
        if (position < 0) position = 0;
        else if (position > (kHHTable - 1)) position = kHHTable - 1;
        return &v[(size_t)position * 6];
    }
};

template <typename R> struct HH : PopBase<R> {
    R e_exc, e_inh, e_leak, g_leak, c_m, inv_c_m, v_thr, tau_exc, inv_tau_exc, tau_inh,
        inv_tau_inh, tau_nmda, inv_tau_nmda, g_na, g_kd, e_na, e_k;
    NmdaTable nmda;
    HHTable table;
    R cond[3];

    /* Default parameters: HHTimeDrivenModel.cpp GetDefaultParameters (order of p[]):
     * e_exc e_inh e_leak g_leak c_m v_thr tau_exc tau_inh tau_nmda g_na g_kd e_na e_k */
    void set_params(const float *p) {
        e_exc = p[0]; e_inh = p[1]; e_leak = p[2]; g_leak = p[3]; c_m = p[4]; v_thr = p[5];
        tau_exc = p[6]; tau_inh = p[7]; tau_nmda = p[8]; g_na = p[9]; g_kd = p[10];
        e_na = p[11]; e_k = p[12];
        inv_c_m = Num<R>::inv(c_m);
        inv_tau_exc = Num<R>::inv(tau_exc);
        inv_tau_inh = Num<R>::inv(tau_inh);
        inv_tau_nmda = Num<R>::inv(tau_nmda);
        nmda.build(p[0], p[1]);
        table.build(p[5]);
        this->nvars = 8;
        this->ndiff = 4;
    }
    R g_nmda_inf(R V) const { return (R)nmda.get((float)V, (float)e_inh); }
    /* HHTimeDrivenModel.cpp:136-160: gates start at their steady state at e_leak */
    void init_states() {
        const float *c = table.at((float)e_leak);
        float m_inf = c[0] / (c[0] + c[1]);
        float h_inf = c[2] / (c[2] + c[3]);
        float n_inf = c[4] / (c[4] + c[5]);
        R init[8] = {e_leak, (R)m_inf, (R)h_inf, (R)n_inf, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* HHTimeDrivenModel.cpp:197-264: clamps the state it is given, in place. */
    void deriv(R *S, R *D, int /*index*/, R /*elapsed*/) const {
        if (S[0] < e_k) S[0] = e_k;
        if (S[0] > e_na - 1.0f) S[0] = e_na - 1.0f;
        if (S[1] < 0) S[1] = 0.0f;
        if (S[1] > 1) S[1] = 1.0f;
        if (S[2] < 0) S[2] = 0.0f;
        if (S[2] > 1) S[2] = 1.0f;
        if (S[3] < 0) S[3] = 0.0f;
        if (S[3] > 0.72) S[3] = 0.72f;
        const R V = S[0], m = S[1], h = S[2], n = S[3];
        const float *values = table.at((float)V);
        R current = 0.0;
        if (this->EXC) current += S[4] * (e_exc - S[0]);
        if (this->INH) current += S[5] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[6] * ginf * (e_exc - S[0]);
        }
        current += S[7];
        D[0] = (current + g_leak * (e_leak - V) + g_na * m * m * m * h * (e_na - V) +
                g_kd * n * n * n * n * (e_k - V)) * inv_c_m;
        const R am = values[0], bm = values[1], ah = values[2], bh = values[3], an = values[4],
                bn = values[5];
        D[1] = (am * (1.0f - m) - bm * m);
        D[2] = (ah * (1.0f - h) - bh * h);
        D[3] = (an * (1.0f - n) - bn * n);
    }
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[4] < limit) S[4] = 0.0f; else S[4] *= cond[0]; }
        if (this->INH) { if (S[5] < limit) S[5] = 0.0f; else S[5] *= cond[1]; }
        if (this->NMDA) { if (S[6] < limit) S[6] = 0.0f; else S[6] *= cond[2]; }
    }
    /* HHTimeDrivenModel.cpp:187-193: upward crossing of -30 mV, no reset. */
    bool spike(R prev_V, R *S, int index) {
        if (S[0] > -30.0f && prev_V < -30.0f) {
            this->fired(index);
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------------ */
/* logf of the granule's Nernst potential.  The reference calls glibc's      */
/* logf (EgidioGranuleCell_TimeDriven.h:157-159, float overload of log under */
/* `using namespace std`).  That routine is not correctly rounded in every   */
/* case and has no published bit-exact description, so the restatement     */
/* uses the correctly rounded logarithm instead: ln x in double from         */
/* x = 2^e m (m in [sqrt(1/2), sqrt(2))), ln m = 2 atanh((m-1)/(m+1)) by its */
/* series through s^23, e ln 2 with a split ln 2, rounded to float.  The    */
/* pin tests measure how often glibc's logf differs from it (by one ulp).   */
/* ------------------------------------------------------------------------ */
/* g_glibc_logf (orc_set_glibc_logf): the Nernst potential calls the C
 * library's logf as the reference does, to measure how often the two
 * choices change a raster.  Off by default (parity runs). */
int g_glibc_logf = 0;

float log_cr(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (std::isinf(x)) return x;
    uint32_t b;
    std::memcpy(&b, &x, 4);
    int e = 0;
    if (b < 0x00800000u) {
        x *= 8388608.0f;
        std::memcpy(&b, &x, 4);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    const uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
    float mf;
    std::memcpy(&mf, &mb, 4);
    double m = mf;
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    static const double inv_odd[12] = {1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0,
                                       1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0,
                                       1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0,  1.0};
    double p = inv_odd[0];
    for (int k = 1; k < 12; k++) p = p * s2 + inv_odd[k];
    const double lm = (2.0 * s) * p;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    return (float)((double)e * ln2_hi + ((double)e * ln2_lo + lm));
}

/* ------------------------------------------------------------------------ */
/* Egidio granule cell: Kernel/src/neuron_model/EgidioGranuleCell_TimeDriven */
/* .cpp, Kernel/include/neuron_model/EgidioGranuleCell_TimeDriven.h.        */
/* state: V, xNa_f, yNa_f, xNa_r, yNa_r, xNa_p, xK_V, xK_A, yK_A, xK_IR,      */
/*        xK_Ca, xCa, yCa, xK_sl, Ca | g_exc, g_inh, g_nmda, ext_I           */
/* Millisecond model with fixed constants (.cpp:26-58).  Channel rates come  */
/* from a 1024 x 26 table on [-100, 50] mV (.h:402-552, float exp).          */
/* ------------------------------------------------------------------------ */
namespace gr {
const float gMAXNa_f = 0.013f, gMAXNa_r = 0.0005f, gMAXNa_p = 0.0002f, gMAXK_V = 0.003f,
            gMAXK_A = 0.004f, gMAXK_IR = 0.0009f, gMAXK_Ca = 0.004f, gMAXCa = 0.00046f,
            gMAXK_sl = 0.00035f, gLkg1 = 5.68e-5f, gLkg2 = 2.17e-5f, VNa = 87.39f, VK = -84.69f,
            VLkg1 = -58.0f, VLkg2 = -65.0f, V0_xK_Ai = -46.7f, K_xK_Ai = -19.8f,
            V0_yK_Ai = -78.8f, K_yK_Ai = 8.4f, V0_xK_sli = -30.0f, B_xK_sli = 6.0f,
            F = 96485.309f, A = 1e-04f, d = 0.2f, betaCa = 1.5f, Ca0 = 1e-04f, R = 8.3134f,
            cao = 2.0f, Cm = 1.0e-3f, inv_Cm = 1.0f / 1.0e-3f, temper = 30.0f;
const int kTable = 1024, kVars = 26;
const float Max_V = 50.0f, Min_V = -100.0f;
const float aux = (kTable - 1) / (Max_V - Min_V);
inline float q10(float t0) { return std::pow(3.0f, (temper - t0) / 10.0f); }
/* .h:170-178 */
inline float linoid(float x, float y) {
    if (std::fabs(x / y) < 1e-06f) return y * (1 - x / y / 2);
    return x / (std::exp(x / y) - 1);
}
}  // namespace gr

struct GranuleTable {
    std::vector<float> v;
    void build() {
        using namespace gr;
        const float Q10_20 = q10(20.0f), Q10_22 = q10(22.0f), Q10_30 = q10(30.0f), Q10_6_3 = q10(6.3f);
        v.assign((size_t)kTable * kVars, 0.0f);
        for (int i = 0; i < kTable; i++) {
            const float V = Min_V + ((Max_V - Min_V) * i) / (kTable - 1);
            float *t = &v[(size_t)i * kVars];
            float a, b;
            a = Q10_20 * (-0.3f) * linoid(V + 19.0f, -10.0f);
            b = Q10_20 * 12.0f * std::exp(-(V + 44.0f) / 18.182f);
            t[0] = a / (a + b); t[1] = a + b;
            a = Q10_20 * 0.105f * std::exp(-(V + 44.0f) / 3.333f);
            b = Q10_20 * 1.5f / (1.0f + std::exp(-(V + 11.0f) / 5.0f));
            t[2] = a / (a + b); t[3] = a + b;
            a = Q10_20 * (0.00008f - 0.00493f * linoid(V - 4.48754f, -6.81881f));
            b = Q10_20 * (0.04752f + 0.01558f * linoid(V + 43.97494f, 0.10818f));
            t[4] = a / (a + b); t[5] = a + b;
            a = Q10_20 * 0.31836f * std::exp(-(V + 80.0f) / 62.52621f);
            b = Q10_20 * 0.01014f * std::exp((V + 83.3332f) / 16.05379f);
            t[6] = a / (a + b); t[7] = a + b;
            t[8] = 1.0f / (1.0f + std::exp(-(V + 42.0f) / 5.0f));
            a = Q10_30 * (-0.091f) * linoid(V + 42.0f, -5.0f);
            b = Q10_30 * 0.062f * linoid(V + 42.0f, 5.0f);
            t[9] = (a + b) * 0.2f;
            a = Q10_6_3 * (-0.01f) * linoid(V + 25.0f, -10.0f);
            b = Q10_6_3 * 0.125f * std::exp(-0.0125f * (V + 35.0f));
            t[10] = a / (a + b); t[11] = a + b;
            t[12] = 1.0f / (1.0f + std::exp((V - V0_xK_Ai) / K_xK_Ai));
            a = (Q10_20 * 4.88826f) / (1 + std::exp(-(V + 9.17203f) / 23.32708f));
            b = (Q10_20 * 0.99285f) / std::exp((V + 18.27914f) / 19.47175f);
            t[13] = a + b;
            t[14] = 1.0f / (1.0f + std::exp((V - V0_yK_Ai) / K_yK_Ai));
            a = (Q10_20 * 0.11042f) / (1.0f + std::exp((V + 111.33209f) / 12.8433f));
            b = (Q10_20 * 0.10353f) / (1.0f + std::exp(-(V + 49.9537f) / 8.90123f));
            t[15] = a + b;
            a = Q10_20 * 0.13289f * std::exp(-(V + 83.94f) / 24.3902f);
            b = Q10_20 * 0.16994f * std::exp((V + 83.94f) / 35.714f);
            t[16] = a / (a + b); t[17] = a + b;
            t[18] = 0.0015f * std::exp(-V / 11.765f);
            t[19] = 1.0f / (0.00015f * std::exp(-V / 11.765f));
            a = Q10_20 * 0.04944f * std::exp((V + 29.06f) / 15.87301587302f);
            b = Q10_20 * 0.08298f * std::exp(-(V + 18.66f) / 25.641f);
            t[20] = a / (a + b); t[21] = a + b;
            a = Q10_20 * 0.0013f * std::exp(-(V + 48.0f) / 18.183f);
            b = Q10_20 * 0.0013f * std::exp((V + 48.0f) / 83.33f);
            t[22] = a / (a + b); t[23] = a + b;
            t[24] = 1.0f / (1.0f + std::exp(-(V - V0_xK_sli) / B_xK_sli));
            a = Q10_22 * 0.0033f * std::exp((V + 30.0f) / 40.0f);
            b = Q10_22 * 0.0033f * (-(V + 30.0f) / 20.0f);
            t[25] = a + b;
        }
    }
    /* Get_channel_values (.h:563-571) */
    const float *at(float V) const {
        int position = trunc_i32_x86((V - gr::Min_V) * gr::aux);
        if (position < 0) position = 0;
        else if (position > gr::kTable - 1) position = gr::kTable - 1;
        return &v[(size_t)position * gr::kVars];
    }
};

template <typename R> struct Granule : PopBase<R> {
    const float e_exc = 0.0f, e_inh = -80.0f, tau_exc = 0.5f, tau_inh = 10.0f, tau_nmda = 15.0f,
                v_thr = 0.0f;
    float Q10_30 = 1.0f;
    NmdaTable nmda;
    GranuleTable table;
    R cond[3];

    void set_params(const float *) {
        Q10_30 = gr::q10(30.0f);
        nmda.build(e_exc, e_inh);
        table.build();
        this->nvars = 19;
        this->ndiff = 15;
    }
    R g_nmda_inf(R V) const { return (R)nmda.get((float)V, e_inh); }
    /* .cpp:176-235 */
    void init_states() {
        const R init[19] = {-80.0f, 0.00047309535f, 1.0f, 0.00013423511f, 0.96227829f,
                            0.00050020111f, 0.010183001f, 0.15685486f, 0.53565367f, 0.37337035f,
                            0.00012384122f, 0.0021951104f, 0.89509747f, 0.00024031171f, gr::Ca0,
                            0.0f, 0.0f, 0.0f, 0.0f};
        PopBase<R>::init_states(init);
    }
    /* .cpp:382-389: exp(-t/tau) (float division, float exp) */
    void calc_cond(R elapsed) {
        if (sizeof(R) == sizeof(float)) {
            cond[0] = std::exp(-(float)elapsed / tau_exc);
            cond[1] = std::exp(-(float)elapsed / tau_inh);
            cond[2] = expf(-(float)elapsed / tau_nmda);
        } else {
            cond[0] = std::exp(-(double)elapsed / tau_exc);
            cond[1] = std::exp(-(double)elapsed / tau_inh);
            cond[2] = std::exp(-(double)elapsed / tau_nmda);
        }
    }
    /* nernst(Ca, cao, 2, temper) (.h:157-159) */
    R nernst(R ci) const {
        using namespace gr;
        const float k = 1000 * (R_gas() * (temper + 273.15f) / F) / 2.0f;
        if (sizeof(R) == sizeof(float)) {
            const float r = std::fabs(cao / (float)ci);
            return (R)(k * (g_glibc_logf ? ::logf(r) : log_cr(r)));
        }
        return (R)k * std::log(std::fabs((double)cao / (double)ci));
    }
    static float R_gas() { return gr::R; }
    /* .cpp:252-365 */
    void deriv(R *S, R *D, int /*index*/, R /*elapsed*/) const {
        using namespace gr;
        R current = 0.0;
        if (this->EXC) current += S[15] * (e_exc - S[0]);
        if (this->INH) current += S[16] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[17] * ginf * (e_exc - S[0]);
        }
        current += S[18];
        current = (R)(current * 3.34156941e-4);   /* float *= double literal */
        const R pV = S[0];
        const R VCa = nernst(S[14]);
        const float *values = table.at((float)pV);
        const R alpha_xK_Ca = (Q10_30 * 2.5f) / (1.0f + values[18] / S[14]);
        const R beta_xK_Ca = (Q10_30 * 1.5f) / (1.0f + S[14] * values[19]);
        const R xK_Ca_inf = alpha_xK_Ca / (alpha_xK_Ca + beta_xK_Ca);
        const R inv_tau_xK_Ca = (alpha_xK_Ca + beta_xK_Ca);
        const R gNa_f = gMAXNa_f * S[1] * S[1] * S[1] * S[2];
        const R gNa_r = gMAXNa_r * S[3] * S[4];
        const R gNa_p = gMAXNa_p * S[5];
        const R gK_V = gMAXK_V * S[6] * S[6] * S[6] * S[6];
        const R gK_A = gMAXK_A * S[7] * S[7] * S[7] * S[8];
        const R gK_IR = gMAXK_IR * S[9];
        const R gK_Ca = gMAXK_Ca * S[10];
        const R gCa = gMAXCa * S[11] * S[11] * S[12];
        const R gK_sl = gMAXK_sl * S[13];
        D[1] = (values[0] - S[1]) * values[1];
        D[2] = (values[2] - S[2]) * values[3];
        D[3] = (values[4] - S[3]) * values[5];
        D[4] = (values[6] - S[4]) * values[7];
        D[5] = (values[8] - S[5]) * values[9];
        D[6] = (values[10] - S[6]) * values[11];
        D[7] = (values[12] - S[7]) * values[13];
        D[8] = (values[14] - S[8]) * values[15];
        D[9] = (values[16] - S[9]) * values[17];
        D[10] = (xK_Ca_inf - S[10]) * inv_tau_xK_Ca;
        D[11] = (values[20] - S[11]) * values[21];
        D[12] = (values[22] - S[12]) * values[23];
        D[13] = (values[24] - S[13]) * values[25];
        D[14] = (-gCa * (pV - VCa) / (2 * F * A * d) - (betaCa * (S[14] - Ca0)));
        D[0] = (current + gNa_f * (VNa - pV) + gNa_r * (VNa - pV) + gNa_p * (VNa - pV) +
                gK_V * (VK - pV) + gK_A * (VK - pV) + gK_IR * (VK - pV) + gK_Ca * (VK - pV) +
                gCa * (VCa - pV) + gK_sl * (VK - pV) + gLkg1 * (VLkg1 - pV) + gLkg2 * (VLkg2 - pV)) *
               inv_Cm;
    }
    /* .cpp:368-380 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[15] < limit) S[15] = 0.0f; else S[15] *= cond[0]; }
        if (this->INH) { if (S[16] < limit) S[16] = 0.0f; else S[16] *= cond[1]; }
        if (this->NMDA) { if (S[17] < limit) S[17] = 0.0f; else S[17] *= cond[2]; }
    }
    /* .cpp:243-249: upward crossing of v_thr, no reset */
    bool spike(R prev_V, R *S, int index) {
        if (S[0] > v_thr && prev_V < v_thr) {
            this->fired(index);
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------------ */
/* Purkinje cell: Kernel/src/neuron_model/TimeDrivenPurkinjeCell.cpp,        */
/* Kernel/include/neuron_model/TimeDrivenPurkinjeCell.h:396-441.             */
/* state: V, ca, M | g_exc, g_inh, g_nmda, ext_I                             */
/* A spike starts a stereotyped waveform (linear rise to spk_peak, linear    */
/* fall to e_leak over tau_ref) timed by the double LastSpikeTime.           */
/* ------------------------------------------------------------------------ */
namespace pk {
const int kTable = 1024;
const float Max_V = 35.0f, Min_V = -100.0f;
const float aux = (kTable - 1) / (Max_V - Min_V);
}  // namespace pk

struct PurkinjeTable {
    std::vector<float> v;
    void build() {
        v.assign((size_t)pk::kTable * 4, 0.0f);
        for (int i = 0; i < pk::kTable; i++) {
            const float V = pk::Min_V + ((pk::Max_V - pk::Min_V) * i) / (pk::kTable - 1);
            const float alpha_ca = 1.6f / (1 + std::exp(-0.072f * (V - 5.0f)));
            const float beta_ca = (0.02f * (V + 8.9f)) / (std::exp((V + 8.9f) * 0.2f) - 1.0f);
            const float alpha_M = 0.3f / (1 + std::exp((-V - 2.0f) * 0.2f));
            const float beta_M = 0.001f * std::exp((-V - 60.0f) * 0.055555555555555f);
            float *t = &v[(size_t)i * 4];
            t[0] = alpha_ca;
            t[1] = alpha_ca + beta_ca;
            t[2] = alpha_M;
            t[3] = alpha_M + beta_M;
        }
    }
    const float *at(float V) const {
        int position = trunc_i32_x86((V - pk::Min_V) * pk::aux);
        if (position < 0) position = 0;
        else if (position > pk::kTable - 1) position = pk::kTable - 1;
        return &v[(size_t)position * 4];
    }
};

template <typename R> struct Purkinje : PopBase<R> {
    const float g_leak = 0.02f, g_Ca = 0.001f, g_M = 0.75f, c_m = 0.95f, inv_c_m = 1.0f / 0.95f,
                spk_peak = 31.0f;
    const float inv_area = 1.0f / (3.141592f * 0.0015f * 2.0f * 0.0008f);
    const float time_scale = 1000.0f;
    float e_exc, e_inh, v_thr, e_leak, tau_exc, inv_tau_exc, tau_inh, inv_tau_inh, tau_nmda,
        inv_tau_nmda, tau_ref, tau_ref_0_5, inv_tau_ref_0_5;
    NmdaTable nmda;
    PurkinjeTable table;
    R cond[3];

    /* Parameters (order of p[]): e_exc e_inh v_thr e_leak tau_exc tau_inh
     * tau_nmda tau_ref (.cpp:342-404, defaults :423-434) */
    void set_params(const float *p) {
        e_exc = p[0]; e_inh = p[1]; v_thr = p[2]; e_leak = p[3];
        tau_exc = p[4]; inv_tau_exc = (float)(1.0 / p[4]);
        tau_inh = p[5]; inv_tau_inh = (float)(1.0 / p[5]);
        tau_nmda = p[6]; inv_tau_nmda = (float)(1.0 / p[6]);
        tau_ref = p[7]; tau_ref_0_5 = p[7] / 2; inv_tau_ref_0_5 = (float)(1.0 / tau_ref_0_5);
        nmda.build(e_exc, e_inh);
        table.build();
        this->nvars = 7;
        this->ndiff = 3;
    }
    R g_nmda_inf(R V) const { return (R)nmda.get((float)V, e_inh); }
    /* .cpp:139-161 */
    void init_states() {
        const float *values = table.at(e_leak);
        const float c_inf = values[0] / values[1];
        const float M_inf = values[2] / values[3];
        const R init[7] = {e_leak, c_inf, M_inf, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    /* .cpp:270-274 */
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* .cpp:193-239 */
    void deriv(R *S, R *D, int index, R /*elapsed*/) const {
        const R V = S[0], ca = S[1], M = S[2];
        const float last_spike = (float)(time_scale * this->last_spike[index]);
        R current = 0.0;
        if (this->EXC) current += S[3] * (e_exc - S[0]);
        if (this->INH) current += S[4] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[5] * ginf * (e_exc - S[0]);
        }
        current += S[6];
        current = (R)(current * (1e-6 * inv_area));
        if (last_spike >= tau_ref) {
            D[0] = (-g_leak * (V + 70.0f) - g_Ca * ca * ca * (V - 125.0f) - g_M * M * (V + 95.0f) + current) *
                   inv_c_m;
        } else if (last_spike <= tau_ref_0_5) {
            D[0] = (spk_peak - v_thr) * inv_tau_ref_0_5;
        } else {
            D[0] = (e_leak - spk_peak) * inv_tau_ref_0_5;
        }
        const float *values = table.at((float)V);
        D[1] = values[0] - ca * values[1];
        D[2] = values[2] - M * values[3];
    }
    /* .cpp:241-268 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[3] < limit) S[3] = 0.0f; else S[3] *= cond[0]; }
        if (this->INH) { if (S[4] < limit) S[4] = 0.0f; else S[4] *= cond[1]; }
        if (this->NMDA) { if (S[5] < limit) S[5] = 0.0f; else S[5] *= cond[2]; }
    }
    /* .cpp:172-190: the crossing fires; then the clock (advanced this step,
     * or restarted by the spike) imposes the waveform on V. */
    bool spike(R prev_V, R *S, int index) {
        bool f = false;
        if (S[0] >= v_thr && prev_V < v_thr) {
            this->fired(index);
            f = true;
        }
        const double last_spike = this->last_spike[index] * time_scale;
        if (last_spike < tau_ref) {
            if (last_spike <= tau_ref_0_5)
                S[0] = (R)(v_thr + (spk_peak - v_thr) * (last_spike * inv_tau_ref_0_5));
            else
                S[0] = (R)(spk_peak - (spk_peak - e_leak) * ((last_spike - tau_ref_0_5) * inv_tau_ref_0_5));
        } else if ((last_spike - tau_ref) < this->dt_model) {
            S[0] = e_leak;
        }
        return f;
    }
};

/* ------------------------------------------------------------------------ */
/* LIF in SI units: Kernel/src/neuron_model/LIFTimeDrivenModel_IS.cpp         */
/* (second time scale; V, F, S, s).  state: V | g_exc, g_inh, g_nmda, ext_I   */
/* ------------------------------------------------------------------------ */
template <typename R> struct LIF_IS : PopBase<R> {
    float e_exc, e_inh, e_leak, v_thr, c_m, inv_c_m_nF, tau_exc, inv_tau_exc, tau_inh, inv_tau_inh,
        tau_ref, g_leak, g_leak_nS, tau_nmda, inv_tau_nmda;
    NmdaTable nmda;
    R cond[3];

    /* SecondScale (.cpp:56) */
    float time_scale_factor() const { return 1.0f; }
    /* Parameters (order of p[]): e_exc e_inh e_leak v_thr c_m tau_exc tau_inh
     * tau_ref g_leak tau_nmda (.cpp:298-375, defaults :392-404) */
    void set_params(const float *p) {
        e_exc = p[0]; e_inh = p[1]; e_leak = p[2]; v_thr = p[3];
        c_m = p[4]; inv_c_m_nF = (float)(1. / (c_m * 1.e9));
        tau_exc = p[5]; inv_tau_exc = (float)(1.0 / tau_exc);
        tau_inh = p[6]; inv_tau_inh = (float)(1.0 / tau_inh);
        tau_ref = p[7];
        g_leak = p[8]; g_leak_nS = (float)(g_leak * 1.e9);
        tau_nmda = p[9]; inv_tau_nmda = (float)(1.0 / tau_nmda);
        nmda.build(e_exc, e_inh, -62.0f);
        this->nvars = 5;
        this->ndiff = 1;
    }
    R g_nmda_inf(R V) const { return (R)nmda.get((float)V, e_inh); }
    /* .cpp:132-136 */
    void init_states() {
        R init[5] = {e_leak, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* .cpp:166-193: the clock in seconds, compared with tau_ref directly */
    void deriv(R *S, R *D, int index, R elapsed_time) const {
        R current = 0.0;
        if (this->EXC) current += S[1] * (e_exc - S[0]);
        if (this->INH) current += S[2] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[3] * ginf * (e_exc - S[0]);
        }
        const R external_current_nA = (R)(0.001 * S[4]);
        current += external_current_nA;
        const double lst = this->last_spike[index];
        if (lst > tau_ref) {
            D[0] = (current + g_leak_nS * (e_leak - S[0])) * inv_c_m_nF;
        } else if ((lst + elapsed_time) > tau_ref) {
            R fraction = (R)((lst + elapsed_time - tau_ref) / elapsed_time);
            D[0] = fraction * ((current + g_leak_nS * (e_leak - S[0])) * inv_c_m_nF);
        } else {
            D[0] = 0;
        }
    }
    /* .cpp:195-220 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[1] < limit) S[1] = 0.0f; else S[1] *= cond[0]; }
        if (this->INH) { if (S[2] < limit) S[2] = 0.0f; else S[2] *= cond[1]; }
        if (this->NMDA) { if (S[3] < limit) S[3] = 0.0f; else S[3] *= cond[2]; }
    }
    /* .cpp:155-163 */
    bool spike(R, R *S, int index) {
        if (S[0] > v_thr) {
            S[0] = e_leak;
            this->fired(index);
            this->reset_integrator(index);
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------------ */
/* Izhikevich: Kernel/src/neuron_model/IzhikevichTimeDrivenModel.cpp          */
/* state: V, u | g_exc, g_inh, g_nmda, ext_I                                  */
/* ------------------------------------------------------------------------ */
template <typename R> struct Izhikevich : PopBase<R> {
    float a, b, c, d, e_exc, e_inh, c_m, inv_c_m, tau_exc, inv_tau_exc, tau_inh, inv_tau_inh,
        tau_nmda, inv_tau_nmda;
    NmdaTable nmda;
    R cond[3];

    /* Parameters (order of p[]): a b c d e_exc e_inh c_m tau_exc tau_inh
     * tau_nmda (.cpp:298-375, defaults :389-401) */
    void set_params(const float *p) {
        a = p[0]; b = p[1]; c = p[2]; d = p[3]; e_exc = p[4]; e_inh = p[5];
        c_m = p[6]; inv_c_m = (float)(1. / c_m);
        tau_exc = p[7]; inv_tau_exc = (float)(1.0 / tau_exc);
        tau_inh = p[8]; inv_tau_inh = (float)(1.0 / tau_inh);
        tau_nmda = p[9]; inv_tau_nmda = (float)(1.0 / tau_nmda);
        nmda.build(e_exc, e_inh);
        this->nvars = 6;
        this->ndiff = 2;
    }
    R g_nmda_inf(R V) const { return (R)nmda.get((float)V, e_inh); }
    /* .cpp:131-136: the resting equilibrium */
    void init_states() {
        const float Veq = (((b - 5.0f) - std::sqrt((5.0f - b) * (5.0f - b) - 22.4f)) / 0.08f);
        const float Ueq = Veq * b;
        const R init[6] = {Veq, Ueq, 0, 0, 0, 0};
        PopBase<R>::init_states(init);
    }
    /* .cpp:224-231 */
    void calc_cond(R elapsed) {
        cond[0] = Num<R>::decay(elapsed, inv_tau_exc);
        cond[1] = Num<R>::decay(elapsed, inv_tau_inh);
        cond[2] = Num<R>::decay(elapsed, inv_tau_nmda);
    }
    /* .cpp:169-195 */
    void deriv(R *S, R *D, int, R) const {
        R current = 0.0;
        if (this->EXC) current += S[2] * (e_exc - S[0]);
        if (this->INH) current += S[3] * (e_inh - S[0]);
        if (this->NMDA) {
            R ginf = g_nmda_inf(S[0]);
            current += S[4] * ginf * (e_exc - S[0]);
        }
        current += S[5];
        if (S[0] <= 30.0f) {
            D[0] = 0.04f * S[0] * S[0] + 5 * S[0] + 140 - S[1] + (current * inv_c_m);
            D[1] = a * (b * S[0] - S[1]);
        } else {
            D[0] = 0;
            D[1] = 0;
        }
    }
    /* .cpp:197-222 */
    void decay(R *S) const {
        const R limit = (R)1e-9f;
        if (this->EXC) { if (S[2] < limit) S[2] = 0.0f; else S[2] *= cond[0]; }
        if (this->INH) { if (S[3] < limit) S[3] = 0.0f; else S[3] *= cond[1]; }
        if (this->NMDA) { if (S[4] < limit) S[4] = 0.0f; else S[4] *= cond[2]; }
    }
    /* .cpp:156-165 */
    bool spike(R, R *S, int index) {
        if (S[0] > 30.0f) {
            S[0] = c;
            S[1] += d;
            this->fired(index);
            this->reset_integrator(index);
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------------ */
/* Fixed-step integrators over the whole population.                        */
/* ------------------------------------------------------------------------ */
const int kMaxVars = 32;

/* Neurons are independent within an update: the loop runs on the OpenMP
 * threads set by orc_set_threads (one by default), and the spike list is
 * gathered in ascending index afterwards, as a serial loop would build it. */
inline int g_threads = 1;
template <typename M, typename F> void for_each_neuron(M &m, F &&body) {
    m.fired_tmp.assign((size_t)m.N, 0);
    uint8_t *f = m.fired_tmp.data();
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && m.N >= 2048)
    for (int index = 0; index < m.N; index++)
        if (body(index)) f[index] = 1;
    for (int i = 0; i < m.N; i++)
        if (f[i]) m.spikes.push_back(i);
}

/* Euler: Kernel/include/integration_method/Euler.h:123-146, decay factors
 * for elapsed = dt (Euler.h:230-234). */
template <typename M, typename R> void step_euler(M &m) {
    const R dt = m.dt_model;
    for_each_neuron(m, [&](int index) {
        R D[kMaxVars];
        R *S = m.at(index);
        R prev = S[0];
        m.deriv(S, D, index, dt);
        for (int j = 0; j < m.ndiff; j++) S[j] += dt * D[j];
        m.decay(S);
        m.add_elapsed(index);
        return m.spike(prev, S, index);
    });
}

/* RK2 (Heun): Kernel/include/integration_method/RK2.h:133-165, decay factors
 * for elapsed = dt (RK2.h:257-261). */
template <typename M, typename R> void step_rk2(M &m) {
    const R dt = m.dt_model;
    for_each_neuron(m, [&](int index) {
        R D1[kMaxVars], D2[kMaxVars];
        R *S = m.at(index);
        R prev = S[0];
        m.deriv(S, D1, index, dt);
        for (int j = 0; j < m.ndiff; j++) S[j] += D1[j] * dt;
        m.decay(S);
        m.deriv(S, D2, index, dt);
        for (int j = 0; j < m.ndiff; j++) S[j] += (-D1[j] + D2[j]) * dt * 0.5f;
        m.add_elapsed(index);
        return m.spike(prev, S, index);
    });
}

/* RK4: Kernel/include/integration_method/RK4.h:167-231, decay factors for
 * elapsed = dt*0.5 applied twice (RK4.h:354-358). */
template <typename M, typename R> void step_rk4(M &m) {
    const R dt = m.dt_model;
    const R dt_0_5 = dt * 0.5f;
    const R dt_0_16 = dt * 0.166666666667f;
    for_each_neuron(m, [&](int index) {
        R A[kMaxVars], D1[kMaxVars], D2[kMaxVars], D3[kMaxVars], D4[kMaxVars];
        R *S = m.at(index);
        R prev = S[0];
        m.deriv(S, D1, index, dt);
        for (int j = 0; j < m.ndiff; j++) A[j] = S[j] + D1[j] * dt_0_5;
        for (int j = m.ndiff; j < m.nvars; j++) A[j] = S[j];
        m.decay(A);
        m.deriv(A, D2, index, dt);
        for (int j = 0; j < m.ndiff; j++) A[j] = S[j] + D2[j] * dt_0_5;
        m.deriv(A, D3, index, dt);
        for (int j = 0; j < m.ndiff; j++) A[j] = S[j] + D3[j] * dt;
        m.decay(A);
        m.deriv(A, D4, index, dt);
        for (int j = 0; j < m.ndiff; j++)
            S[j] += (D1[j] + 2.0f * (D2[j] + D3[j]) + D4[j]) * dt_0_16;
        for (int j = m.ndiff; j < m.nvars; j++) S[j] = A[j];
        m.add_elapsed(index);
        return m.spike(prev, S, index);
    });
}


/* BDFn: Kernel/include/integration_method/BDFn.h:288-423 (step), 604-669
 * (coefficients), 682-722 (Jacobian), 724-786 (Gauss-Jordan inverse).
 * Per neuron: Euler predictor while the method has no history (state 0),
 * else the difference-table predictor; then up to 10 Newton iterations on
 * the BDF corrector until the max-norm change is <= 1e-6; the order rises by
 * one per step up to bdf_order and restarts at 0 on a spike.  The
 * conductances of the iterate are decayed copies; the stored ones decay once
 * after the update. */
template <typename R> struct BdfStore {
    int order = 1;
    std::vector<std::vector<R>> D, Prev;   /* [order][N*nd], [order-1][N*nd] */
    float C[7][7];
    void init(int ord, int N, int nd) {
        order = ord;
        D.assign(ord, std::vector<R>((size_t)N * nd, 0));
        Prev.assign(ord > 1 ? ord - 1 : 0, std::vector<R>((size_t)N * nd, 0));
        std::memset(C, 0, sizeof(C));
        C[0][0] = 1.0f; C[0][1] = 1.0f;
        C[1][0] = 1.0f; C[1][1] = 1.0f;
        C[2][0] = 2.0f / 3.0f; C[2][1] = 4.0f / 3.0f; C[2][2] = -1.0f / 3.0f;
        C[3][0] = 6.0f / 11.0f; C[3][1] = 18.0f / 11.0f; C[3][2] = -9.0f / 11.0f; C[3][3] = 2.0f / 11.0f;
        C[4][0] = 12.0f / 25.0f; C[4][1] = 48.0f / 25.0f; C[4][2] = -36.0f / 25.0f;
        C[4][3] = 16.0f / 25.0f; C[4][4] = -3.0f / 25.0f;
        C[5][0] = 60.0f / 137.0f; C[5][1] = 300.0f / 137.0f; C[5][2] = -300.0f / 137.0f;
        C[5][3] = 200.0f / 137.0f; C[5][4] = -75.0f / 137.0f; C[5][5] = 12.0f / 137.0f;
        C[6][0] = 60.0f / 147.0f; C[6][1] = 360.0f / 147.0f; C[6][2] = -450.0f / 147.0f;
        C[6][3] = 400.0f / 147.0f; C[6][4] = -225.0f / 147.0f; C[6][5] = 72.0f / 147.0f;
        C[6][6] = -10.0f / 147.0f;
    }
};

/* invermat (BDFn.h:724-786): Gauss-Jordan with row swap on a zero pivot. */
template <typename R> void invermat(int n, const R *a, R *ainv) {
    if (n == 1) {
        ainv[0] = 1.0f / a[0];
        return;
    }
    R la[kMaxVars * kMaxVars], li[kMaxVars * kMaxVars];
    for (int i = 0; i < n * n; i++) { la[i] = a[i]; li[i] = 0; }
    for (int i = 0; i < n; i++) li[i * n + i] = 1.0f;
    for (int s = 0; s < n; s++) {
        R element = la[s * n + s];
        if (element == 0) {
            for (int r = s + 1; r < n; r++) {
                element = la[r * n + s];
                if (element != 0) {
                    for (int m = 0; m < n; m++) {
                        R v = la[r * n + m]; la[r * n + m] = la[s * n + m]; la[s * n + m] = v;
                        v = li[r * n + m]; li[r * n + m] = li[s * n + m]; li[s * n + m] = v;
                    }
                    break;
                }
            }
        }
        R inv_element = 1.0f / element;
        for (int j = 0; j < n; j++) {
            la[s * n + j] *= inv_element;
            li[s * n + j] *= inv_element;
        }
        for (int i = 0; i < n; i++) {
            if (i != s) {
                R coef = -la[i * n + s];
                for (int j = 0; j < n; j++) {
                    la[i * n + j] += la[s * n + j] * coef;
                    li[i * n + j] += li[s * n + j] * coef;
                }
            }
        }
    }
    for (int i = 0; i < n * n; i++) ainv[i] = li[i];
}

template <typename M, typename R> void step_bdf(M &m, BdfStore<R> &b) {
    const int nd = m.ndiff, nv = m.nvars;
    const R dt = m.dt_model;
    for_each_neuron(m, [&](int index) {
        R F[kMaxVars], P[kMaxVars], P1[kMaxVars], Cc[kMaxVars];
        R jac[kMaxVars * kMaxVars], J[kMaxVars * kMaxVars], iJ[kMaxVars * kMaxVars];
        R Jx[kMaxVars], Jpos[kMaxVars], Jneg[kMaxVars];
        R *S = m.at(index);
        R prev = S[0];
        int &st = m.bdf_state[index];
        if (st == 0) {
            m.deriv(S, F, index, dt);
            for (int j = 0; j < nd; j++) P[j] = S[j] + dt * F[j];
        } else {
            for (int j = 0; j < nd; j++) {
                P[j] = S[j];
                for (int i = 0; i < st; i++) P[j] += b.D[i][(size_t)index * nd + j];
            }
        }
        for (int i = nd; i < nv; i++) P[i] = S[i];
        m.decay(P);
        R epsi = 1.0f;
        int k = 0;
        while (epsi > 1e-6 && k < 10) {
            m.deriv(P, F, index, dt);
            for (int j = 0; j < nd; j++) {
                Cc[j] = b.C[st][0] * dt * F[j] + b.C[st][1] * S[j];
                for (int i = 1; i < st; i++) Cc[j] += b.C[st][i + 1] * b.Prev[i - 1][(size_t)index * nd + j];
            }
            /* Jacobian (one-sided, eps = 0.1 dt) at the iterate */
            const R eps = dt * 0.1f;
            const R inv_eps = 1.0f / eps;
            for (int q = 0; q < nv; q++) Jx[q] = P[q];
            m.deriv(Jx, Jpos, index, dt);
            for (int j = 0; j < nd; j++) {
                for (int q = 0; q < nv; q++) Jx[q] = P[q];
                Jx[j] -= eps;
                m.deriv(Jx, Jneg, index, dt);
                for (int z = 0; z < nd; z++) jac[z * nd + j] = (Jpos[z] - Jneg[z]) * inv_eps;
            }
            for (int z = 0; z < nd; z++)
                for (int t = 0; t < nd; t++) {
                    J[z * nd + t] = b.C[st][0] * dt * jac[z * nd + t];
                    if (z == t) J[z * nd + t] -= 1;
                }
            invermat<R>(nd, J, iJ);
            for (int z = 0; z < nd; z++) {
                R aux = 0.0;
                for (int t = 0; t < nd; t++) aux += iJ[z * nd + t] * (P[t] - Cc[t]);
                P1[z] = aux + P[z];
            }
            R aux2 = 0.0f;
            for (int z = 0; z < nd; z++) {
                R aux = std::fabs(P1[z] - P[z]);
                if (aux > aux2) aux2 = aux;
            }
            for (int z = 0; z < nd; z++) P[z] = P1[z];
            epsi = aux2;
            k++;
        }
        if (st < b.order) st++;
        for (int j = 0; j < nd; j++) {
            const size_t q = (size_t)index * nd + j;
            for (int i = st - 1; i > 0; i--) b.D[i][q] = -b.D[i - 1][q];
            b.D[0][q] = P[j] - S[j];
            for (int i = 1; i < st; i++) b.D[i][q] += b.D[i - 1][q];
        }
        if (st > 1) {
            for (int i = st - 2; i > 0; i--)
                for (int j = 0; j < nd; j++) b.Prev[i][(size_t)index * nd + j] = b.Prev[i - 1][(size_t)index * nd + j];
            for (int j = 0; j < nd; j++) b.Prev[0][(size_t)index * nd + j] = S[j];
        }
        for (int j = 0; j < nd; j++) S[j] = P[j];
        m.decay(S);
        m.add_elapsed(index);
        return m.spike(prev, S, index);
    });
}

/* ------------------------------------------------------------------------ */
/* Type-erased population handle.                                           */
/* ------------------------------------------------------------------------ */
struct PopIface {
    virtual ~PopIface() {}
    virtual int nvars() = 0;
    virtual int ndiff() = 0;
    virtual int size() = 0;
    virtual void get_state(int var, double *out) = 0;
    virtual void set_state(int var, const double *in) = 0;
    virtual void add_input(int target, int type, float w) = 0;
    virtual void set_ext(const float *I) = 0;
    virtual int update(int64_t step, int32_t *spikes) = 0;
    virtual bool exc_enabled() = 0;
    virtual float dt_s() = 0;
};

template <typename M, typename R> struct Pop : PopIface {
    M m;
    int integrator = ORC_EULER;
    BdfStore<R> bdf;
    bool round_to_float = false;   /* FP64 Newton mode: state kept in float between steps */
    /* Bifixed methods (Bifixed_{Euler,RK2,RK4}_GPU2.cuh): every update takes
     * n_sub steps of the base method with the step's seconds divided by
     * n_sub in float (BifixedStep_GPU_C_INTERFACE.cuh:157-158); the model
     * scale is that times the time scale (BifixedStep_GPU2.cuh:81-82), the
     * decay factors are those of the sub-step (Bifixed_RK4_GPU2.cuh:237-241),
     * the spike clock gains the sub-step and the spike condition is checked
     * after every sub-step (Bifixed_RK4_GPU2.cuh:143-149); a neuron is
     * reported once per update (a flag per neuron), and the NaN check runs
     * once, after the last sub-step. */
    int n_sub = 1;
    float update_dt_s = 0.0f;
    std::vector<uint8_t> fired_flag;

    Pop(int integ, int bdf_order, const float *params, float dt_s, int n, uint32_t mask) {
        if (integ >= ORC_BIFIXED_EULER) {
            n_sub = bdf_order;
            integ -= ORC_BIFIXED_EULER;
            fired_flag.assign(n, 0);
        }
        integrator = integ;
        update_dt_s = dt_s;
        dt_s = dt_s / (float)n_sub;
        m.set_params(params);
        m.N = n;
        m.dt_s = dt_s;
        /* IntegrationMethodFast.h:191-193: seconds * time scale, in float. */
        float dt_model = dt_s * m.time_scale_factor();
        m.dt_model = (R)dt_model;
        m.EXC = (mask & 1u) != 0;
        m.INH = (mask & 2u) != 0;
        m.NMDA = (mask & 4u) != 0;
        m.EXT_I = (mask & 8u) != 0;
        m.init_states();
        if (integ == ORC_RK4) m.calc_cond((R)(dt_model * 0.5f));
        else m.calc_cond((R)dt_model);
        if (integ == ORC_BDF) {
            bdf.init(bdf_order, n, m.ndiff);
            m.bdf_state.assign(n, 0);
        }
    }
    int nvars() override { return m.nvars; }
    int ndiff() override { return m.ndiff; }
    int size() override { return m.N; }
    void get_state(int var, double *out) override {
        for (int i = 0; i < m.N; i++) out[i] = (double)m.state[(size_t)i * m.nvars + var];
    }
    void set_state(int var, const double *in) override {
        for (int i = 0; i < m.N; i++) m.state[(size_t)i * m.nvars + var] = (R)in[i];
    }
    /* AdExTimeDrivenModel.cpp:74-79: one add per input, in arrival order. */
    void add_input(int target, int type, float w) override {
        R &g = m.state[(size_t)target * m.nvars + m.ndiff + type];
        if (round_to_float) g = (R)((float)g + w);   /* state held in float */
        else g += (R)w;
    }
    /* AdExTimeDrivenModel.cpp:82-89: ext_I holds the total input current. */
    void set_ext(const float *I) override {
        for (int i = 0; i < m.N; i++) m.state[(size_t)i * m.nvars + m.ndiff + 3] = (R)I[i];
    }
    bool exc_enabled() override { return m.EXC; }
    float dt_s() override { return update_dt_s; }
    void step_once() {
        if (integrator == ORC_EULER) step_euler<M, R>(m);
        else if (integrator == ORC_RK2) step_rk2<M, R>(m);
        else if (integrator == ORC_RK4) step_rk4<M, R>(m);
        else step_bdf<M, R>(m, bdf);
    }
    /* AdExTimeDrivenModel.cpp:92-101 + TimeDrivenNeuronModel.cpp:75-88. */
    int update(int64_t /*step*/, int32_t *spikes) override {
        m.spikes.clear();
        if (fired_flag.empty()) {
            step_once();
        } else {
            for (int it = 0; it < n_sub; it++) {
                step_once();
                for (int32_t i : m.spikes) fired_flag[i] = 1;
                m.spikes.clear();
            }
            for (int i = 0; i < m.N; i++)
                if (fired_flag[i]) {
                    m.spikes.push_back(i);
                    fired_flag[i] = 0;
                }
        }
        if (round_to_float)
            for (auto &v : m.state) v = (R)(float)v;
        /* A NaN membrane potential resets that neuron to its initial state. */
        for (int i = 0; i < m.N; i++) {
            R v = m.state[(size_t)i * m.nvars];
            if (v != v) m.reset(i);
        }
        int n = (int)m.spikes.size();
        if (spikes) std::copy(m.spikes.begin(), m.spikes.end(), spikes);
        return n;
    }
};

/* ------------------------------------------------------------------------ */
/* RandomGenerator: Kernel/include/simulation/RandomGenerator.h:35-121,      */
/* Kernel/src/simulation/RandomGenerator.cpp:21-27.  Both recurrences are    */
/* taken modulo 2^31 (the & 0x7fffffff mask).                                */
/* ------------------------------------------------------------------------ */
inline uint32_t lcg_next(uint32_t x) {
    return (uint32_t)(((uint64_t)x * 1103515245ull + 12345ull) & 0x7fffffffull);
}
inline uint32_t seed_next(uint32_t g) {
    return (uint32_t)(((uint64_t)g * 765426813ull + 54321ull) & 0x7fffffffull);
}
inline float lcg_frand(uint32_t x) {
    /* ((float)next_element)/MAX_RAND, MAX_RAND converted to float. */
    return ((float)x) / 2147483647;
}

/* ------------------------------------------------------------------------ */
/* Network: canonical delivery order.                                        */
/* A spike emitted at update s through a synapse of delay d steps arrives     */
/* at t_s + d*dt (double), which falls strictly between updates s+d and      */
/* s+d+1 because float(dt) < dt (TimeEventUpdateNeurons.cpp:78,              */
/* Network.cpp:1311-1317, EventQueue.cpp:121-198 with TIMEEVENT priority      */
/* above PROPAGATEDSPIKE, Event.h:37).  Within one inter-update interval the  */
/* arrival times rise with falling emission step, i.e. with ascending d.     */
/* Each group lists its sources in ascending neuron index                     */
/* (TimeDrivenInternalSpike.cpp:109-160) and each source's synapses of that   */
/* delay in ascending target index (Network.cpp:45-68).  Groups of equal     */
/* arrival time from different source models are taken in the order the     */
/* sources were added.                                                       */
/* ------------------------------------------------------------------------ */
struct Source {
    bool self = true;
    int n = 0;
    std::vector<int64_t> row_ptr;
    std::vector<uint32_t> tgt_type;
    std::vector<float> w;
    /* Poisson device (PoissonGeneratorDeviceVector: one frequency per
     * generator, `rates`, or `rate` for all) */
    std::vector<uint32_t> lcg;
    std::vector<uint8_t> fired_tmp;
    float rate = 0.0f;
    std::vector<float> rates;
    /* spike train (input neurons placed on update steps): the spikes of
     * update k are train_idx[train_ptr[k-1] .. train_ptr[k]) */
    bool train = false;
    std::vector<int64_t> train_ptr;
    std::vector<int32_t> train_idx;
    /* event engine: sources of several networks with the same key are one
     * emitting model (one layer type of the reference); rank orders the
     * models' first update events (the network file's type order).  A train
     * may carry the input spikes' times, in load order. */
    int emitter_key = -1;
    int rank = -1;
    std::vector<std::pair<int32_t, int32_t>> poisson_log;
    std::vector<double> train_time;      /* per schedule entry (train_idx order) */
    std::vector<int64_t> train_order;    /* load rank of each entry */
    std::vector<int64_t> train_load;     /* first update of the run that queues it */
    /* projection from another network's population: its spikes */
    struct Net *from = nullptr;
    /* spikes per step: ring indexed by step % ring */
    std::vector<std::vector<int32_t>> ring;
    /* additive STDP on every synapse of the source (STDPWeightChange,
     * STDPState): traces, last update time, and per target the positions of
     * its plastic input synapses */
    bool stdp = false;
    float a_plus = 0, inv_tau_plus = 0, a_minus = 0, inv_tau_minus = 0, w_max = 0;
    std::vector<float> pre, post;
    std::vector<double> last;
    std::vector<std::vector<int64_t>> in_syn;
};

/* STDPState::SetNewUpdateTime (STDPState.cpp:74-82): both traces decay by
 * ExponentialTable::GetResult(-elapsed / tau) with elapsed = float(new - last). */
inline void stdp_decay(Source &s, int64_t q, double t) {
    const float elapsed = (float)(t - s.last[q]);
    s.pre[q] *= exp_lut(-elapsed * s.inv_tau_plus);
    s.post[q] *= exp_lut(-elapsed * s.inv_tau_minus);
    s.last[q] = t;
}
/* Interconnection::IncrementWeight (Interconnection.h:400-409) */
inline void stdp_increment(Source &s, int64_t q, float inc) {
    s.w[q] += inc;
    if (s.w[q] > s.w_max) s.w[q] = s.w_max;
    else if (s.w[q] < 0.0f) s.w[q] = 0.0f;
}
/* STDPWeightChange::ApplyPreSynapticSpike (STDPWeightChange.cpp:43-55) */
inline void stdp_pre(Source &s, int64_t q, double t) {
    stdp_decay(s, q, t);
    s.pre[q] += 1.0f;
    stdp_increment(s, q, -s.a_minus * s.post[q]);
}
/* STDPWeightChange::ApplyPostSynapticSpike (STDPWeightChange.cpp:57-76) */
inline void stdp_post(Source &s, int64_t q, double t) {
    stdp_decay(s, q, t);
    s.post[q] += 1.0f;
    stdp_increment(s, q, s.a_plus * s.pre[q]);
}

/* Network::RoundPropagationDelay (Network.cpp:1311-1317) of a delay of d
 * steps, minpropagationdelay = 0.0001 s. */
inline double delay_seconds(int d, float dt_s) {
    const double m = 0.0001;
    double r = std::floor((double)d * (double)dt_s * (1.0 / 0.0001) + m * 0.5) * m;
    return r <= 0 ? m : r;
}

/* PoissonGeneratorDeviceVector's step: 0.1f / 1000.0f in float
 * (PoissonGeneratorDeviceVector.cpp:26,137). */
inline double evq_poisson_step() { return (double)(0.1f / 1000.0f); }

struct TraceRec {
    int32_t step;
    int32_t tgt;
    uint8_t type;
    float w;
};

struct Net {
    bool trace = false;
    std::vector<TraceRec> trace_buf;
    PopIface *pop = nullptr;
    int max_delay = 1;
    int ring = 2;
    std::vector<Source> sources;
    /* external drive */
    std::vector<uint32_t> thr;
    float w_ext = 0.0f;
    uint64_t ext_seed = 0;
    bool has_ext = false;
    /* the population's spikes of each step, by step % ring (read by
     * projections into other networks), and the spikes logged by circuit runs */
    std::vector<std::vector<int32_t>> out_ring;
    std::vector<std::pair<int32_t, int32_t>> log;
    /* event engine: the simulation this network runs in, the rank of its
     * population among the models, and an explicit stop time for its next run */
    struct Sim *sim = nullptr;
    int rank = -1;
    int model_rank = -1;
    double stop_time = -1.0;
    /* update times: t[k] = t[k-1] + double(float dt), t[0] = 0
     * (TimeEventUpdateNeurons.cpp:78, Simulation.cpp:228-268) */
    std::vector<double> t;
    double time_of(int64_t k) {
        if (t.empty()) t.push_back(0.0);
        while ((int64_t)t.size() <= k) t.push_back(t.back() + (double)pop->dt_s());
        return t[(size_t)k];
    }
};

/* Counter-based external drive (the synthetic workloads' Poisson drive; not a
 * reference feature): u = mix(neuron*0x9E3779B9 + key(seed, step)), where mix
 * is a 32-bit multiply-xorshift finaliser. */
inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
uint32_t drive_hash(uint64_t seed, int64_t step, uint32_t neuron) {
    uint32_t key = mix32((uint32_t)seed ^ mix32((uint32_t)(seed >> 32) ^ mix32((uint32_t)step)));
    return mix32(neuron * 0x9E3779B9u + key);
}

PopIface *make_pop(int model, int integ, int bdf_order, const float *p, float dt_s, int n,
                   uint32_t mask, bool fp64) {
    if (model == ORC_ADEX) {
        if (fp64) return new Pop<AdEx<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<AdEx<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_LIF) {
        if (fp64) return new Pop<LIF<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<LIF<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_HH) {
        if (fp64) return new Pop<HH<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<HH<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_GRANULE) {
        if (fp64) return new Pop<Granule<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<Granule<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_PURKINJE) {
        if (fp64) return new Pop<Purkinje<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<Purkinje<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_LIF_IS) {
        if (fp64) return new Pop<LIF_IS<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<LIF_IS<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    if (model == ORC_IZHIKEVICH) {
        if (fp64) return new Pop<Izhikevich<double>, double>(integ, bdf_order, p, dt_s, n, mask);
        return new Pop<Izhikevich<float>, float>(integ, bdf_order, p, dt_s, n, mask);
    }
    return nullptr;
}

}  // namespace

/* ======================================================================== */
extern "C" {

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

void orc_exp_table(float *out) { std::memcpy(out, exp_table().v, sizeof(float) * kExpSize); }

void orc_exp_lut(const float *x, float *out, int n) {
    for (int i = 0; i < n; i++) out[i] = exp_lut(x[i]);
}

void orc_nmda_table(float e_exc, float e_inh, float *out) {
    NmdaTable t;
    t.build(e_exc, e_inh);
    std::memcpy(out, t.v, sizeof(float) * kNmdaSize);
}

void orc_hh_table(float v_thr, float *out) {
    HHTable t;
    t.build(v_thr);
    std::memcpy(out, t.v.data(), sizeof(float) * t.v.size());
}

void orc_granule_table(float *out) {
    GranuleTable t;
    t.build();
    std::memcpy(out, t.v.data(), sizeof(float) * t.v.size());
}

void orc_purkinje_table(float *out) {
    PurkinjeTable t;
    t.build();
    std::memcpy(out, t.v.data(), sizeof(float) * t.v.size());
}

void orc_set_glibc_logf(int on) { g_glibc_logf = on ? 1 : 0; }

void orc_log_cr(const float *x, float *out, int n) {
    for (int i = 0; i < n; i++) out[i] = log_cr(x[i]);
}

void orc_lcg_seeds(uint32_t seed, int n, uint32_t *out) {
    /* RandomGenerator::srand then n constructions (RandomGenerator.cpp:21-24). */
    uint32_t g = seed;
    for (int i = 0; i < n; i++) {
        out[i] = g;
        g = seed_next(g);
    }
}

void orc_lcg_frand(uint32_t x0, int n, float *out) {
    uint32_t x = x0;
    for (int i = 0; i < n; i++) {
        x = lcg_next(x);
        out[i] = lcg_frand(x);
    }
}

uint32_t orc_drive_hash(uint64_t seed, int64_t step, uint32_t neuron) {
    return drive_hash(seed, step, neuron);
}

void *orc_pop_create(int model, int integrator, int bdf_order, const float *params,
                     int /*n_params*/, float dt_s, int n_neurons, uint32_t syn_mask, int fp64) {
    if (integrator == ORC_BDF && (bdf_order < 1 || bdf_order > 6)) return nullptr;
    if (integrator < ORC_EULER || integrator > ORC_BIFIXED_RK4) return nullptr;
    if (integrator >= ORC_BIFIXED_EULER && bdf_order < 1) return nullptr;
    /* fp64: 0 float restatement, 1 FP64 oracle, 2 FP64 with the state rounded
     * to float after every step (the HIP path's FP64 Newton mode for HH). */
    PopIface *pi = make_pop(model, integrator, bdf_order, params, dt_s, n_neurons, syn_mask, fp64 != 0);
    if (pi && fp64 == 2) {
        if (model != ORC_HH) {
            delete pi;
            return nullptr;
        }
        ((Pop<HH<double>, double> *)pi)->round_to_float = true;
    }
    return pi;
}

void orc_pop_destroy(void *pop) { delete (PopIface *)pop; }
int orc_pop_nvars(void *pop) { return ((PopIface *)pop)->nvars(); }
void orc_pop_get_state(void *pop, int var, double *out) { ((PopIface *)pop)->get_state(var, out); }
void orc_pop_set_state(void *pop, int var, const double *in) { ((PopIface *)pop)->set_state(var, in); }

void orc_pop_add_inputs(void *pop, const int32_t *target, const uint8_t *type, const float *w,
                        int n) {
    PopIface *p = (PopIface *)pop;
    for (int i = 0; i < n; i++) p->add_input(target[i], type[i], w[i]);
}

void orc_pop_set_ext_current(void *pop, const float *current) {
    ((PopIface *)pop)->set_ext(current);
}

int orc_pop_update(void *pop, int64_t step, int32_t *spikes) {
    return ((PopIface *)pop)->update(step, spikes);
}

void *orc_net_create(void *pop, int max_delay) {
    Net *n = new Net;
    n->pop = (PopIface *)pop;
    n->max_delay = max_delay;
    n->ring = max_delay + 2;
    return n;
}

void orc_net_destroy(void *net) { delete (Net *)net; }

static int add_source(Net *n, Source &s, const int64_t *row_ptr, const uint32_t *tgt_type,
                      const float *w) {
    size_t nrow = (size_t)s.n * n->max_delay + 1;
    s.row_ptr.assign(row_ptr, row_ptr + nrow);
    int64_t nsyn = s.row_ptr.back();
    s.tgt_type.assign(tgt_type, tgt_type + nsyn);
    s.w.assign(w, w + nsyn);
    s.ring.assign(n->ring, std::vector<int32_t>());
    n->sources.push_back(s);
    return (int)n->sources.size() - 1;
}

int orc_net_add_source_self(void *net, const int64_t *row_ptr, const uint32_t *tgt_type,
                            const float *w) {
    Net *n = (Net *)net;
    Source s;
    s.self = true;
    s.n = n->pop->size();
    return add_source(n, s, row_ptr, tgt_type, w);
}

int orc_net_add_source_poisson(void *net, int n_src, const uint32_t *lcg_state, float rate_hz,
                               const int64_t *row_ptr, const uint32_t *tgt_type,
                               const float *w) {
    Net *n = (Net *)net;
    Source s;
    s.self = false;
    s.n = n_src;
    s.lcg.assign(lcg_state, lcg_state + n_src);
    s.rate = rate_hz;
    return add_source(n, s, row_ptr, tgt_type, w);
}

int orc_net_add_source_train(void *net, int n_src, int64_t n_steps, const int64_t *step_ptr,
                             const int32_t *idx, const int64_t *row_ptr, const uint32_t *tgt_type,
                             const float *w) {
    Net *n = (Net *)net;
    Source s;
    s.self = false;
    s.train = true;
    s.n = n_src;
    s.train_ptr.assign(step_ptr, step_ptr + n_steps + 1);
    s.train_idx.assign(idx, idx + (n_steps > 0 ? step_ptr[n_steps] : 0));
    return add_source(n, s, row_ptr, tgt_type, w);
}

void orc_net_set_external(void *net, int n_thr, const uint32_t *thresholds, float w_ext,
                          uint64_t seed) {
    Net *n = (Net *)net;
    n->thr.assign(thresholds, thresholds + n_thr);
    n->w_ext = w_ext;
    n->ext_seed = seed;
    n->has_ext = n_thr > 0;
}

int orc_net_set_stdp(void *net, int source, float max_ltp, float tau_ltp, float max_ltd,
                     float tau_ltd, float max_weight) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size() || !(tau_ltp > 0) || !(tau_ltd > 0)) return -1;
    Source &s = n->sources[(size_t)source];
    s.stdp = true;
    s.a_plus = max_ltp;
    s.a_minus = max_ltd;
    /* STDPState.cpp:24-27 */
    s.inv_tau_plus = 1.0f / tau_ltp;
    s.inv_tau_minus = 1.0f / tau_ltd;
    s.w_max = max_weight;
    const size_t nsyn = s.w.size();
    s.pre.assign(nsyn, 0.0f);
    s.post.assign(nsyn, 0.0f);
    s.last.assign(nsyn, 0.0);   /* ConnectionState.cpp:22 */
    s.in_syn.assign((size_t)n->pop->size(), std::vector<int64_t>());
    for (size_t q = 0; q < nsyn; q++) s.in_syn[s.tgt_type[q] & 0x3fffffffu].push_back((int64_t)q);
    return 0;
}

int orc_net_get_synapses(void *net, int source, float *w, float *pre, float *post, double *last) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size()) return -1;
    Source &s = n->sources[(size_t)source];
    for (size_t q = 0; q < s.w.size(); q++) {
        if (w) w[q] = s.w[q];
        if (pre) pre[q] = s.stdp ? s.pre[q] : 0.0f;
        if (post) post[q] = s.stdp ? s.post[q] : 0.0f;
        if (last) last[q] = s.stdp ? s.last[q] : 0.0;
    }
    return (int)s.w.size();
}

void orc_net_set_trace(void *net, int on) {
    Net *n = (Net *)net;
    n->trace = on != 0;
    n->trace_buf.clear();
}

int64_t orc_net_get_trace(void *net, int32_t *step, int32_t *tgt, uint8_t *type, float *w,
                          int64_t cap) {
    Net *n = (Net *)net;
    const int64_t m = (int64_t)n->trace_buf.size();
    for (int64_t i = 0; i < m && i < cap; i++) {
        step[i] = n->trace_buf[i].step;
        tgt[i] = n->trace_buf[i].tgt;
        type[i] = n->trace_buf[i].type;
        w[i] = n->trace_buf[i].w;
    }
    return m;
}

/* One update k of a network: delivery, external drive, Poisson devices,
 * population update, post-synaptic learning.  Returns the spike count; the
 * spikes are in spk[0..ns) and in out_ring[k % ring]. */
static int net_step(Net *n, int64_t k, std::vector<int32_t> &spk) {
    PopIface *pop = n->pop;
    const int N = pop->size();
    const int D = n->max_delay;
    /* 1. delivery of every event consumed by update k: by delay, then source
     * (in the order added), then spiking source index, then synapse */
    for (int d = 1; d <= D; d++) {
        int64_t s = k - 1 - d;
        if (s < 1) continue;
        for (Source &src : n->sources) {
            const std::vector<int32_t> &list =
                src.from ? src.from->out_ring[s % src.from->ring] : src.ring[s % n->ring];
            for (int32_t id : list) {
                int64_t b = src.row_ptr[(size_t)id * D + d - 1];
                int64_t e = src.row_ptr[(size_t)id * D + d];
                for (int64_t q = b; q < e; q++) {
                    uint32_t tt = src.tgt_type[q];
                    pop->add_input((int)(tt & 0x3fffffffu), (int)(tt >> 30), src.w[q]);
                    if (n->trace)
                        n->trace_buf.push_back({(int32_t)k, (int32_t)(tt & 0x3fffffffu),
                                                (uint8_t)(tt >> 30), src.w[q]});
                    /* PropagatedSpikeGroup.cpp:76-97: the learning rule runs
                     * after ProcessInputSpike, at the arrival time */
                    if (src.stdp) stdp_pre(src, q, n->time_of(s) + delay_seconds(d, pop->dt_s()));
                }
            }
        }
    }
    /* 2. external drive: after the recurrent events of the interval */
    if (n->has_ext) {
        for (int i = 0; i < N; i++) {
            uint32_t u = drive_hash(n->ext_seed, k, (uint32_t)i);
            int cnt = 0;
            while (cnt < (int)n->thr.size() && u < n->thr[cnt]) cnt++;
            for (int c = 0; c < cnt; c++) pop->add_input(i, 0, n->w_ext);
        }
    }
    /* 3. Poisson devices at update k (PoissonGeneratorDeviceVector.cpp:56-70):
     * a source fires when frequency * step (float * double) exceeds its
     * next frand(); frand is not drawn when the frequency is not positive. */
    for (Source &src : n->sources) {
        if (src.self || src.from) continue;
        std::vector<int32_t> &list = src.ring[k % n->ring];
        list.clear();
        if (src.train) {
            /* InputSpikeNeuronModel: the spikes the host placed on update k */
            const int64_t steps = (int64_t)src.train_ptr.size() - 1;
            if (k >= 1 && k <= steps)
                for (int64_t j = src.train_ptr[(size_t)k - 1]; j < src.train_ptr[(size_t)k]; j++)
                    list.push_back(src.train_idx[(size_t)j]);
            continue;
        }
        const double step_s = evq_poisson_step();
        src.fired_tmp.assign((size_t)src.n, 0);
        uint8_t *fl = src.fired_tmp.data();
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && src.n >= 4096)
        for (int i = 0; i < src.n; i++) {
            const float f = src.rates.empty() ? src.rate : src.rates[(size_t)i];
            if (f > 0.0f) {
                src.lcg[i] = lcg_next(src.lcg[i]);
                if ((double)f * step_s > (double)lcg_frand(src.lcg[i])) fl[i] = 1;
            }
        }
        for (int i = 0; i < src.n; i++)
            if (fl[i]) list.push_back(i);
    }
    /* 4. population update k */
    int ns = pop->update(k, spk.data());
    for (Source &src : n->sources) {
        if (!src.self) continue;
        src.ring[k % n->ring].assign(spk.begin(), spk.begin() + ns);
    }
    if (n->out_ring.empty()) n->out_ring.assign(n->ring, std::vector<int32_t>());
    n->out_ring[k % n->ring].assign(spk.begin(), spk.begin() + ns);
    /* post-synaptic learning of the neurons that fired, at the update
     * time (TimeDrivenInternalSpike.cpp:94-103) */
    for (Source &src : n->sources) {
        if (!src.stdp) continue;
        const double tk = n->time_of(k);
        for (int q = 0; q < ns; q++)
            for (int64_t syn : src.in_syn[(size_t)spk[q]]) stdp_post(src, syn, tk);
    }
    return ns;
}

}  // extern "C"

namespace {

/* ======================================================================== */
/* Event engine: the reference's event loop on one queue.                    */
/*                                                                          */
/* Every time-driven model (a population, a Poisson generator layer) has a  */
/* TimeEventUpdateNeurons stream at its own step (TimeEventUpdateNeurons.cpp */
/* 45-79, first event at one step, Simulation.cpp:228-268); an update       */
/* gathers its spiking neurons, in ascending index, into one                */
/* PropagatedSpikeGroup per delay of the model's delay structure, queueing  */
/* a group as soon as it holds 1,024 sources and the rest afterwards in     */
/* ascending delay (TimeDrivenInternalSpike.cpp:66-104,163-178,             */
/* PropagatedSpikeGroup.h:54); a group delivers, source by source, the      */
/* source's synapses of its delay (PropagatedSpikeGroup.cpp:66-108).  Input */
/* neurons' spikes are InputSpike events that chain one PropagatedSpike     */
/* per delay of the neuron (InputSpike.cpp, PropagatedSpike.cpp:119-160).   */
/* All of them share one binary heap (EventQueue.cpp:121-198) ordered by   */
/* time, then EventPriority (Event.h:37); equal keys leave in the order the */
/* heap's layout gives.  A run is RunSimulationSlot (Simulation.cpp:322-373):*/
/* a StopSimulationEvent at the stop time, then events until it leaves.     */
/* ======================================================================== */
namespace evq {

enum { PRIO_STOP = 1, PRIO_SPIKE = 8, PRIO_TIME = 10 };   /* Event.h:37 */
enum { K_TIME = 0, K_GROUP = 1, K_INPUT = 2, K_PROP = 3, K_STOP = 4 };
enum { E_POP = 0, E_POISSON = 1, E_TRAIN = 2 };

struct Item {
    double t;
    int prio;
    int32_t ev;
};

/* EventQueue::InsertEvent / RemoveEvent restricted to one queue. */
struct Heap {
    std::vector<Item> e{Item{0.0, 0, -1}};
    size_t size() const { return e.size() - 1; }
    void insert(const Item &it) {
        e.push_back(it);
        for (size_t c = size(); c > 1 && (e[c / 2].t > e[c].t || (e[c / 2].t == e[c].t && e[c / 2].prio < e[c].prio));
             c /= 2)
            std::swap(e[c], e[c / 2]);
    }
    bool remove(Item &out) {
        const size_t n = e.size();
        if (n > 2) {
            out = e[1];
            e[1] = e[size()];
            e.pop_back();
            size_t p = 1, c;
            for (c = p * 2; c < size(); p = c, c = p * 2) {
                if (e[c].t > e[c + 1].t || (e[c].t == e[c + 1].t && e[c].prio < e[c + 1].prio)) c++;
                if (e[c].t < e[p].t || (e[c].t == e[p].t && e[c].prio > e[p].prio)) std::swap(e[p], e[c]);
                else break;
            }
            if (c == size() && (e[c].t < e[p].t || (e[c].t == e[p].t && e[c].prio > e[p].prio)))
                std::swap(e[p], e[c]);
            return true;
        }
        if (n == 2) {
            out = e[1];
            e.pop_back();
            return true;
        }
        return false;
    }
};

struct Ev {
    int kind = 0;
    int em = 0;
    int d = 0;          /* delay (steps) of a group */
    int32_t neuron = 0; /* input spike: source neuron */
    int j = 0;          /* propagated spike: index into the neuron's delays */
    std::vector<int32_t> srcs;
};

struct Emitter {
    int kind = E_POP;
    int rank = 0;
    int key = -1;
    Net *net = nullptr;                              /* population: its network */
    std::vector<std::pair<Net *, int>> copies;      /* (network, source index) it feeds */
    double h = 0.0;                                  /* step (s) of its update events */
    int64_t k = 1;                                   /* update index of its next update */
    int n = 0;
    int max_delay = 1;
    float dt_s = 1e-4f;                              /* step of the networks it feeds */
    std::vector<int> delays;                         /* model delay structure, ascending */
    std::vector<std::vector<int>> ndel;              /* delays of each neuron, ascending */
};

/* PoissonGeneratorDeviceVector samples every sampling_period = 0.1f ms of
 * its millisecond time scale: 0.1f / 1000.0f in float, held as a double
 * (PoissonGeneratorDeviceVector.cpp:26,137). */
inline double poisson_step() { return (double)(0.1f / 1000.0f); }

}  // namespace evq

struct Sim {
    std::vector<Net *> nets;
    std::vector<evq::Emitter> em;
    evq::Heap heap;
    std::vector<evq::Ev> pool;
    std::vector<int32_t> free_ev;
    int64_t next_first = 1;
    int32_t new_ev(int kind, int em) {
        int32_t id;
        if (!free_ev.empty()) {
            id = free_ev.back();
            free_ev.pop_back();
        } else {
            id = (int32_t)pool.size();
            pool.emplace_back();
        }
        evq::Ev &e = pool[(size_t)id];
        e.kind = kind;
        e.em = em;
        e.d = 0;
        e.neuron = 0;
        e.j = 0;
        e.srcs.clear();
        return id;
    }
};

namespace evq {

/* The delays (steps) of neuron i of a source, in ascending order. */
inline void add_neuron_delays(Emitter &E, const Net *n, const Source &s) {
    const int D = n->max_delay;
    for (int i = 0; i < E.n; i++)
        for (int d = 1; d <= D; d++)
            if (s.row_ptr[(size_t)i * D + d] > s.row_ptr[(size_t)i * D + d - 1]) E.ndel[(size_t)i].push_back(d);
}

inline void finish_delays(Emitter &E) {
    std::vector<char> any((size_t)E.max_delay + 1, 0);
    for (auto &v : E.ndel) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (int d : v) any[(size_t)d] = 1;
    }
    for (int d = 1; d <= E.max_delay; d++)
        if (any[(size_t)d]) E.delays.push_back(d);
}

/* Delivery of neuron `id`'s synapses of delay d in every network the
 * emitter feeds, at event time t (ProcessInputSpike + pre-synaptic STDP). */
inline void deliver(Emitter &E, int32_t id, int d, double t) {
    for (auto &cp : E.copies) {
        Net *n = cp.first;
        Source &src = n->sources[(size_t)cp.second];
        const int D = n->max_delay;
        const int64_t b = src.row_ptr[(size_t)id * D + d - 1];
        const int64_t e = src.row_ptr[(size_t)id * D + d];
        const int32_t k_next = (int32_t)n->sim->em[(size_t)n->rank].k;
        for (int64_t q = b; q < e; q++) {
            const uint32_t tt = src.tgt_type[(size_t)q];
            n->pop->add_input((int)(tt & 0x3fffffffu), (int)(tt >> 30), src.w[(size_t)q]);
            if (n->trace) n->trace_buf.push_back({k_next, (int32_t)(tt & 0x3fffffffu), (uint8_t)(tt >> 30), src.w[(size_t)q]});
            if (src.stdp) stdp_pre(src, q, t);
        }
    }
}

/* TimeDrivenInternalSpike::ProcessEvent: groups of the spiking neurons. */
inline void queue_groups(Sim &S, int e, const std::vector<int32_t> &spk, double t) {
    Emitter &E = S.em[(size_t)e];
    std::vector<int32_t> open((size_t)E.max_delay + 1, -1);
    auto push = [&](int d, int32_t id) {
        if (open[(size_t)d] < 0) {
            open[(size_t)d] = S.new_ev(K_GROUP, e);
            S.pool[(size_t)open[(size_t)d]].d = d;
        }
        const int32_t g = open[(size_t)d];
        S.pool[(size_t)g].srcs.push_back(id);
        if (S.pool[(size_t)g].srcs.size() >= 1024) {
            S.heap.insert(Item{t + delay_seconds(d, E.dt_s), PRIO_SPIKE, g});
            open[(size_t)d] = -1;
        }
    };
    for (int32_t id : spk)
        for (int d : E.ndel[(size_t)id]) push(d, id);
    for (int d : E.delays)
        if (open[(size_t)d] >= 0) S.heap.insert(Item{t + delay_seconds(d, E.dt_s), PRIO_SPIKE, open[(size_t)d]});
}

}  // namespace evq

/* Population update k of network n at time t (external drive, update,
 * post-synaptic learning); returns the spikes. */
static void engine_update_pop(Net *n, int64_t k, double t, std::vector<int32_t> &spk) {
    PopIface *pop = n->pop;
    const int N = pop->size();
    if (n->has_ext) {
        for (int i = 0; i < N; i++) {
            uint32_t u = drive_hash(n->ext_seed, k, (uint32_t)i);
            int cnt = 0;
            while (cnt < (int)n->thr.size() && u < n->thr[cnt]) cnt++;
            for (int c = 0; c < cnt; c++) pop->add_input(i, 0, n->w_ext);
        }
    }
    spk.resize((size_t)N);
    const int ns = pop->update(k, spk.data());
    spk.resize((size_t)ns);
    for (int32_t i : spk) n->log.push_back({(int32_t)k, i});
    for (Source &src : n->sources) {
        if (!src.stdp) continue;
        for (int32_t i : spk)
            for (int64_t syn : src.in_syn[(size_t)i]) stdp_post(src, syn, t);
    }
}

/* Input spikes of the spike trains queued when the simulation starts (load
 * update 0, before the models' update events, as FileInputSpikeDriver's) or
 * by the run starting at update `first` (Simulation_API::
 * AddExternalSpikeActivity), merged across trains by load rank.  Without
 * explicit times a spike sits at its update's time, is loaded at the start,
 * and the schedule order is its load order. */
static void engine_queue_inputs(Sim &S, int64_t first, bool initial) {
    using namespace evq;
    struct Rec {
        int64_t rank;
        int em;
        double t;
        int32_t neuron;
    };
    std::vector<Rec> recs;
    int64_t base = 0;
    for (size_t e = 0; e < S.em.size(); e++) {
        Emitter &E = S.em[e];
        if (E.kind != E_TRAIN) continue;
        const Source &s = E.copies[0].first->sources[(size_t)E.copies[0].second];
        const Net *n0 = E.copies[0].first;
        const int64_t steps = (int64_t)s.train_ptr.size() - 1;
        double t = 0.0;
        for (int64_t k = 1; k <= steps; k++) {
            t += (double)n0->pop->dt_s();
            for (int64_t j = s.train_ptr[(size_t)k - 1]; j < s.train_ptr[(size_t)k]; j++) {
                const int64_t load = s.train_load.empty() ? 0 : s.train_load[(size_t)j];
                if (initial ? load != 0 : load != first) continue;
                Rec r;
                r.rank = s.train_order.empty() ? base + j : s.train_order[(size_t)j];
                r.em = (int)e;
                r.t = s.train_time.empty() ? t : s.train_time[(size_t)j];
                r.neuron = s.train_idx[(size_t)j];
                recs.push_back(r);
            }
        }
        base += (int64_t)s.train_idx.size();
    }
    std::stable_sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) { return a.rank < b.rank; });
    for (const Rec &r : recs) {
        if (S.em[(size_t)r.em].ndel[(size_t)r.neuron].empty()) continue;
        const int32_t id = S.new_ev(K_INPUT, r.em);
        S.pool[(size_t)id].neuron = r.neuron;
        S.heap.insert(Item{r.t, PRIO_SPIKE, id});
    }
}

/* Builds the simulation of a set of networks: one emitter per population
 * and per Poisson / spike-train model; the initial events. */
static Sim *engine_create(Net **nets, int n_nets) {
    using namespace evq;
    Sim *S = new Sim;
    S->nets.assign(nets, nets + n_nets);
    for (int i = 0; i < n_nets; i++) {
        Net *n = nets[i];
        n->sim = S;
        n->rank = i;
        Emitter E;
        E.kind = E_POP;
        E.rank = n->model_rank >= 0 ? n->model_rank : i;
        E.net = n;
        E.h = (double)n->pop->dt_s();
        E.n = n->pop->size();
        E.max_delay = n->max_delay;
        E.dt_s = n->pop->dt_s();
        E.ndel.assign((size_t)E.n, {});
        S->em.push_back(E);
    }
    int next_key = 1 << 20;
    for (int i = 0; i < n_nets; i++) {
        Net *n = nets[i];
        for (size_t si = 0; si < n->sources.size(); si++) {
            Source &s = n->sources[si];
            int e = -1;
            if (s.self) e = i;
            else if (s.from) e = s.from->rank;   /* projections: the source network is in the circuit */
            else {
                const int key = s.emitter_key >= 0 ? s.emitter_key : next_key++;
                for (size_t q = (size_t)n_nets; q < S->em.size(); q++)
                    if (S->em[q].key == key) e = (int)q;
                if (e < 0) {
                    Emitter E;
                    E.kind = s.train ? E_TRAIN : E_POISSON;
                    E.key = key;
                    E.rank = s.rank >= 0 ? s.rank : 1000 + (int)S->em.size();
                    E.h = s.train ? 0.0 : poisson_step();
                    E.n = s.n;
                    E.max_delay = n->max_delay;
                    E.dt_s = n->pop->dt_s();
                    E.ndel.assign((size_t)E.n, {});
                    S->em.push_back(E);
                    e = (int)S->em.size() - 1;
                }
            }
            S->em[(size_t)e].copies.push_back({n, (int)si});
            add_neuron_delays(S->em[(size_t)e], n, s);
        }
    }
    for (Emitter &E : S->em) finish_delays(E);
    /* Simulation::InitSimulation: the input spikes first (GetInput), then one
     * update event per time-driven model in type order. */
    engine_queue_inputs(*S, 0, true);
    std::vector<int> order;
    for (size_t e = 0; e < S->em.size(); e++)
        if (S->em[e].kind != E_TRAIN) order.push_back((int)e);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return S->em[(size_t)a].rank < S->em[(size_t)b].rank; });
    for (int e : order) S->heap.insert(Item{S->em[(size_t)e].h, PRIO_TIME, S->new_ev(K_TIME, e)});
    return S;
}

/* RunSimulationSlot(stop): updates first..first+n_steps-1 of the
 * populations; the stop time defaults to the last update's time. */
static int engine_run(Sim *S, int64_t first, int n_steps, double stop) {
    using namespace evq;
    if (first != S->next_first) return -1;
    const int64_t last = first + n_steps - 1;
    /* input spikes loaded for this run, then its stop event; stop = -2: no
     * stop event (the run is continued by the next one as one reference run) */
    engine_queue_inputs(*S, first, false);
    const bool with_stop = stop != -2.0;
    if (stop < 0) {
        const Emitter &P = S->em[0];
        double t = 0.0;
        for (int64_t k = 1; k <= last; k++) t += P.h;
        stop = t;
    }
    if (with_stop) S->heap.insert(Item{stop, PRIO_STOP, S->new_ev(K_STOP, -1)});
    std::vector<int32_t> spk;
    Item it;
    while (S->heap.size() > 0) {
        /* the window after update `last` ends where the populations' next
         * update starts */
        const Item top = S->heap.e[1];
        const Ev &tev = S->pool[(size_t)top.ev];
        if (tev.kind == K_TIME && S->em[(size_t)tev.em].kind == E_POP && S->em[(size_t)tev.em].k > last) break;
        S->heap.remove(it);
        const int kind = S->pool[(size_t)it.ev].kind;
        const int e = S->pool[(size_t)it.ev].em;
        if (kind == K_STOP) {
            S->free_ev.push_back(it.ev);
            break;
        }
        if (kind == K_TIME) {
            Emitter &E = S->em[(size_t)e];
            const int64_t k = E.k;
            if (E.kind == E_POP) {
                engine_update_pop(E.net, k, it.t, spk);
            } else {
                /* PoissonGeneratorDeviceVector::UpdateState (.cpp:56-70) with
                 * the first copy's generators; the other copies follow */
                Source &s0 = E.copies[0].first->sources[(size_t)E.copies[0].second];
                s0.fired_tmp.assign((size_t)s0.n, 0);
                uint8_t *fl = s0.fired_tmp.data();
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && s0.n >= 4096)
                for (int i = 0; i < s0.n; i++) {
                    const float f = s0.rates.empty() ? s0.rate : s0.rates[(size_t)i];
                    if (f > 0.0f) {
                        s0.lcg[(size_t)i] = lcg_next(s0.lcg[(size_t)i]);
                        if ((double)f * E.h > (double)lcg_frand(s0.lcg[(size_t)i])) fl[i] = 1;
                    }
                }
                spk.clear();
                for (int i = 0; i < s0.n; i++)
                    if (fl[i]) spk.push_back(i);
                for (size_t c = 1; c < E.copies.size(); c++)
                    E.copies[c].first->sources[(size_t)E.copies[c].second].lcg = s0.lcg;
                for (auto &cp : E.copies) {
                    Source &s = cp.first->sources[(size_t)cp.second];
                    for (int32_t i : spk) s.poisson_log.push_back({(int32_t)k, i});
                }
            }
            queue_groups(*S, e, spk, it.t);
            E.k = k + 1;
            S->heap.insert(Item{it.t + E.h, PRIO_TIME, it.ev});
            continue;
        }
        Ev &ev = S->pool[(size_t)it.ev];
        Emitter &E = S->em[(size_t)e];
        if (kind == K_GROUP) {
            for (int32_t id : ev.srcs) deliver(E, id, ev.d, it.t);
            S->free_ev.push_back(it.ev);
        } else if (kind == K_INPUT) {
            /* InputSpike::ProcessEvent: the first delay of the neuron */
            ev.kind = K_PROP;
            ev.j = 0;
            S->heap.insert(Item{it.t + delay_seconds(E.ndel[(size_t)ev.neuron][0], E.dt_s), PRIO_SPIKE, it.ev});
        } else {
            /* PropagatedSpike::ProcessEvent: this delay, then the next one at
             * time - delay_j + delay_{j+1} */
            const std::vector<int> &dl = E.ndel[(size_t)ev.neuron];
            deliver(E, ev.neuron, dl[(size_t)ev.j], it.t);
            if (ev.j + 1 < (int)dl.size()) {
                const double t2 = it.t - delay_seconds(dl[(size_t)ev.j], E.dt_s) + delay_seconds(dl[(size_t)ev.j + 1], E.dt_s);
                ev.j++;
                S->heap.insert(Item{t2, PRIO_SPIKE, it.ev});
            } else {
                S->free_ev.push_back(it.ev);
            }
        }
    }
    S->next_first = last + 1;
    return 0;
}

}  // namespace

extern "C" {

/* Order of delivery: 1 (default) the reference's event queue, 0 the
 * canonical order (delay, source, spiking index, synapse). */
static int g_order = 1;
void orc_set_order(int mode) { g_order = mode; }

/* Sources with the same key (across networks) are one model; rank orders
 * the models' first update events. */
int orc_net_set_source_model(void *net, int source, int key, int rank) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size() || n->sim) return -1;
    n->sources[(size_t)source].emitter_key = key;
    n->sources[(size_t)source].rank = rank;
    return 0;
}

/* Times (s), load ranks and load runs of a spike train's input spikes,
 * aligned with its schedule's entries (null: the update's time, schedule
 * order, queued at the start). */
int orc_net_set_train_times(void *net, int source, int64_t n_spikes, const double *time, const int64_t *order,
                            const int64_t *load_first) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size() || !n->sources[(size_t)source].train) return -1;
    Source &s = n->sources[(size_t)source];
    if (n_spikes != (int64_t)s.train_idx.size()) return -1;
    if (time) s.train_time.assign(time, time + n_spikes);
    else s.train_time.clear();
    if (order) s.train_order.assign(order, order + n_spikes);
    else s.train_order.clear();
    if (load_first) s.train_load.assign(load_first, load_first + n_spikes);
    else s.train_load.clear();
    return 0;
}

/* The event heap alone: op[i] = 1 inserts an element (time t[i], priority
 * prio[i], id = number of earlier inserts), op[i] = 0 removes the top and
 * writes its id to removed[]; returns the number removed. */
int64_t orc_heap_replay(int64_t n_ops, const int8_t *op, const double *t, const int32_t *prio,
                        int32_t *removed) {
    evq::Heap h;
    int32_t next = 0;
    int64_t r = 0;
    for (int64_t i = 0; i < n_ops; i++) {
        if (op[i]) {
            h.insert(evq::Item{t[i], prio[i], next++});
        } else {
            evq::Item it;
            if (h.remove(it)) removed[r++] = it.ev;
        }
    }
    return r;
}

/* Replaces a spike train's schedule (its times, ranks and load runs reset). */
int orc_net_set_train(void *net, int source, int64_t n_steps, const int64_t *step_ptr, const int32_t *idx) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size() || !n->sources[(size_t)source].train) return -1;
    Source &s = n->sources[(size_t)source];
    s.train_ptr.assign(step_ptr, step_ptr + n_steps + 1);
    s.train_idx.assign(idx, idx + (n_steps > 0 ? step_ptr[n_steps] : 0));
    s.train_time.clear();
    s.train_order.clear();
    s.train_load.clear();
    return 0;
}

/* The event queue of the simulation the network runs in, as stored (heap
 * array order): time, priority, kind (0 update, 1 group, 2 input spike,
 * 3 propagated spike, 4 stop), model and delay (steps; propagated spikes:
 * the delay's index).  Returns the number of events. */
int64_t orc_net_event_queue(void *net, double *t, int32_t *prio, int32_t *kind, int32_t *em, int32_t *d,
                            int64_t cap) {
    Net *n = (Net *)net;
    if (!n->sim) return 0;
    const auto &h = n->sim->heap.e;
    const int64_t m = (int64_t)h.size() - 1;
    for (int64_t i = 0; i < m && i < cap; i++) {
        const evq::Item &it = h[(size_t)i + 1];
        const evq::Ev &ev = n->sim->pool[(size_t)it.ev];
        t[i] = it.t;
        prio[i] = it.prio;
        kind[i] = ev.kind;
        em[i] = ev.em < 0 ? 0 : ev.em;
        d[i] = ev.kind == evq::K_PROP ? ev.j : ev.d;
    }
    return m;
}

/* Rank of the network's population among the models (default: its
 * position in the circuit). */
int orc_net_set_model_rank(void *net, int rank) {
    Net *n = (Net *)net;
    if (n->sim) return -1;
    n->model_rank = rank;
    return 0;
}

/* Stop time (s) of the network's (or circuit's) next run. */
void orc_net_set_stop_time(void *net, double t) { ((Net *)net)->stop_time = t; }

int64_t orc_net_run(void *net, int64_t first_step, int n_steps, int32_t *spike_step,
                    int32_t *spike_idx, int64_t cap) {
    Net *n = (Net *)net;
    int64_t total = 0;
    if (g_order == 1) {
        if (!n->sim) {
            for (const Source &s : n->sources)
                if (s.from) return -2;   /* projections run in a circuit */
            engine_create(&n, 1);
        }
        const size_t l0 = n->log.size();
        if (engine_run(n->sim, first_step, n_steps, n->stop_time) != 0) return -2;
        n->stop_time = -1.0;
        for (size_t q = l0; q < n->log.size(); q++) {
            if (total < cap) {
                spike_step[total] = n->log[q].first;
                spike_idx[total] = n->log[q].second;
            }
            total++;
        }
        n->log.resize(l0);
        return total > cap ? -1 : total;
    }
    std::vector<int32_t> spk(n->pop->size());
    for (int64_t k = first_step; k < first_step + n_steps; k++) {
        const int ns = net_step(n, k, spk);
        for (int q = 0; q < ns; q++) {
            if (total < cap) {
                spike_step[total] = (int32_t)k;
                spike_idx[total] = spk[q];
            }
            total++;
        }
    }
    return total > cap ? -1 : total;
}

/* Projection of `source`'s population onto `net`'s (its CSR rows are the
 * source population's neurons).  Both networks must have the same maximum
 * delay and step. */
int orc_net_add_link(void *net, void *source, const int64_t *row_ptr, const uint32_t *tgt_type,
                     const float *w) {
    Net *n = (Net *)net;
    Net *f = (Net *)source;
    if (f->max_delay != n->max_delay) return -1;
    Source s;
    s.self = false;
    s.from = f;
    s.n = f->pop->size();
    if (f->out_ring.empty()) f->out_ring.assign(f->ring, std::vector<int32_t>());
    return add_source(n, s, row_ptr, tgt_type, w);
}

/* One frequency per generator of a Poisson source. */
int orc_net_set_poisson_rates(void *net, int source, const float *rates) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size()) return -1;
    Source &s = n->sources[(size_t)source];
    if (s.self || s.from) return -1;
    s.rates.assign(rates, rates + s.n);
    return 0;
}

/* Lockstep run of several networks: update k of every network (in the order
 * given) before update k + 1 of any; the spikes go to each network's log. */
void orc_circuit_run(void **nets, int n_nets, int64_t first_step, int n_steps) {
    if (g_order == 1) {
        Net *n0 = (Net *)nets[0];
        if (!n0->sim) engine_create((Net **)nets, n_nets);
        engine_run(n0->sim, first_step, n_steps, n0->stop_time);
        n0->stop_time = -1.0;
        return;
    }
    std::vector<std::vector<int32_t>> spk((size_t)n_nets);
    for (int i = 0; i < n_nets; i++) spk[(size_t)i].resize(((Net *)nets[i])->pop->size());
    for (int64_t k = first_step; k < first_step + n_steps; k++)
        for (int i = 0; i < n_nets; i++) {
            Net *n = (Net *)nets[i];
            const int ns = net_step(n, k, spk[(size_t)i]);
            for (int q = 0; q < ns; q++) n->log.push_back({(int32_t)k, spk[(size_t)i][(size_t)q]});
        }
}

/* Spikes of a Poisson source's generators at each of its updates (event
 * engine); moved out when cap holds them all; returns the number logged. */
int64_t orc_net_take_poisson_log(void *net, int source, int32_t *step, int32_t *idx, int64_t cap) {
    Net *n = (Net *)net;
    if (source < 0 || source >= (int)n->sources.size()) return -1;
    auto &lg = n->sources[(size_t)source].poisson_log;
    const int64_t m = (int64_t)lg.size();
    if (cap < m) return m;
    for (int64_t i = 0; i < m; i++) {
        step[i] = lg[(size_t)i].first;
        idx[i] = lg[(size_t)i].second;
    }
    lg.clear();
    return m;
}

/* Moves the logged spikes out when cap holds them all (else leaves them);
 * returns the number logged. */
int64_t orc_net_take_log(void *net, int32_t *step, int32_t *idx, int64_t cap) {
    Net *n = (Net *)net;
    const int64_t m = (int64_t)n->log.size();
    if (cap < m) return m;
    for (int64_t i = 0; i < m; i++) {
        step[i] = n->log[(size_t)i].first;
        idx[i] = n->log[(size_t)i].second;
    }
    n->log.clear();
    return m;
}

}  // extern "C"
