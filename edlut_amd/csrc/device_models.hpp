// device_models.hpp -- per-neuron right-hand sides of the time-driven models
// on gfx950.  State is held in registers; the update kernel loads it from the
// structure-of-arrays layout and stores it back once per step.
//
// Float expressions keep the association of the reference's CPU models
// (every file compiled with -ffp-contract=off, IEEE division, no denormal
// flush) so that results are bit-identical to them.  Look-up tables are built
// on the host with the same routines as the reference and uploaded once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edl {

// ExponentialTable geometry (Kernel/include/simulation/ExponentialTable.h:60-103,
// Kernel/src/simulation/ExponentialTable.cpp:23-27).
constexpr int kExpSize = 1024 * 16;
constexpr float kExpMin = -20.0f;
constexpr float kExpMax = 20.0f;
constexpr float kExpAux = (kExpSize - 1) / (kExpMax - kExpMin);
constexpr int kNmdaSize = 256;   // TimeDrivenNeuronModel.h:103

// ExponentialTable::GetResult: truncating index into the table on [Min, Max],
// exp above, zero below.  The table lives in global memory and is served from
// L2/L1 (neighbouring lanes read neighbouring entries).
__device__ __forceinline__ float lut_exp(const float *__restrict__ t, float x) {
    if (x >= kExpMin && x <= kExpMax) return t[(int)((x - kExpMin) * kExpAux)];
    if (x > kExpMax) return expf(x);
    return 0.0f;
}

// float -> int32 truncation as the reference's x86-64 build performs it
// (cvttss2si): NaN and values outside the int32 range give INT_MIN, where the
// GPU's own conversion would saturate.
__device__ __forceinline__ int trunc_i32_x86(float x) {
    if (!(x < 2147483648.0f) || x < -2147483648.0f) return (int)0x80000000u;
    return (int)x;
}

// Get_g_nmda_inf (AdExTimeDrivenModel.cpp:38-47): clamped truncating index.
__device__ __forceinline__ float nmda_inf(const float *__restrict__ t, float aux, float e_inh,
                                          float v) {
    int p = trunc_i32_x86((v - e_inh) * aux);
    p = p < 0 ? 0 : (p > kNmdaSize - 1 ? kNmdaSize - 1 : p);
    return t[p];
}

// Conductance decay: values below 1e-9 snap to zero, others scale by the
// precomputed factor (AdExTimeDrivenModel.cpp:220-245).
__device__ __forceinline__ float decay_g(float g, float f) {
    return (g < 1e-9f) ? 0.0f : g * f;
}
__device__ __forceinline__ double decay_g(double g, double f) {
    return (g < (double)1e-9f) ? 0.0 : g * f;
}

// Enabled synapse types.  AdEx, LIF and HH kernels take them from the
// compile-time MASK (EDLUT_SYN_* bits), or, when MASK has kRtMask set, from
// the run-time mask of the parameters (one kernel for every combination).
constexpr unsigned kRtMask = 16u;
template <unsigned MASK>
__device__ __forceinline__ bool syn(uint32_t rt, unsigned bit) {
    return (MASK & kRtMask) ? (rt & bit) != 0u : (MASK & bit) != 0u;
}

// ---------------------------------------------------------------------------
// Model descriptors.  Each exposes NV (state variables), ND (differential
// variables), the variable rows that hold g_exc/g_inh/g_nmda/ext_I, the
// right-hand side, the decay and the threshold/reset.
// ---------------------------------------------------------------------------
struct AdExParams {
    float a, b, inv_thr, v_thr, v_cap, inv_tau_w, e_exc, e_inh, e_reset, e_leak, g_leak,
        glk_thr, inv_c_m, nmda_aux;
    float cf[3];     // decay factors for the integrator's elapsed time
    float init[6];   // initial state (reset target of the NaN check)
    uint32_t mask;   // run-time synapse mask (kernels compiled with kRtMask)
};

struct LIFParams {
    float e_exc, e_inh, e_leak, v_thr, g_leak, inv_c_m, nmda_aux;
    float cf[3];
    float init[5];
    uint32_t mask;
};

struct HHParams {
    float e_exc, e_inh, e_leak, g_leak, inv_c_m, g_na, g_kd, e_na, e_k, nmda_aux;
    float cf[3];
    float init[8];
    double inv_c_m_d;   // FP64 Newton mode: 1/c_m and decay factors in double
    double cf_d[3];
    uint32_t mask;
};

// Granule and Purkinje cells keep the enabled synapse types as a run-time
// mask (`mask`, EDLUT_SYN_* bits): one kernel serves every combination.
struct GranuleParams {
    float e_exc, e_inh, nmda_aux;
    float a_xkca, b_xkca;   // Q10_30*2.5f, Q10_30*1.5f
    float k_nernst;         // 1000*(R*(temper+273.15)/F)/z for calcium (z = 2)
    float cf[3];
    float init[19];
    uint32_t mask;
};

struct PurkinjeParams {
    float e_exc, e_inh, e_leak, v_thr, nmda_aux;
    float rise, fall;       // refractory dV/dt: (spk_peak-v_thr)/tau_ref_0_5, (e_leak-spk_peak)/tau_ref_0_5
    double cur_scale;       // 1e-6 * inv_area (pA -> model units, in double)
    float cf[3];
    float init[7];
    uint32_t mask;
};

struct IzhParams {
    float a, b, c, d, e_exc, e_inh, inv_c_m, nmda_aux;
    float cf[3];
    float init[6];
    uint32_t mask;   // run-time synapse mask, as for the granule
};

struct LifIsParams {
    float e_exc, e_inh, e_leak, v_thr, g_leak_nS, inv_c_m_nF, nmda_aux;
    float cf[3];
    float init[5];
    uint32_t mask;   // run-time synapse mask
};

struct ModelParams {
    AdExParams adex;
    LIFParams lif;
    HHParams hh;
    GranuleParams gran;
    PurkinjeParams pkj;
    IzhParams izh;
    LifIsParams lifis;
};

// Per-thread context the right-hand sides need besides the state.
struct RhsCtx {
    const float *chan;   // channel-rate table (HH 6, granule 26, Purkinje 4 floats per entry)
    const float *exp_lut;
    const float *nmda_lut;
    const float *refr;   // LIF: refractory fractions by steps since the last spike;
                         // Purkinje: (derivative mode, V override) pairs (see PurkinjeModel)
    int refr_len;
    int since_spike;     // steps since the last spike (see the clock tables)
};

struct AdExModel {
    static constexpr int NV = 6, ND = 2;
    static constexpr bool kResetsIntegrator = true;   // spike -> resetState
    static constexpr bool kClockOverride = false;     // no post-spike V waveform
    static constexpr int EXC = 2, INH = 3, NMDA = 4, EXT = 5;

    // AdExTimeDrivenModel.cpp:170-218.
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s,
                                               T *d) {
        const AdExParams &p = mp.adex;
        const float V = s[0];
        if (V <= p.v_cap) {
            float cur = 0.0f;
            if (syn<MASK>(p.mask, 1u)) cur += s[EXC] * (p.e_exc - V);
            if (syn<MASK>(p.mask, 2u)) cur += s[INH] * (p.e_inh - V);
            if (syn<MASK>(p.mask, 4u)) {
                const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
                cur += s[NMDA] * gi * (p.e_exc - V);
            }
            if (syn<MASK>(p.mask, 8u)) cur += s[EXT];
            d[0] = (p.g_leak * (p.e_leak - V) + p.glk_thr * lut_exp(c.exp_lut, (V - p.v_thr) * p.inv_thr) -
                    s[1] + cur) * p.inv_c_m;
            d[1] = (p.a * (V - p.e_leak) - s[1]) * p.inv_tau_w;
        } else if (V <= 0.0f) {
            const float Vm = p.v_cap;
            float cur = 0.0f;
            if (syn<MASK>(p.mask, 1u)) cur += s[EXC] * (p.e_exc - Vm);
            if (syn<MASK>(p.mask, 2u)) cur += s[INH] * (p.e_inh - Vm);
            if (syn<MASK>(p.mask, 4u)) {
                const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, Vm);
                cur += s[NMDA] * gi * (p.e_exc - Vm);
            }
            if (syn<MASK>(p.mask, 8u)) cur += s[EXT];
            d[0] = (p.g_leak * (p.e_leak - Vm) + p.glk_thr * lut_exp(c.exp_lut, (Vm - p.v_thr) * p.inv_thr) -
                    s[1] + cur) * p.inv_c_m;
            d[1] = (p.a * (V - p.e_leak) - s[1]) * p.inv_tau_w;
        } else {
            d[0] = 0.0f;
            d[1] = 0.0f;
        }
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const AdExParams &p = mp.adex;
        if (syn<MASK>(p.mask, 1u)) s[EXC] = decay_g(s[EXC], mp.adex.cf[0]);
        if (syn<MASK>(p.mask, 2u)) s[INH] = decay_g(s[INH], mp.adex.cf[1]);
        if (syn<MASK>(p.mask, 4u)) s[NMDA] = decay_g(s[NMDA], mp.adex.cf[2]);
    }
    // AdExTimeDrivenModel.cpp:159-168.
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &mp, T prev_v, T *s) {
        if (s[0] > 0.0f) {
            s[0] = mp.adex.e_reset;
            s[1] += mp.adex.b;
            return true;
        }
        return false;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.adex.init[v];
    }
};

struct LIFModel {
    static constexpr int NV = 5, ND = 1;
    static constexpr bool kResetsIntegrator = true;
    static constexpr bool kClockOverride = false;
    static constexpr int EXC = 1, INH = 2, NMDA = 3, EXT = 4;

    // LIFTimeDrivenModel.cpp:170-202.  The double-precision refractory clock
    // of the reference is replaced by an integer count of steps since the
    // last spike and a host-built table of the fraction the clock yields for
    // each count (negative = derivative forced to zero); counts past the
    // table's end are outside the refractory period.
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s,
                                               T *d) {
        const LIFParams &p = mp.lif;
        const float V = s[0];
        float cur = 0.0f;
        if (syn<MASK>(p.mask, 1u)) cur += s[EXC] * (p.e_exc - V);
        if (syn<MASK>(p.mask, 2u)) cur += s[INH] * (p.e_inh - V);
        if (syn<MASK>(p.mask, 4u)) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        if (syn<MASK>(p.mask, 8u)) cur += s[EXT];
        if (c.since_spike >= c.refr_len) {
            d[0] = (cur + p.g_leak * (p.e_leak - V)) * p.inv_c_m;
        } else {
            const float f = c.refr[c.since_spike];
            d[0] = (f < 0.0f) ? 0.0f : f * ((cur + p.g_leak * (p.e_leak - V)) * p.inv_c_m);
        }
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const LIFParams &p = mp.lif;
        if (syn<MASK>(p.mask, 1u)) s[EXC] = decay_g(s[EXC], mp.lif.cf[0]);
        if (syn<MASK>(p.mask, 2u)) s[INH] = decay_g(s[INH], mp.lif.cf[1]);
        if (syn<MASK>(p.mask, 4u)) s[NMDA] = decay_g(s[NMDA], mp.lif.cf[2]);
    }
    // LIFTimeDrivenModel.cpp:158-168.
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &mp, T prev_v, T *s) {
        if (s[0] > mp.lif.v_thr) {
            s[0] = mp.lif.e_leak;
            return true;
        }
        return false;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.lif.init[v];
    }
};


// HH channel-rate table geometry (HHTimeDrivenModel.h:64-79, .cpp:27-30).
constexpr int kHHTable = 1024 * 256;
constexpr float kHHMinV = -90.0f;
constexpr float kHHMaxV = 50.0f;
constexpr float kHHAux = (kHHTable - 1) / (kHHMaxV - kHHMinV);

struct HHModel {
    static constexpr int NV = 8, ND = 4;
    static constexpr bool kResetsIntegrator = false;
    static constexpr bool kClockOverride = false;
    static constexpr int EXC = 4, INH = 5, NMDA = 6, EXT = 7;

    // HHTimeDrivenModel.cpp:197-264.  The state it is given is clamped in
    // place (V to [e_k, e_na-1], gates to [0,1], n to 0.72), as in the
    // reference, where the first stage receives the neuron's own state.
    // T = double is the FP64 Newton mode: same expression order in double,
    // channel rates from the same float table.
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s,
                                               T *d) {
        const HHParams &p = mp.hh;
        if (s[0] < (T)p.e_k) s[0] = p.e_k;
        if (s[0] > (T)(p.e_na - 1.0f)) s[0] = p.e_na - 1.0f;
        if (s[1] < 0) s[1] = 0.0f;
        if (s[1] > 1) s[1] = 1.0f;
        if (s[2] < 0) s[2] = 0.0f;
        if (s[2] > 1) s[2] = 1.0f;
        if (s[3] < 0) s[3] = 0.0f;
        if ((double)s[3] > 0.72) s[3] = 0.72f;
        const T V = s[0], m = s[1], h = s[2], n = s[3];
        int pos = trunc_i32_x86(((float)V - kHHMinV) * kHHAux);
        pos = pos < 0 ? 0 : (pos > kHHTable - 1 ? kHHTable - 1 : pos);
        const float *v = c.chan + (int64_t)pos * 6;
        T cur = 0.0f;
        if (syn<MASK>(p.mask, 1u)) cur += s[EXC] * ((T)p.e_exc - s[0]);
        if (syn<MASK>(p.mask, 2u)) cur += s[INH] * ((T)p.e_inh - s[0]);
        if (syn<MASK>(p.mask, 4u)) {
            const T gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, (float)s[0]);
            cur += s[NMDA] * gi * ((T)p.e_exc - s[0]);
        }
        if (syn<MASK>(p.mask, 8u)) cur += s[EXT];
        const T inv_c_m = sizeof(T) == sizeof(double) ? (T)p.inv_c_m_d : (T)p.inv_c_m;
        d[0] = (cur + (T)p.g_leak * ((T)p.e_leak - V) + (T)p.g_na * m * m * m * h * ((T)p.e_na - V) +
                (T)p.g_kd * n * n * n * n * ((T)p.e_k - V)) * inv_c_m;
        const T am = v[0], bm = v[1], ah = v[2], bh = v[3], an = v[4], bn = v[5];
        d[1] = (am * (1.0f - m) - bm * m);
        d[2] = (ah * (1.0f - h) - bh * h);
        d[3] = (an * (1.0f - n) - bn * n);
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const bool dbl = sizeof(T) == sizeof(double);
        const HHParams &p = mp.hh;
        if (syn<MASK>(p.mask, 1u)) s[EXC] = decay_g(s[EXC], dbl ? (T)mp.hh.cf_d[0] : (T)mp.hh.cf[0]);
        if (syn<MASK>(p.mask, 2u)) s[INH] = decay_g(s[INH], dbl ? (T)mp.hh.cf_d[1] : (T)mp.hh.cf[1]);
        if (syn<MASK>(p.mask, 4u)) s[NMDA] = decay_g(s[NMDA], dbl ? (T)mp.hh.cf_d[2] : (T)mp.hh.cf[2]);
    }
    // HHTimeDrivenModel.cpp:187-193: upward crossing of -30 mV, no reset.
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &, T prev_v, T *s) {
        return s[0] > -30.0f && prev_v < -30.0f;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.hh.init[v];
    }
};

// LIF in SI units (LIFTimeDrivenModel_IS.cpp, second time scale): V | g_exc,
// g_inh, g_nmda, ext_I.  Conductances in nS, V in volts, the external current
// in pA (taken as nA x 1000), the refractory clock in seconds (table over
// steps since the spike, as for LIFModel).
struct LifIsModel {
    static constexpr int NV = 5, ND = 1;
    static constexpr bool kResetsIntegrator = true;
    static constexpr bool kClockOverride = false;
    static constexpr int EXC = 1, INH = 2, NMDA = 3, EXT = 4;

    // LIFTimeDrivenModel_IS.cpp:166-193
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s, T *d) {
        const LifIsParams &p = mp.lifis;
        const float V = s[0];
        float cur = 0.0f;
        if (p.mask & 1u) cur += s[EXC] * (p.e_exc - V);
        if (p.mask & 2u) cur += s[INH] * (p.e_inh - V);
        if (p.mask & 4u) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        const float ext_na = (float)(0.001 * (double)s[EXT]);
        cur += ext_na;
        if (c.since_spike >= c.refr_len) {
            d[0] = (cur + p.g_leak_nS * (p.e_leak - V)) * p.inv_c_m_nF;
        } else {
            const float f = c.refr[c.since_spike];
            d[0] = (f < 0.0f) ? 0.0f : f * ((cur + p.g_leak_nS * (p.e_leak - V)) * p.inv_c_m_nF);
        }
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const LifIsParams &p = mp.lifis;
        if (p.mask & 1u) s[EXC] = decay_g(s[EXC], p.cf[0]);
        if (p.mask & 2u) s[INH] = decay_g(s[INH], p.cf[1]);
        if (p.mask & 4u) s[NMDA] = decay_g(s[NMDA], p.cf[2]);
    }
    // :155-163
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &mp, T, T *s) {
        if (s[0] > mp.lifis.v_thr) {
            s[0] = mp.lifis.e_leak;
            return true;
        }
        return false;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.lifis.init[v];
    }
};

// Izhikevich (IzhikevichTimeDrivenModel.cpp): V, u | g_exc, g_inh, g_nmda, ext_I.
struct IzhModel {
    static constexpr int NV = 6, ND = 2;
    static constexpr bool kResetsIntegrator = true;
    static constexpr bool kClockOverride = false;
    static constexpr int EXC = 2, INH = 3, NMDA = 4, EXT = 5;

    // IzhikevichTimeDrivenModel.cpp:169-195 (run-time synapse mask; MASK unused)
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s, T *d) {
        const IzhParams &p = mp.izh;
        const float V = s[0];
        float cur = 0.0f;
        if (p.mask & 1u) cur += s[EXC] * (p.e_exc - V);
        if (p.mask & 2u) cur += s[INH] * (p.e_inh - V);
        if (p.mask & 4u) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        cur += s[EXT];
        if (V <= 30.0f) {
            d[0] = 0.04f * V * V + 5 * V + 140 - s[1] + (cur * p.inv_c_m);
            d[1] = p.a * (p.b * V - s[1]);
        } else {
            d[0] = 0.0f;
            d[1] = 0.0f;
        }
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const IzhParams &p = mp.izh;
        if (p.mask & 1u) s[EXC] = decay_g(s[EXC], p.cf[0]);
        if (p.mask & 2u) s[INH] = decay_g(s[INH], p.cf[1]);
        if (p.mask & 4u) s[NMDA] = decay_g(s[NMDA], p.cf[2]);
    }
    // :156-165: above 30 mV, V resets to c and u grows by d
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &mp, T, T *s) {
        if (s[0] > 30.0f) {
            s[0] = mp.izh.c;
            s[1] += mp.izh.d;
            return true;
        }
        return false;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.izh.init[v];
    }
};

// Natural logarithm of a positive float, evaluated in double and rounded to
// float: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
// s = (m-1)/(m+1) (|s| <= 0.1716), atanh by its Taylor series through s^23
// (truncation below 1e-19), e ln 2 with ln 2 split so that e*ln2_hi is exact.
// The result is the correctly rounded logf except where the exact value lies
// within ~1e-16 (relative) of a rounding midpoint.  Only +,-,*,/ in a fixed
// order, so a host build with the same operation order gives the same bits.
__device__ __forceinline__ float log_f(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    uint32_t b = __float_as_uint(x);
    int e = 0;
    if (b < 0x00800000u) {   // subnormal: scale by 2^23 (exact)
        x *= 8388608.0f;
        b = __float_as_uint(x);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    double m = (double)__uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double lm = (2.0 * s) * p;
    const double ln2_hi = 6.93147180369123816490e-01;   // 0x3fe62e42fee00000
    const double ln2_lo = 1.90821492927058770002e-10;
    const double de = (double)e;
    return (float)(de * ln2_hi + (de * ln2_lo + lm));
}

// Granule channel table (EgidioGranuleCell_TimeDriven.h:133-146,402-552):
// 1024 entries of 26 floats on [-100, 50] mV.
constexpr int kGrTable = 1024;
constexpr int kGrVars = 26;
constexpr float kGrMinV = -100.0f;
constexpr float kGrMaxV = 50.0f;
constexpr float kGrAux = (kGrTable - 1) / (kGrMaxV - kGrMinV);

struct GranuleModel {
    // V, xNa_f, yNa_f, xNa_r, yNa_r, xNa_p, xK_V, xK_A, yK_A, xK_IR, xK_Ca,
    // xCa, yCa, xK_sl, Ca | g_exc, g_inh, g_nmda, ext_I
    static constexpr int NV = 19, ND = 15;
    static constexpr bool kResetsIntegrator = false;
    static constexpr bool kClockOverride = false;
    static constexpr int EXC = 15, INH = 16, NMDA = 17, EXT = 18;

    // EgidioGranuleCell_TimeDriven.cpp:252-365 (constants :26-58), split into
    // the parts the BDF Newton step reuses: the voltage-indexed table row and
    // the calcium Nernst potential (Terms), the membrane equation (dv) and
    // the gate / calcium equations.  The synapse types come from the run-time
    // mask; MASK is unused.
    struct Terms {
        float v[kGrVars];   // channel table row at V
        float vca;          // nernst(Ca, cao = 2, z = 2, 30 C)
    };
    __device__ static __forceinline__ void load_row(const RhsCtx &c, float V, float *v) {
        int pos = trunc_i32_x86((V - kGrMinV) * kGrAux);
        pos = pos < 0 ? 0 : (pos > kGrTable - 1 ? kGrTable - 1 : pos);
        // a row is 104 bytes: 13 aligned float2 loads
        const float2 *r2 = (const float2 *)(c.chan + pos * kGrVars);
#pragma unroll
        for (int q = 0; q < kGrVars / 2; q++) {
            const float2 t = r2[q];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
    __device__ static __forceinline__ float nernst_ca(const GranuleParams &p, float ca) {
        return p.k_nernst * log_f(fabsf(2.0f / ca));
    }
    __device__ static __forceinline__ void prep(const ModelParams &mp, const RhsCtx &c, const float *s,
                                                Terms &t) {
        load_row(c, s[0], t.v);
        t.vca = nernst_ca(mp.gran, s[14]);
    }
    // dV/dt (.cpp:342-348 with the synaptic current of :255-272)
    __device__ static __forceinline__ float dv(const ModelParams &mp, const RhsCtx &c, const float *s,
                                               float vca) {
        const GranuleParams &p = mp.gran;
        const float V = s[0];
        float cur = 0.0f;
        if (p.mask & 1u) cur += s[EXC] * (p.e_exc - V);
        if (p.mask & 2u) cur += s[INH] * (p.e_inh - V);
        if (p.mask & 4u) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        cur += s[EXT];
        cur = (float)((double)cur * 3.34156941e-4);   // float *= double literal
        const float pv = V;
        const float gNa_f = 0.013f * s[1] * s[1] * s[1] * s[2];
        const float gNa_r = 0.0005f * s[3] * s[4];
        const float gNa_p = 0.0002f * s[5];
        const float gK_V = 0.003f * s[6] * s[6] * s[6] * s[6];
        const float gK_A = 0.004f * s[7] * s[7] * s[7] * s[8];
        const float gK_IR = 0.0009f * s[9];
        const float gK_Ca = 0.004f * s[10];
        const float gCa = 0.00046f * s[11] * s[11] * s[12];
        const float gK_sl = 0.00035f * s[13];
        const float VNa = 87.39f, VK = -84.69f;
        return (cur + gNa_f * (VNa - pv) + gNa_r * (VNa - pv) + gNa_p * (VNa - pv) + gK_V * (VK - pv) +
                gK_A * (VK - pv) + gK_IR * (VK - pv) + gK_Ca * (VK - pv) + gCa * (vca - pv) +
                gK_sl * (VK - pv) + 5.68e-5f * (-58.0f - pv) + 2.17e-5f * (-65.0f - pv)) *
               (1.0f / 1.0e-3f);
    }
    // voltage-gated variable k (1..9, 11..13): (x_inf - x)/tau from its table columns
    __device__ static __forceinline__ float dgate(int k, const float *s, const float *v) {
        const int col = k < 10 ? 2 * (k - 1) : 2 * k - 2;
        return (v[col] - s[k]) * v[col + 1];
    }
    // Ca-dependent K gate (.cpp:300-305)
    __device__ static __forceinline__ float dxkca(const GranuleParams &p, const float *s, const float *v) {
        const float ca = s[14];
        const float a_kca = p.a_xkca / (1.0f + v[18] / ca);
        const float b_kca = p.b_xkca / (1.0f + ca * v[19]);
        const float xkca_inf = a_kca / (a_kca + b_kca);
        const float inv_tau_xkca = a_kca + b_kca;
        return (xkca_inf - s[10]) * inv_tau_xkca;
    }
    // calcium (.cpp:340-341): 2*F*A*d = ((2*96485.309f)*1e-04f)*0.2f, betaCa 1.5, Ca0 1e-4
    __device__ static __forceinline__ float dca(const float *s, float vca) {
        constexpr float kFAd = ((2 * 96485.309f) * 1e-04f) * 0.2f;
        const float gCa = 0.00046f * s[11] * s[11] * s[12];
        return (-gCa * (s[0] - vca) / kFAd - (1.5f * (s[14] - 1e-04f)));
    }
    __device__ static __forceinline__ void rhs_terms(const ModelParams &mp, const RhsCtx &c, const float *s,
                                                     const Terms &t, float *d) {
#pragma unroll
        for (int k = 1; k < 14; k++) d[k] = k == 10 ? dxkca(mp.gran, s, t.v) : dgate(k, s, t.v);
        d[14] = dca(s, t.vca);
        d[0] = dv(mp, c, s, t.vca);
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s, T *d) {
        Terms t;
        prep(mp, c, s, t);
        rhs_terms(mp, c, s, t, d);
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const GranuleParams &p = mp.gran;
        if (p.mask & 1u) s[EXC] = decay_g(s[EXC], p.cf[0]);
        if (p.mask & 2u) s[INH] = decay_g(s[INH], p.cf[1]);
        if (p.mask & 4u) s[NMDA] = decay_g(s[NMDA], p.cf[2]);
    }
    // EgidioGranuleCell_TimeDriven.cpp:243-249: upward crossing of v_thr = 0 mV.
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &, T prev_v, T *s) {
        return s[0] > 0.0f && prev_v < 0.0f;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.gran.init[v];
    }
};

// Purkinje channel table (TimeDrivenPurkinjeCell.h:76-91,396-441): 1024
// entries of (alpha_ca, inv_tau_ca, alpha_M, inv_tau_M) on [-100, 35] mV.
constexpr int kPkTable = 1024;
constexpr float kPkMinV = -100.0f;
constexpr float kPkMaxV = 35.0f;
constexpr float kPkAux = (kPkTable - 1) / (kPkMaxV - kPkMinV);

struct PurkinjeModel {
    // V, ca, M | g_exc, g_inh, g_nmda, ext_I
    static constexpr int NV = 7, ND = 3;
    static constexpr bool kResetsIntegrator = false;
    static constexpr bool kClockOverride = true;
    static constexpr int EXC = 3, INH = 4, NMDA = 5, EXT = 6;

    // The reference's double LastSpikeTime is replaced by the count of steps
    // since the last spike and a host-built table over that count: entry j
    // (clock S_j = j-fold double sum of the step, 0 right after a spike)
    // holds the derivative mode of S_j (0 full, 1 rising, 2 falling flank;
    // TimeDrivenPurkinjeCell.cpp:218-224) and the membrane potential the
    // spike condition imposes at S_j (NaN: none; :181-189).  Counts past the
    // table's end are outside the refractory period.
    __device__ static __forceinline__ float clock_entry(const RhsCtx &c, int j, int q, float dflt) {
        return (j >= 0 && j < c.refr_len) ? c.refr[2 * j + q] : dflt;
    }

    // TimeDrivenPurkinjeCell.cpp:193-239.
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c, T *s,
                                               T *d) {
        const PurkinjeParams &p = mp.pkj;
        const float V = s[0], ca = s[1], M = s[2];
        float cur = 0.0f;
        if (p.mask & 1u) cur += s[EXC] * (p.e_exc - V);
        if (p.mask & 2u) cur += s[INH] * (p.e_inh - V);
        if (p.mask & 4u) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        cur += s[EXT];
        cur = (float)((double)cur * p.cur_scale);
        const float mode = clock_entry(c, c.since_spike, 0, 0.0f);
        if (mode == 0.0f) {
            d[0] = (-0.02f * (V + 70.0f) - 0.001f * ca * ca * (V - 125.0f) - 0.75f * M * (V + 95.0f) + cur) *
                   (1.0f / 0.95f);
        } else {
            d[0] = mode == 1.0f ? p.rise : p.fall;
        }
        int pos = trunc_i32_x86((V - kPkMinV) * kPkAux);
        pos = pos < 0 ? 0 : (pos > kPkTable - 1 ? kPkTable - 1 : pos);
        const float4 t = *(const float4 *)(c.chan + pos * 4);
        d[1] = t.x - ca * t.y;
        d[2] = t.z - M * t.w;
    }
    template <unsigned MASK, typename T = float>
    __device__ static __forceinline__ void decay(const ModelParams &mp, T *s) {
        const PurkinjeParams &p = mp.pkj;
        if (p.mask & 1u) s[EXC] = decay_g(s[EXC], p.cf[0]);
        if (p.mask & 2u) s[INH] = decay_g(s[INH], p.cf[1]);
        if (p.mask & 4u) s[NMDA] = decay_g(s[NMDA], p.cf[2]);
    }
    // TimeDrivenPurkinjeCell.cpp:172-176: V reaches v_thr from below.
    template <typename T = float>
    __device__ static __forceinline__ bool spike(const ModelParams &mp, T prev_v, T *s) {
        return s[0] >= mp.pkj.v_thr && prev_v < mp.pkj.v_thr;
    }
    // :177-189: after the clock advanced (or restarted at a spike), the
    // waveform value of the new clock replaces V.
    template <typename T = float>
    __device__ static __forceinline__ void clock_override(const RhsCtx &c, bool fired, T *s) {
        const float v = clock_entry(c, fired ? 0 : c.since_spike + 1, 1, __builtin_nanf(""));
        if (v == v) s[0] = v;
    }
    __device__ static __forceinline__ float init(const ModelParams &mp, int v) {
        return mp.pkj.init[v];
    }
};

// ---------------------------------------------------------------------------
// Fixed-step integrators on I neurons held in registers.  Work is issued
// stage-major across the I neurons so that their look-up-table gathers are in
// flight together; each neuron's arithmetic is exactly the one-neuron
// sequence of the reference.
// dt  = elapsedTimeInNeuronModelScale; h = dt*0.5f; s6 = dt*0.166666666667f.
// ---------------------------------------------------------------------------
struct StepConsts {
    float dt, h, s6;
};

// Euler.h:123-146.
template <class M, unsigned MASK, int I>
__device__ __forceinline__ void integrate_euler(const ModelParams &mp, const RhsCtx *c,
                                                const StepConsts &k, float (*s)[M::NV]) {
    float d[I][M::ND];
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], s[n], d[n]);
#pragma unroll
    for (int n = 0; n < I; n++) {
#pragma unroll
        for (int j = 0; j < M::ND; j++) s[n][j] += k.dt * d[n][j];
        M::template decay<MASK>(mp, s[n]);
    }
}

// RK2.h:133-165 (Heun with one decay between the stages).
template <class M, unsigned MASK, int I>
__device__ __forceinline__ void integrate_rk2(const ModelParams &mp, const RhsCtx *c,
                                              const StepConsts &k, float (*s)[M::NV]) {
    float d1[I][M::ND], d2[I][M::ND];
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], s[n], d1[n]);
#pragma unroll
    for (int n = 0; n < I; n++) {
#pragma unroll
        for (int j = 0; j < M::ND; j++) s[n][j] += d1[n][j] * k.dt;
        M::template decay<MASK>(mp, s[n]);
    }
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], s[n], d2[n]);
#pragma unroll
    for (int n = 0; n < I; n++)
#pragma unroll
        for (int j = 0; j < M::ND; j++) s[n][j] += (-d1[n][j] + d2[n][j]) * k.dt * 0.5f;
}

// RK4.h:167-231: half-step decays before stages 2 and 4.
template <class M, unsigned MASK, int I>
__device__ __forceinline__ void integrate_rk4(const ModelParams &mp, const RhsCtx *c,
                                              const StepConsts &k, float (*s)[M::NV]) {
    float a[I][M::NV], d1[I][M::ND], d2[I][M::ND], d3[I][M::ND], d4[I][M::ND];
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], s[n], d1[n]);
#pragma unroll
    for (int n = 0; n < I; n++) {
#pragma unroll
        for (int j = 0; j < M::ND; j++) a[n][j] = s[n][j] + d1[n][j] * k.h;
#pragma unroll
        for (int j = M::ND; j < M::NV; j++) a[n][j] = s[n][j];
        M::template decay<MASK>(mp, a[n]);
    }
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], a[n], d2[n]);
#pragma unroll
    for (int n = 0; n < I; n++)
#pragma unroll
        for (int j = 0; j < M::ND; j++) a[n][j] = s[n][j] + d2[n][j] * k.h;
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], a[n], d3[n]);
#pragma unroll
    for (int n = 0; n < I; n++) {
#pragma unroll
        for (int j = 0; j < M::ND; j++) a[n][j] = s[n][j] + d3[n][j] * k.dt;
        M::template decay<MASK>(mp, a[n]);
    }
#pragma unroll
    for (int n = 0; n < I; n++) M::template rhs<MASK>(mp, c[n], a[n], d4[n]);
#pragma unroll
    for (int n = 0; n < I; n++) {
#pragma unroll
        for (int j = 0; j < M::ND; j++)
            s[n][j] += (d1[n][j] + 2.0f * (d2[n][j] + d3[n][j]) + d4[n][j]) * k.s6;
#pragma unroll
        for (int j = M::ND; j < M::NV; j++) s[n][j] = a[n][j];
    }
}


// ---------------------------------------------------------------------------
// BDF (BDFn.h:288-423, Jacobian 682-722, invermat 724-786) on one neuron
// held in registers, in arithmetic type T (float: the reference's; double:
// the FP64 Newton mode).  D and Prev are the difference table and previous
// states (row-major in the history index, SoA over neurons).
// ---------------------------------------------------------------------------
template <int N, typename T>
__device__ __forceinline__ void invermat_dev(const T *a, T *ainv) {
    if (N == 1) {
        ainv[0] = (T)1.0f / a[0];
        return;
    }
    T la[N * N], li[N * N];
#pragma unroll
    for (int q = 0; q < N * N; q++) {
        la[q] = a[q];
        li[q] = 0;
    }
#pragma unroll
    for (int q = 0; q < N; q++) li[q * N + q] = 1.0f;
#pragma unroll
    for (int s = 0; s < N; s++) {
        T element = la[s * N + s];
        if (element == 0) {
            bool done = false;
#pragma unroll
            for (int r = s + 1; r < N; r++) {
                if (!done && la[r * N + s] != 0) {
                    element = la[r * N + s];
#pragma unroll
                    for (int m = 0; m < N; m++) {
                        T v = la[r * N + m];
                        la[r * N + m] = la[s * N + m];
                        la[s * N + m] = v;
                        v = li[r * N + m];
                        li[r * N + m] = li[s * N + m];
                        li[s * N + m] = v;
                    }
                    done = true;
                }
            }
        }
        const T inv_element = (T)1.0f / element;
#pragma unroll
        for (int j = 0; j < N; j++) {
            la[s * N + j] *= inv_element;
            li[s * N + j] *= inv_element;
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i != s) {
                const T coef = -la[i * N + s];
#pragma unroll
                for (int j = 0; j < N; j++) {
                    la[i * N + j] += la[s * N + j] * coef;
                    li[i * N + j] += li[s * N + j] * coef;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < N * N; q++) ainv[q] = li[q];
}

struct BdfState {
    int order;
    const float *C;    // [7][7] coefficients (BDFn.h:604-669)
    void *d;           // T[order][ND][stride]
    void *prev;        // T[order-1][ND][stride]
    int32_t *st;       // [stride]
    // granule: working storage of the reference-arranged Newton correction,
    // ws_cap slots per step handed out by ws_count[step parity]
    float *ws;
    uint32_t *ws_count;
    int ws_cap;
    uint32_t *err;
    // granule: spike bits of the launch, one word per 32 neurons, and the
    // count of finished workgroups (both zero between launches)
    uint32_t *spk_bits;
    uint32_t *done;
};

template <class M, unsigned MASK, typename T>
__device__ __forceinline__ void integrate_bdf(const ModelParams &mp, const RhsCtx &c, T dt,
                                              const BdfState &b, int64_t ld, uint32_t i,
                                              int &st, T *S) {
    constexpr int ND = M::ND, NV = M::NV;
    T *Dh = (T *)b.d;
    T *Ph = (T *)b.prev;
    T F[ND], P[NV], P1[ND], Cc[ND], jac[ND * ND], J[ND * ND], iJ[ND * ND], X[NV], Fp[ND], Fn[ND];
    if (st == 0) {
        M::template rhs<MASK, T>(mp, c, S, F);
#pragma unroll
        for (int j = 0; j < ND; j++) P[j] = S[j] + dt * F[j];
    } else {
#pragma unroll
        for (int j = 0; j < ND; j++) {
            P[j] = S[j];
            for (int q = 0; q < st; q++) P[j] += Dh[((int64_t)q * ND + j) * ld + i];
        }
    }
#pragma unroll
    for (int j = ND; j < NV; j++) P[j] = S[j];
    M::template decay<MASK, T>(mp, P);
    T epsi = 1.0f;
    int k = 0;
    const float *Cs = b.C + st * 7;
    while (epsi > 1e-6 && k < 10) {
        M::template rhs<MASK, T>(mp, c, P, F);
#pragma unroll
        for (int j = 0; j < ND; j++) {
            Cc[j] = (T)Cs[0] * dt * F[j] + (T)Cs[1] * S[j];
            for (int q = 1; q < st; q++) Cc[j] += (T)Cs[q + 1] * Ph[((int64_t)(q - 1) * ND + j) * ld + i];
        }
        const T eps = dt * 0.1f;
        const T inv_eps = (T)1.0f / eps;
#pragma unroll
        for (int q = 0; q < NV; q++) X[q] = P[q];
        M::template rhs<MASK, T>(mp, c, X, Fp);
#pragma unroll
        for (int j = 0; j < ND; j++) {
#pragma unroll
            for (int q = 0; q < NV; q++) X[q] = P[q];
            X[j] -= eps;
            M::template rhs<MASK, T>(mp, c, X, Fn);
#pragma unroll
            for (int z = 0; z < ND; z++) jac[z * ND + j] = (Fp[z] - Fn[z]) * inv_eps;
        }
#pragma unroll
        for (int z = 0; z < ND; z++)
#pragma unroll
            for (int t = 0; t < ND; t++) {
                J[z * ND + t] = (T)Cs[0] * dt * jac[z * ND + t];
                if (z == t) J[z * ND + t] -= 1;
            }
        invermat_dev<ND, T>(J, iJ);
#pragma unroll
        for (int z = 0; z < ND; z++) {
            T aux = 0.0f;
#pragma unroll
            for (int t = 0; t < ND; t++) aux += iJ[z * ND + t] * (P[t] - Cc[t]);
            P1[z] = aux + P[z];
        }
        T aux2 = 0.0f;
#pragma unroll
        for (int z = 0; z < ND; z++) {
            const T aux = fabs(P1[z] - P[z]);
            if (aux > aux2) aux2 = aux;
        }
#pragma unroll
        for (int z = 0; z < ND; z++) P[z] = P1[z];
        epsi = aux2;
        k++;
    }
    if (st < b.order) st++;
#pragma unroll
    for (int j = 0; j < ND; j++) {
        for (int q = st - 1; q > 0; q--) Dh[((int64_t)q * ND + j) * ld + i] = -Dh[((int64_t)(q - 1) * ND + j) * ld + i];
        Dh[(int64_t)j * ld + i] = P[j] - S[j];
        for (int q = 1; q < st; q++)
            Dh[((int64_t)q * ND + j) * ld + i] += Dh[((int64_t)(q - 1) * ND + j) * ld + i];
    }
    if (st > 1) {
        for (int q = st - 2; q > 0; q--)
#pragma unroll
            for (int j = 0; j < ND; j++)
                Ph[((int64_t)q * ND + j) * ld + i] = Ph[((int64_t)(q - 1) * ND + j) * ld + i];
#pragma unroll
        for (int j = 0; j < ND; j++) Ph[(int64_t)j * ld + i] = S[j];
    }
#pragma unroll
    for (int j = 0; j < ND; j++) S[j] = P[j];
    M::template decay<MASK, T>(mp, S);
}

// ---------------------------------------------------------------------------
// Granule BDF step (the same algorithm as integrate_bdf, BDFn.h:288-423 and
// 682-786), arranged for a 15-variable system held in registers:
//   * the Jacobian column of a gate perturbs one variable: only the rows
//     whose equations read it are evaluated (the others difference to +0
//     exactly, as in the reference), and the table row and the Nernst
//     potential are reused unless V or Ca is the perturbed variable;
//   * the inverse is formed in place: at pivot s the dead column s of the
//     eliminated matrix becomes column s of the inverse, and the inverse's
//     still-unit columns are not stored.  Every stored value takes exactly
//     the reference's operations (li[i][s] = 0 + li[s][s]*coef included);
//   * a zero pivot (the reference then swaps rows) or a non-finite value
//     sends that Newton iteration through the reference arrangement
//     (granule_newton_reference) on a global working-storage slot, so that
//     the kernel keeps no large private frame.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ bool invert_in_place(float *A) {
    bool ok = true;
#pragma unroll
    for (int s = 0; s < N; s++) {
        const float element = A[s * N + s];
        ok = ok && element != 0.0f;
        const float inv = 1.0f / element;
#pragma unroll
        for (int j = 0; j < N; j++)
            if (j != s) A[s * N + j] *= inv;
        A[s * N + s] = inv;
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i == s) continue;
            const float coef = -A[i * N + s];
#pragma unroll
            for (int j = 0; j < N; j++)
                if (j != s) A[i * N + j] += A[s * N + j] * coef;
            A[i * N + s] = 0.0f + A[s * N + s] * coef;
        }
    }
    bool fin = true;
#pragma unroll
    for (int q = 0; q < N * N; q++) fin = fin && __builtin_isfinite(A[q]);
    return ok && fin;
}

template <int N>
__device__ __forceinline__ bool any_nan(const float *x) {
    bool r = false;
#pragma unroll
    for (int q = 0; q < N; q++) r = r || x[q] != x[q];
    return r;
}

// Floats of one working-storage slot: la, li (15 x 15 each), X (19), Fp, Fn,
// Cc, P1 (15 each).
constexpr int kGrWsFloats = 2 * 225 + 19 + 4 * 15;

// One Newton correction exactly as the reference arranges it (BDFn.h:
// 372-400, 682-786): full Jacobian by one-sided differences, Gauss-Jordan
// inverse with row swaps, product.  Works on a global working-storage slot
// w, whose X holds the iterate and Cc the corrector constants on entry and
// whose P1 receives the corrected iterate.
__device__ __forceinline__ void granule_newton_reference(const ModelParams &mp, const RhsCtx &c, float dt,
                                                         float c0, float *w) {
    constexpr int ND = GranuleModel::ND, NV = GranuleModel::NV;
    float *la = w, *li = w + 225, *X = w + 450, *Fp = X + NV, *Fn = Fp + ND, *Cc = Fn + ND, *P1 = Cc + ND;
    const float eps = dt * 0.1f;
    const float inv_eps = 1.0f / eps;
    GranuleModel::rhs<15u>(mp, c, X, Fp);
    for (int j = 0; j < ND; j++) {
        const float xj = X[j];
        X[j] = xj - eps;
        GranuleModel::rhs<15u>(mp, c, X, Fn);
        X[j] = xj;
        for (int z = 0; z < ND; z++) la[z * ND + j] = c0 * dt * ((Fp[z] - Fn[z]) * inv_eps);
    }
    for (int z = 0; z < ND; z++) la[z * ND + z] -= 1;
    for (int q = 0; q < ND * ND; q++) li[q] = 0;
    for (int q = 0; q < ND; q++) li[q * ND + q] = 1.0f;
    for (int s = 0; s < ND; s++) {
        float element = la[s * ND + s];
        if (element == 0) {
            for (int r = s + 1; r < ND; r++) {
                element = la[r * ND + s];
                if (element != 0) {
                    for (int m = 0; m < ND; m++) {
                        float v = la[r * ND + m];
                        la[r * ND + m] = la[s * ND + m];
                        la[s * ND + m] = v;
                        v = li[r * ND + m];
                        li[r * ND + m] = li[s * ND + m];
                        li[s * ND + m] = v;
                    }
                    break;
                }
            }
        }
        const float inv_element = 1.0f / element;
        for (int j = 0; j < ND; j++) {
            la[s * ND + j] *= inv_element;
            li[s * ND + j] *= inv_element;
        }
        for (int i = 0; i < ND; i++) {
            if (i == s) continue;
            const float coef = -la[i * ND + s];
            for (int j = 0; j < ND; j++) {
                la[i * ND + j] += la[s * ND + j] * coef;
                li[i * ND + j] += li[s * ND + j] * coef;
            }
        }
    }
    for (int z = 0; z < ND; z++) {
        float aux = 0.0f;
        for (int t = 0; t < ND; t++) aux += li[z * ND + t] * (X[t] - Cc[t]);
        P1[z] = aux + X[z];
    }
}

// The granule BDF step in three parts, so that the update kernel can hand a
// lane its next neuron as soon as that lane's corrector has converged:
// granule_bdf_predict (the order-st predictor), granule_bdf_newton (one
// corrector iteration, repeated while epsi > 1e-6 and k < 10) and
// granule_bdf_commit (history shift and new state).  Run in that order they
// are the step of integrate_bdf above, BDFn.h:288-423.  The corrector reads
// the first previous-state row from ph0 (a lane's copy in LDS, stride kBlock,
// taken when the lane took the neuron), the older rows from memory.
template <unsigned MASK>
__device__ __forceinline__ void granule_bdf_predict(const ModelParams &mp, const RhsCtx &c, float dt,
                                                    const BdfState &b, int64_t ld, uint32_t i, int st,
                                                    const float *S, float *P) {
    using G = GranuleModel;
    constexpr int ND = G::ND, NV = G::NV;
    const float *Dh = (const float *)b.d;
    if (st == 0) {
        float F[ND];
        G::rhs<MASK>(mp, c, (float *)S, F);
#pragma unroll
        for (int j = 0; j < ND; j++) P[j] = S[j] + dt * F[j];
    } else {
#pragma unroll
        for (int j = 0; j < ND; j++) {
            P[j] = S[j];
            for (int q = 0; q < st; q++) P[j] += Dh[((int64_t)q * ND + j) * ld + i];
        }
    }
#pragma unroll
    for (int j = ND; j < NV; j++) P[j] = S[j];
    G::decay<MASK>(mp, P);
}

template <unsigned MASK>
__device__ __forceinline__ void granule_bdf_newton(const ModelParams &mp, const RhsCtx &c, float dt,
                                                   const BdfState &b, int par, int64_t ld, uint32_t i, int st,
                                                   const float *S, float *P, float &epsi, int &k,
                                                   const float *ph0) {
    using G = GranuleModel;
    constexpr int ND = G::ND, NV = G::NV;
    const float *Ph = (const float *)b.prev;
    float F[ND], P1[ND], Cc[ND], A[ND * ND];
    const float *Cs = b.C + st * 7;
    const float c0 = Cs[0], c0dt = c0 * dt;
    const float eps = dt * 0.1f;
    const float inv_eps = 1.0f / eps;
    {
        G::Terms t;
        G::prep(mp, c, P, t);
        G::rhs_terms(mp, c, P, t, F);
        bool fin = true;
#pragma unroll
        for (int j = 0; j < ND; j++) {
            Cc[j] = c0 * dt * F[j] + Cs[1] * S[j];
            if (st > 1) Cc[j] += Cs[2] * ph0[j * kBlock];
            for (int q = 2; q < st; q++) Cc[j] += Cs[q + 1] * Ph[((int64_t)(q - 1) * ND + j) * ld + i];
            fin = fin && __builtin_isfinite(F[j]);
        }
        // J = c0*dt*jac - I, jac[z][j] = (F[z] - Fn_j[z]) / eps (one-sided, at P)
#pragma unroll
        for (int q = 0; q < ND * ND; q++) A[q] = 0.0f;
        {   // V: every equation reads it (table row at the perturbed V)
            float X[NV], Fn[ND];
            G::Terms tv;
#pragma unroll
            for (int q = 0; q < NV; q++) X[q] = P[q];
            X[0] -= eps;
            G::load_row(c, X[0], tv.v);
            tv.vca = t.vca;
            G::rhs_terms(mp, c, X, tv, Fn);
#pragma unroll
            for (int z = 0; z < ND; z++) A[z * ND] = c0dt * ((F[z] - Fn[z]) * inv_eps);
        }
#pragma unroll
        for (int j = 1; j < 14; j++) {   // gates: the membrane equation and their own (+ calcium for xCa, yCa)
            float X[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) X[q] = P[q];
            X[j] -= eps;
            A[j] = c0dt * ((F[0] - G::dv(mp, c, X, t.vca)) * inv_eps);
            const float fj = j == 10 ? G::dxkca(mp.gran, X, t.v) : G::dgate(j, X, t.v);
            A[j * ND + j] = c0dt * ((F[j] - fj) * inv_eps);
            if (j == 11 || j == 12) A[14 * ND + j] = c0dt * ((F[14] - G::dca(X, t.vca)) * inv_eps);
        }
        {   // Ca: the membrane equation (Nernst potential), the Ca-dependent K gate, calcium
            float X[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) X[q] = P[q];
            X[14] -= eps;
            const float vca = G::nernst_ca(mp.gran, X[14]);
            A[14] = c0dt * ((F[0] - G::dv(mp, c, X, vca)) * inv_eps);
            A[10 * ND + 14] = c0dt * ((F[10] - G::dxkca(mp.gran, X, t.v)) * inv_eps);
            A[14 * ND + 14] = c0dt * ((F[14] - G::dca(X, vca)) * inv_eps);
        }
#pragma unroll
        for (int z = 0; z < ND; z++) A[z * ND + z] -= 1;
        if (fin && invert_in_place<ND>(A)) {
#pragma unroll
            for (int z = 0; z < ND; z++) {
                float aux = 0.0f;
#pragma unroll
                for (int q = 0; q < ND; q++) aux += A[z * ND + q] * (P[q] - Cc[q]);
                P1[z] = aux + P[z];
            }
        } else if (any_nan<ND>(P)) {
            // An iterate with a NaN component corrects to all NaN whatever
            // the Jacobian: every P1[z] = sum_t iJ[z][t] * (P[t] - Cc[t]) +
            // P[z] takes the NaN term (and the max-norm change below is then
            // 0: no comparison with a NaN holds).  This is the standing case
            // of a neuron whose step once went NaN: the NaN reset restores
            // the state but not the BDF history (TimeDrivenNeuronModel.cpp:
            // 75-88), so every later predictor is NaN.  Only the NaN payload
            // can differ from the full correction's.
#pragma unroll
            for (int q = 0; q < ND; q++) P1[q] = __builtin_nanf("");
        } else {
            const uint32_t slot = atomicAdd(&b.ws_count[par], 1u);
            if (slot < (uint32_t)b.ws_cap) {
                float *w = b.ws + (size_t)slot * kGrWsFloats;
                float *X = w + 450, *Cw = X + NV + 2 * ND, *P1w = Cw + ND;
#pragma unroll
                for (int q = 0; q < NV; q++) X[q] = P[q];
#pragma unroll
                for (int q = 0; q < ND; q++) Cw[q] = Cc[q];
                granule_newton_reference(mp, c, dt, c0, w);
#pragma unroll
                for (int q = 0; q < ND; q++) P1[q] = P1w[q];
            } else {
                atomicOr(b.err, 16u);   // reported to the host; the step is not the reference's
#pragma unroll
                for (int q = 0; q < ND; q++) P1[q] = P[q];
            }
        }
        float aux2 = 0.0f;
#pragma unroll
        for (int z = 0; z < ND; z++) {
            const float aux = fabsf(P1[z] - P[z]);
            if (aux > aux2) aux2 = aux;
        }
#pragma unroll
        for (int z = 0; z < ND; z++) P[z] = P1[z];
        epsi = aux2;
        k++;
    }
}

template <unsigned MASK>
__device__ __forceinline__ void granule_bdf_commit(const ModelParams &mp, const BdfState &b, int64_t ld,
                                                   uint32_t i, int &st, float *S, const float *P) {
    using G = GranuleModel;
    constexpr int ND = G::ND;
    float *Dh = (float *)b.d;
    float *Ph = (float *)b.prev;
    if (st < b.order) st++;
#pragma unroll
    for (int j = 0; j < ND; j++) {
        for (int q = st - 1; q > 0; q--) Dh[((int64_t)q * ND + j) * ld + i] = -Dh[((int64_t)(q - 1) * ND + j) * ld + i];
        Dh[(int64_t)j * ld + i] = P[j] - S[j];
        for (int q = 1; q < st; q++)
            Dh[((int64_t)q * ND + j) * ld + i] += Dh[((int64_t)(q - 1) * ND + j) * ld + i];
    }
    if (st > 1) {
        for (int q = st - 2; q > 0; q--)
#pragma unroll
            for (int j = 0; j < ND; j++)
                Ph[((int64_t)q * ND + j) * ld + i] = Ph[((int64_t)(q - 1) * ND + j) * ld + i];
#pragma unroll
        for (int j = 0; j < ND; j++) Ph[(int64_t)j * ld + i] = S[j];
    }
#pragma unroll
    for (int j = 0; j < ND; j++) S[j] = P[j];
    G::decay<MASK>(mp, S);
}

}  // namespace edl
