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
};

struct LIFParams {
    float e_exc, e_inh, e_leak, v_thr, g_leak, inv_c_m, nmda_aux;
    float cf[3];
    float init[5];
};

struct ModelParams {
    AdExParams adex;
    LIFParams lif;
};

// Per-thread context the right-hand sides need besides the state.
struct RhsCtx {
    const float *exp_lut;
    const float *nmda_lut;
    const float *refr;   // LIF: refractory fractions by steps since the last spike
    int refr_len;
    int since_spike;     // LIF: steps since the last spike (see refractory table)
};

struct AdExModel {
    static constexpr int NV = 6, ND = 2;
    static constexpr int EXC = 2, INH = 3, NMDA = 4, EXT = 5;

    // AdExTimeDrivenModel.cpp:170-218.
    template <unsigned MASK>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c,
                                               const float *s, float *d) {
        const AdExParams &p = mp.adex;
        const float V = s[0];
        if (V <= p.v_cap) {
            float cur = 0.0f;
            if (MASK & 1u) cur += s[EXC] * (p.e_exc - V);
            if (MASK & 2u) cur += s[INH] * (p.e_inh - V);
            if (MASK & 4u) {
                const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
                cur += s[NMDA] * gi * (p.e_exc - V);
            }
            if (MASK & 8u) cur += s[EXT];
            d[0] = (p.g_leak * (p.e_leak - V) + p.glk_thr * lut_exp(c.exp_lut, (V - p.v_thr) * p.inv_thr) -
                    s[1] + cur) * p.inv_c_m;
            d[1] = (p.a * (V - p.e_leak) - s[1]) * p.inv_tau_w;
        } else if (V <= 0.0f) {
            const float Vm = p.v_cap;
            float cur = 0.0f;
            if (MASK & 1u) cur += s[EXC] * (p.e_exc - Vm);
            if (MASK & 2u) cur += s[INH] * (p.e_inh - Vm);
            if (MASK & 4u) {
                const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, Vm);
                cur += s[NMDA] * gi * (p.e_exc - Vm);
            }
            if (MASK & 8u) cur += s[EXT];
            d[0] = (p.g_leak * (p.e_leak - Vm) + p.glk_thr * lut_exp(c.exp_lut, (Vm - p.v_thr) * p.inv_thr) -
                    s[1] + cur) * p.inv_c_m;
            d[1] = (p.a * (V - p.e_leak) - s[1]) * p.inv_tau_w;
        } else {
            d[0] = 0.0f;
            d[1] = 0.0f;
        }
    }
    template <unsigned MASK>
    __device__ static __forceinline__ void decay(const ModelParams &mp, float *s) {
        if (MASK & 1u) s[EXC] = decay_g(s[EXC], mp.adex.cf[0]);
        if (MASK & 2u) s[INH] = decay_g(s[INH], mp.adex.cf[1]);
        if (MASK & 4u) s[NMDA] = decay_g(s[NMDA], mp.adex.cf[2]);
    }
    // AdExTimeDrivenModel.cpp:159-168.
    __device__ static __forceinline__ bool spike(const ModelParams &mp, float *s) {
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
    static constexpr int EXC = 1, INH = 2, NMDA = 3, EXT = 4;

    // LIFTimeDrivenModel.cpp:170-202.  The double-precision refractory clock
    // of the reference is replaced by an integer count of steps since the
    // last spike and a host-built table of the fraction the clock yields for
    // each count (negative = derivative forced to zero); counts past the
    // table's end are outside the refractory period.
    template <unsigned MASK>
    __device__ static __forceinline__ void rhs(const ModelParams &mp, const RhsCtx &c,
                                               const float *s, float *d) {
        const LIFParams &p = mp.lif;
        const float V = s[0];
        float cur = 0.0f;
        if (MASK & 1u) cur += s[EXC] * (p.e_exc - V);
        if (MASK & 2u) cur += s[INH] * (p.e_inh - V);
        if (MASK & 4u) {
            const float gi = nmda_inf(c.nmda_lut, p.nmda_aux, p.e_inh, V);
            cur += s[NMDA] * gi * (p.e_exc - V);
        }
        if (MASK & 8u) cur += s[EXT];
        if (c.since_spike >= c.refr_len) {
            d[0] = (cur + p.g_leak * (p.e_leak - V)) * p.inv_c_m;
        } else {
            const float f = c.refr[c.since_spike];
            d[0] = (f < 0.0f) ? 0.0f : f * ((cur + p.g_leak * (p.e_leak - V)) * p.inv_c_m);
        }
    }
    template <unsigned MASK>
    __device__ static __forceinline__ void decay(const ModelParams &mp, float *s) {
        if (MASK & 1u) s[EXC] = decay_g(s[EXC], mp.lif.cf[0]);
        if (MASK & 2u) s[INH] = decay_g(s[INH], mp.lif.cf[1]);
        if (MASK & 4u) s[NMDA] = decay_g(s[NMDA], mp.lif.cf[2]);
    }
    // LIFTimeDrivenModel.cpp:158-168.
    __device__ static __forceinline__ bool spike(const ModelParams &mp, float *s) {
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

}  // namespace edl
