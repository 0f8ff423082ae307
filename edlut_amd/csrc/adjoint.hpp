// adjoint.hpp -- differentiable trace fitting: forward simulation of
// independent LIF (Euler) or AdEx (RK4) neurons driven by fixed input spike
// trains, and the discrete adjoint of that simulation with checkpointed
// recomputation.
//
// Forward map of one step (the plain path's update, made differentiable):
// events of the step add their weights to g_exc / g_inh; LIF takes an Euler
// step scaled by the refractory fraction, AdEx an RK4 step with conductances
// decayed at the half steps; conductances below 1e-9 snap to zero; a spike
// resets V (LIF to e_leak; AdEx to e_reset, w += b).  The AdEx exponential
// term uses expf (the reference's CUDA kernels do the same,
// AdExTimeDrivenModel_GPU2.cuh:307,313; the CPU table's derivative is zero
// almost everywhere) and decay factors are exp(-h/tau) of each neuron's tau.
//
// One lane per neuron: the step recurrence is sequential, so the kernels are
// latency bound; each neuron's events are walked with a cursor, and the
// backward sweep prefetches the next step's stored values.
//
//   k_fit_forward   V trace [steps][n]; state + event cursor every K steps
//   k_fit_backward  per segment (last first): replay K steps from the
//                   checkpoint storing the stage values, then the reverse
//                   sweep (reverse-mode differentiation of each step with
//                   the forward run's spike pattern held fixed: a reset
//                   stops the V adjoint and, for LIF, routes it to e_leak)
#pragma once
#include "device_common.hpp"

namespace edl {

constexpr int kFitBlock = 64;   // one wave per workgroup: neurons spread over CUs
constexpr int kFitStore = 14;   // floats stored per replayed step

struct FitArgs {
    int model;              // 0 LIF (Euler), 1 AdEx (RK4)
    int n, steps, n_syn, ckpt;
    float dt, h, s6;
    // fixed parameters
    float e_exc, e_inh, v_thr, inv_c_m;          // both
    float a, b, thr, inv_thr, inv_tau_w, e_reset, v_cap;   // AdEx
    const float *refr;                           // LIF refractory fractions
    int refr_len;
    // trainable
    const float *gl, *el, *te, *ti;              // [n]
    const float *w;                              // [n][n_syn]
    const uint8_t *syn_type;                     // [n][n_syn] 0 exc, 1 inh
    // inputs: per neuron, events sorted by step
    const int32_t *ev_ptr, *ev_step, *ev_syn;
    // forward outputs
    float *v_trace;                              // [steps][n]
    uint8_t *spikes;                             // [steps][n] (optional)
    float *traj;                                 // [steps][4][n] state before integration (optional)
    int32_t *traj_since;                         // [steps][n] refractory count (optional)
    float *ck_state;                             // [segments+1][4][n]
    int32_t *ck_int;                             // [segments+1][2][n]: event cursor, last spike
    // backward
    const float *grad_v;                         // [steps][n]
    float *grad_par;                             // [4][n]: g_leak e_leak tau_exc tau_inh
    float *grad_w;                               // [n][n_syn]
    float *scratch;                              // [ckpt][kFitStore][n]
    int32_t *scratch_int;                        // [ckpt][n]: refractory count | spike << 30
};

__device__ __forceinline__ float fit_decay(float g, float cf) { return (g < 1e-9f) ? 0.0f : g * cf; }

__device__ __forceinline__ float lif_frac(const FitArgs &a, int since) {
    if (since >= a.refr_len) return 1.0f;
    const float f = a.refr[since];
    return f < 0.0f ? 0.0f : f;
}

// AdEx right-hand side; E returns the exponential term's exp() (0 above 0 mV)
__device__ __forceinline__ void fit_adex_f(const FitArgs &a, float gl, float el, float V, float W, float ge,
                                           float gi, float &dV, float &dW, float &E) {
    if (V <= a.v_cap) {
        E = expf((V - a.v_thr) * a.inv_thr);
        const float cur = ge * (a.e_exc - V) + gi * (a.e_inh - V);
        dV = (gl * (el - V) + gl * a.thr * E - W + cur) * a.inv_c_m;
        dW = (a.a * (V - el) - W) * a.inv_tau_w;
    } else if (V <= 0.0f) {
        const float Vm = a.v_cap;
        E = expf((Vm - a.v_thr) * a.inv_thr);
        const float cur = ge * (a.e_exc - Vm) + gi * (a.e_inh - Vm);
        dV = (gl * (el - Vm) + gl * a.thr * E - W + cur) * a.inv_c_m;
        dW = (a.a * (V - el) - W) * a.inv_tau_w;
    } else {
        E = 0.0f;
        dV = 0.0f;
        dW = 0.0f;
    }
}

// (ldV, ldW) . d(dV, dW)/d(V, W, ge, gi, gl, el) given the stage point and E
__device__ __forceinline__ void fit_adex_vjp(const FitArgs &a, float gl, float el, float V, float ge,
                                             float gi, float E, float ldV, float ldW, float *lx,
                                             double *lp) {
    if (V <= a.v_cap) {
        lx[0] += ldV * (-gl - ge - gi + gl * a.thr * E * a.inv_thr) * a.inv_c_m + ldW * a.a * a.inv_tau_w;
        lx[1] += -ldV * a.inv_c_m - ldW * a.inv_tau_w;
        lx[2] += ldV * (a.e_exc - V) * a.inv_c_m;
        lx[3] += ldV * (a.e_inh - V) * a.inv_c_m;
        lp[0] += (double)(ldV * ((el - V) + a.thr * E) * a.inv_c_m);
        lp[1] += (double)(ldV * gl * a.inv_c_m - ldW * a.a * a.inv_tau_w);
    } else if (V <= 0.0f) {
        const float Vm = a.v_cap;
        lx[0] += ldW * a.a * a.inv_tau_w;
        lx[1] += -ldV * a.inv_c_m - ldW * a.inv_tau_w;
        lx[2] += ldV * (a.e_exc - Vm) * a.inv_c_m;
        lx[3] += ldV * (a.e_inh - Vm) * a.inv_c_m;
        lp[0] += (double)(ldV * ((el - Vm) + a.thr * E) * a.inv_c_m);
        lp[1] += (double)(ldV * gl * a.inv_c_m - ldW * a.a * a.inv_tau_w);
    }
}

// One forward step of neuron i from state (V, W, ge, gi) after the step's
// inputs; `st` (if non-null) receives the stage values for the adjoint.
__device__ __forceinline__ bool fit_step(const FitArgs &a, float gl, float el, float cfe, float cfi,
                                         int since, float &V, float &W, float &ge, float &gi, float *st) {
    if (a.model == 0) {
        const float frac = lif_frac(a, since);
        const float cur = ge * (a.e_exc - V) + gi * (a.e_inh - V);
        const float f = frac * ((cur + gl * (el - V)) * a.inv_c_m);
        V = V + a.dt * f;
        ge = fit_decay(ge, cfe);
        gi = fit_decay(gi, cfi);
        if (V > a.v_thr) {
            V = el;
            return true;
        }
        return false;
    }
    float d1V, d1W, d2V, d2W, d3V, d3W, d4V, d4W, E1, E2, E3, E4;
    fit_adex_f(a, gl, el, V, W, ge, gi, d1V, d1W, E1);
    const float gem = fit_decay(ge, cfe), gim = fit_decay(gi, cfi);
    const float a1V = V + d1V * a.h, a1W = W + d1W * a.h;
    fit_adex_f(a, gl, el, a1V, a1W, gem, gim, d2V, d2W, E2);
    const float a2V = V + d2V * a.h, a2W = W + d2W * a.h;
    fit_adex_f(a, gl, el, a2V, a2W, gem, gim, d3V, d3W, E3);
    const float ge2 = fit_decay(gem, cfe), gi2 = fit_decay(gim, cfi);
    const float a3V = V + d3V * a.dt, a3W = W + d3W * a.dt;
    fit_adex_f(a, gl, el, a3V, a3W, ge2, gi2, d4V, d4W, E4);
    if (st) {
        st[4] = a1V; st[5] = a1W; st[6] = a2V; st[7] = a2W; st[8] = a3V; st[9] = a3W;
        st[10] = E1; st[11] = E2; st[12] = E3; st[13] = E4;
    }
    V = V + (d1V + 2.0f * (d2V + d3V) + d4V) * a.s6;
    W = W + (d1W + 2.0f * (d2W + d3W) + d4W) * a.s6;
    ge = ge2;
    gi = gi2;
    if (V > 0.0f) {
        V = a.e_reset;
        W = W + a.b;
        return true;
    }
    return false;
}

__global__ __launch_bounds__(kFitBlock) void k_fit_forward(FitArgs a) {
    const int i = blockIdx.x * kFitBlock + threadIdx.x;
    if (i >= a.n) return;
    const float gl = a.gl[i], el = a.el[i];
    const float cfe = expf(-a.h / a.te[i]), cfi = expf(-a.h / a.ti[i]);
    const float *w = a.w + (int64_t)i * a.n_syn;
    const uint8_t *ty = a.syn_type + (int64_t)i * a.n_syn;
    float V = el, W = 0.0f, ge = 0.0f, gi = 0.0f;
    int last = -(1 << 30);
    int e = a.ev_ptr[i];
    const int e_end = a.ev_ptr[i + 1];
    int next_step = e < e_end ? a.ev_step[e] : 0x7fffffff;
    for (int k = 1; k <= a.steps; k++) {
        if ((k - 1) % a.ckpt == 0) {
            const int c = (k - 1) / a.ckpt;
            float *cs = a.ck_state + (int64_t)c * 4 * a.n;
            cs[i] = V;
            cs[a.n + i] = W;
            cs[2 * a.n + i] = ge;
            cs[3 * a.n + i] = gi;
            int32_t *ci = a.ck_int + (int64_t)c * 2 * a.n;
            ci[i] = e;
            ci[a.n + i] = last;
        }
        while (next_step == k) {
            const int s = a.ev_syn[e];
            if (ty[s] == 0) ge += w[s];
            else gi += w[s];
            e++;
            next_step = e < e_end ? a.ev_step[e] : 0x7fffffff;
        }
        if (a.traj) {
            float *y = a.traj + (int64_t)(k - 1) * 4 * a.n + i;
            y[0] = V;
            y[a.n] = W;
            y[2 * (int64_t)a.n] = ge;
            y[3 * (int64_t)a.n] = gi;
            a.traj_since[(int64_t)(k - 1) * a.n + i] = k - 1 - last;
        }
        const bool fired = fit_step(a, gl, el, cfe, cfi, k - 1 - last, V, W, ge, gi, nullptr);
        if (fired) last = k;
        a.v_trace[(int64_t)(k - 1) * a.n + i] = V;
        if (a.spikes) a.spikes[(int64_t)(k - 1) * a.n + i] = fired ? 1 : 0;
    }
}

__global__ __launch_bounds__(kFitBlock) void k_fit_backward(FitArgs a) {
    const int i = blockIdx.x * kFitBlock + threadIdx.x;
    if (i >= a.n) return;
    const float gl = a.gl[i], el = a.el[i];
    const float te = a.te[i], ti = a.ti[i];
    const float cfe = expf(-a.h / te), cfi = expf(-a.h / ti);
    const float *w = a.w + (int64_t)i * a.n_syn;
    const uint8_t *ty = a.syn_type + (int64_t)i * a.n_syn;
    float *gw = a.grad_w + (int64_t)i * a.n_syn;
    for (int s = 0; s < a.n_syn; s++) gw[s] = 0.0f;
    const int n = a.n;
    const int segs = (a.steps + a.ckpt - 1) / a.ckpt;
    float lV = 0.0f, lW = 0.0f, lge = 0.0f, lgi = 0.0f;
    double lp[2] = {0.0, 0.0};
    double lcfe = 0.0, lcfi = 0.0;
    const int e_begin = a.ev_ptr[i];
    for (int c = segs - 1; c >= 0; c--) {
        // 1. replay the segment from its checkpoint, storing what the reverse needs
        const float *cs = a.ck_state + (int64_t)c * 4 * n;
        float V = cs[i], W = cs[n + i], ge = cs[2 * n + i], gi = cs[3 * n + i];
        const int32_t *ci = a.ck_int + (int64_t)c * 2 * n;
        int e = ci[i];
        int last = ci[n + i];
        const int e_end = a.ev_ptr[i + 1];
        int next_step = e < e_end ? a.ev_step[e] : 0x7fffffff;
        const int k0 = c * a.ckpt + 1;
        const int k1 = min(a.steps, (c + 1) * a.ckpt);
        for (int k = k0; k <= k1; k++) {
            while (next_step == k) {
                const int s = a.ev_syn[e];
                if (ty[s] == 0) ge += w[s];
                else gi += w[s];
                e++;
                next_step = e < e_end ? a.ev_step[e] : 0x7fffffff;
            }
            float st[kFitStore];
            st[0] = V; st[1] = W; st[2] = ge; st[3] = gi;
            const int since = k - 1 - last;
            const bool fired = fit_step(a, gl, el, cfe, cfi, since, V, W, ge, gi, st);
            if (fired) last = k;
            float *sc = a.scratch + (int64_t)(k - k0) * kFitStore * n + i;
            const int nst = a.model == 0 ? 4 : kFitStore;
            for (int q = 0; q < nst; q++) sc[(int64_t)q * n] = st[q];   // LIF needs the state only
            a.scratch_int[(int64_t)(k - k0) * n + i] = (since < (1 << 29) ? since : (1 << 29)) | (fired ? (1 << 30) : 0);
        }
        // 2. reverse sweep over the segment.  The stored values, dL/dV and
        // the event cursor of step k-1 are fetched while step k is
        // differentiated: nothing on the adjoint's dependency chain waits on
        // memory.
        int er = e - 1;   // last event of the segment
        int er_step = er >= e_begin ? a.ev_step[er] : -1;
        const int nst = a.model == 0 ? 4 : kFitStore;
        float nx[kFitStore];
        float gv_n;
        int32_t si_n;
        {
            const float *sc = a.scratch + (int64_t)(k1 - k0) * kFitStore * n + i;
#pragma unroll
            for (int q = 0; q < kFitStore; q++) nx[q] = q < nst ? sc[(int64_t)q * n] : 0.0f;
            gv_n = a.grad_v[(int64_t)(k1 - 1) * n + i];
            si_n = a.scratch_int[(int64_t)(k1 - k0) * n + i];
        }
        for (int k = k1; k >= k0; k--) {
            float cur[kFitStore];
#pragma unroll
            for (int q = 0; q < kFitStore; q++) cur[q] = nx[q];
            const float gvk = gv_n;
            const int32_t si = si_n;
            if (k > k0) {
                const float *sc = a.scratch + (int64_t)(k - 1 - k0) * kFitStore * n + i;
#pragma unroll
                for (int q = 0; q < kFitStore; q++) nx[q] = q < nst ? sc[(int64_t)q * n] : 0.0f;
                gv_n = a.grad_v[(int64_t)(k - 2) * n + i];
                si_n = a.scratch_int[(int64_t)(k - 1 - k0) * n + i];
            }
            lV += gvk;
            const bool fired = (si >> 30) & 1;
            const float Vy = cur[0], Wy = cur[1], gey = cur[2], giy = cur[3];
            (void)Wy;
            if (fired) {
                if (a.model == 0) lp[1] += (double)lV;
                lV = 0.0f;
            }
            float nV, nW = 0.0f, nge, ngi;
            if (a.model == 0) {
                const float frac = lif_frac(a, si & ((1 << 30) - 1));
                const float ld = a.dt * lV;
                nV = lV + ld * frac * (-gey - giy - gl) * a.inv_c_m;
                nge = ld * frac * (a.e_exc - Vy) * a.inv_c_m;
                ngi = ld * frac * (a.e_inh - Vy) * a.inv_c_m;
                lp[0] += (double)(ld * frac * (el - Vy) * a.inv_c_m);
                lp[1] += (double)(ld * frac * gl * a.inv_c_m);
                if (!(gey < 1e-9f)) {
                    nge += lge * cfe;
                    lcfe += (double)(lge * gey);
                }
                if (!(giy < 1e-9f)) {
                    ngi += lgi * cfi;
                    lcfi += (double)(lgi * giy);
                }
            } else {
                const float a1V = cur[4], a2V = cur[6], a3V = cur[8];
                const float E1 = cur[10], E2 = cur[11], E3 = cur[12], E4 = cur[13];
                const float gem = fit_decay(gey, cfe), gim = fit_decay(giy, cfi);
                const float ge2 = fit_decay(gem, cfe), gi2 = fit_decay(gim, cfi);
                float l4V = lV * a.s6, l4W = lW * a.s6;
                float l3V = 2.0f * lV * a.s6, l3W = 2.0f * lW * a.s6;
                float l2V = l3V, l2W = l3W;
                float l1V = l4V, l1W = l4W;
                nV = lV;
                nW = lW;
                float lg2e = lge, lg2i = lgi, lgme = 0.0f, lgmi = 0.0f;
                float lx[4];
                lx[0] = lx[1] = lx[2] = lx[3] = 0.0f;
                fit_adex_vjp(a, gl, el, a3V, ge2, gi2, E4, l4V, l4W, lx, lp);
                nV += lx[0]; nW += lx[1]; l3V += a.dt * lx[0]; l3W += a.dt * lx[1];
                lg2e += lx[2]; lg2i += lx[3];
                lx[0] = lx[1] = lx[2] = lx[3] = 0.0f;
                fit_adex_vjp(a, gl, el, a2V, gem, gim, E3, l3V, l3W, lx, lp);
                nV += lx[0]; nW += lx[1]; l2V += a.h * lx[0]; l2W += a.h * lx[1];
                lgme += lx[2]; lgmi += lx[3];
                lx[0] = lx[1] = lx[2] = lx[3] = 0.0f;
                fit_adex_vjp(a, gl, el, a1V, gem, gim, E2, l2V, l2W, lx, lp);
                nV += lx[0]; nW += lx[1]; l1V += a.h * lx[0]; l1W += a.h * lx[1];
                lgme += lx[2]; lgmi += lx[3];
                lx[0] = lx[1] = lx[2] = lx[3] = 0.0f;
                fit_adex_vjp(a, gl, el, Vy, gey, giy, E1, l1V, l1W, lx, lp);
                nV += lx[0]; nW += lx[1];
                nge = lx[2];
                ngi = lx[3];
                if (!(gem < 1e-9f)) { lgme += lg2e * cfe; lcfe += (double)(lg2e * gem); }
                if (!(gim < 1e-9f)) { lgmi += lg2i * cfi; lcfi += (double)(lg2i * gim); }
                if (!(gey < 1e-9f)) { nge += lgme * cfe; lcfe += (double)(lgme * gey); }
                if (!(giy < 1e-9f)) { ngi += lgmi * cfi; lcfi += (double)(lgmi * giy); }
            }
            lV = nV;
            lW = nW;
            lge = nge;
            lgi = ngi;
            while (er_step == k) {
                const int s = a.ev_syn[er];
                gw[s] += ty[s] == 0 ? lge : lgi;
                er--;
                er_step = er >= e_begin ? a.ev_step[er] : -1;
            }
        }
    }
    // initial state V = e_leak
    lp[1] += (double)lV;
    a.grad_par[i] = (float)lp[0];
    a.grad_par[n + i] = (float)lp[1];
    // decay factors exp(-h/tau): d/dtau = cf * h / tau^2
    a.grad_par[2 * n + i] = (float)(lcfe * (double)cfe * (double)a.h / ((double)te * (double)te));
    a.grad_par[3 * n + i] = (float)(lcfi * (double)cfi * (double)a.h / ((double)ti * (double)ti));
}

}  // namespace edl
