/*
 * adjoint_oracle.cpp -- CPU reference of the trace-fitting path (FP64 and
 * float): forward simulation of independent LIF (Euler) or AdEx (RK4)
 * neurons driven by fixed input spike trains, and the discrete adjoint of
 * that simulation.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint kernels; used by
 * tests/ alone.  The reference has no gradient path, so this code is pinned
 * by central finite differences of its own forward map, not by reference
 * output.
 *
 * The forward map of one step is the time-driven update of the plain path:
 * input events add their weights to g_exc / g_inh before the step; LIF takes
 * an Euler step (Euler.h:123-146) with the refractory fraction of the step,
 * AdEx an RK4 step with conductances decayed at the half steps (RK4.h:
 * 167-231); conductances below 1e-9 snap to zero (AdExTimeDrivenModel.cpp:
 * 220-245); a spike resets V (LIF to e_leak, AdEx to e_reset with w += b).
 * Two differences from the plain path make the map differentiable: the AdEx
 * exponential term uses exp() instead of the look-up table (as the
 * reference's CUDA kernels do, AdExTimeDrivenModel_GPU2.cuh:307,313), and
 * the decay factors exp(-h/tau) are computed from each neuron's own tau.
 *
 * Trainable per neuron: g_leak, e_leak, tau_exc, tau_inh; per synapse: the
 * weight.  The adjoint is reverse-mode differentiation of the discrete map
 * with the spike pattern of the forward run held fixed: the reset's jump
 * (V set to a constant) stops the V adjoint and, for LIF, routes it to
 * e_leak; the threshold test itself has zero derivative.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <typename T> struct Fixed {
    // LIF: e_exc e_inh v_thr c_m (refractory from `refr`)
    // AdEx: a b thr_slo_fac v_thr tau_w e_exc e_inh e_reset c_m
    T e_exc, e_inh, v_thr, inv_c_m, a, b, thr, inv_thr, inv_tau_w, e_reset, v_cap;
};

template <typename T> struct Par {
    T gl, el, te, ti;
};

template <typename T> inline T decay(T g, T cf, bool &live) {
    live = !(g < (T)1e-9f);
    return live ? g * cf : (T)0;
}

/* ---- LIF ---------------------------------------------------------------- */
template <typename T> inline T lif_f(const Fixed<T> &c, const Par<T> &p, T frac, T V, T ge, T gi) {
    T cur = ge * (c.e_exc - V) + gi * (c.e_inh - V);
    return frac * ((cur + p.gl * (p.el - V)) * c.inv_c_m);
}

/* ---- AdEx --------------------------------------------------------------- */
template <typename T> inline void adex_f(const Fixed<T> &c, const Par<T> &p, T V, T w, T ge, T gi,
                                         T &dV, T &dw) {
    if (V <= c.v_cap) {
        T cur = ge * (c.e_exc - V) + gi * (c.e_inh - V);
        dV = (p.gl * (p.el - V) + p.gl * c.thr * std::exp((V - c.v_thr) * c.inv_thr) - w + cur) * c.inv_c_m;
        dw = (c.a * (V - p.el) - w) * c.inv_tau_w;
    } else if (V <= (T)0) {
        const T Vm = c.v_cap;
        T cur = ge * (c.e_exc - Vm) + gi * (c.e_inh - Vm);
        dV = (p.gl * (p.el - Vm) + p.gl * c.thr * std::exp((Vm - c.v_thr) * c.inv_thr) - w + cur) * c.inv_c_m;
        dw = (c.a * (V - p.el) - w) * c.inv_tau_w;
    } else {
        dV = 0;
        dw = 0;
    }
}

/* VJP of adex_f: adds (ldV, ldw) . d(dV, dw)/d(V, w, ge, gi, gl, el). */
template <typename T> inline void adex_f_vjp(const Fixed<T> &c, const Par<T> &p, T V, T w, T ge, T gi,
                                             T ldV, T ldw, T *lx, T *lp) {
    (void)w;
    if (V <= c.v_cap) {
        const T E = std::exp((V - c.v_thr) * c.inv_thr);
        lx[0] += ldV * (-p.gl - ge - gi + p.gl * c.thr * E * c.inv_thr) * c.inv_c_m + ldw * c.a * c.inv_tau_w;
        lx[1] += -ldV * c.inv_c_m - ldw * c.inv_tau_w;
        lx[2] += ldV * (c.e_exc - V) * c.inv_c_m;
        lx[3] += ldV * (c.e_inh - V) * c.inv_c_m;
        lp[0] += ldV * ((p.el - V) + c.thr * E) * c.inv_c_m;
        lp[1] += ldV * p.gl * c.inv_c_m - ldw * c.a * c.inv_tau_w;
    } else if (V <= (T)0) {
        const T Vm = c.v_cap;
        const T E = std::exp((Vm - c.v_thr) * c.inv_thr);
        lx[0] += ldw * c.a * c.inv_tau_w;
        lx[1] += -ldV * c.inv_c_m - ldw * c.inv_tau_w;
        lx[2] += ldV * (c.e_exc - Vm) * c.inv_c_m;
        lx[3] += ldV * (c.e_inh - Vm) * c.inv_c_m;
        lp[0] += ldV * ((p.el - Vm) + c.thr * E) * c.inv_c_m;
        lp[1] += ldV * p.gl * c.inv_c_m - ldw * c.a * c.inv_tau_w;
    }
}

template <typename T> struct Problem {
    int model;   /* 0 LIF, 1 AdEx */
    int n, steps, n_syn;
    T dt;
    Fixed<T> c;
    std::vector<T> refr;   /* LIF: refractory fraction by steps since the spike */
    const int32_t *ev_ptr, *ev_step, *ev_syn;
    const uint8_t *syn_type;
};

template <typename T> inline T cf_of(T h, T tau) { return std::exp(-h / tau); }

/* one neuron, whole run: V trace, spike flags, and (optionally) the state
 * before each step's integration (after inputs) */
template <typename T> void forward_neuron(const Problem<T> &P, int i, const Par<T> &p, const T *w,
                                          T *v_out, uint8_t *spk_out, size_t stride, size_t col,
                                          std::vector<T> *y_store, std::vector<int> *since_store) {
    const bool adex = P.model == 1;
    const T h = adex ? P.dt * (T)0.5 : P.dt;
    const T cfe = cf_of(h, p.te), cfi = cf_of(h, p.ti);
    T V = p.el, W = 0, ge = 0, gi = 0;
    int last = -(1 << 30);
    int e = P.ev_ptr[i];
    for (int k = 1; k <= P.steps; k++) {
        for (; e < P.ev_ptr[i + 1] && P.ev_step[e] == k; e++) {
            const int s = P.ev_syn[e];
            if (P.syn_type[(size_t)i * P.n_syn + s] == 0) ge += w[s];
            else gi += w[s];
        }
        if (y_store) {
            T *y = &(*y_store)[(size_t)(k - 1) * 4];
            y[0] = V; y[1] = W; y[2] = ge; y[3] = gi;
        }
        bool live;
        bool fired;
        if (!adex) {
            const int since = k - 1 - last;
            if (since_store) (*since_store)[(size_t)k - 1] = since;
            T frac = since >= (int)P.refr.size() ? (T)1 : (P.refr[since] < 0 ? (T)0 : P.refr[since]);
            const T f = lif_f(P.c, p, frac, V, ge, gi);
            V = V + P.dt * f;
            ge = decay(ge, cfe, live);
            gi = decay(gi, cfi, live);
            fired = V > P.c.v_thr;
            if (fired) {
                V = p.el;
                last = k;
            }
        } else {
            T d1V, d1w, d2V, d2w, d3V, d3w, d4V, d4w;
            adex_f(P.c, p, V, W, ge, gi, d1V, d1w);
            const T gem = decay(ge, cfe, live), gim = decay(gi, cfi, live);
            adex_f(P.c, p, V + d1V * h, W + d1w * h, gem, gim, d2V, d2w);
            adex_f(P.c, p, V + d2V * h, W + d2w * h, gem, gim, d3V, d3w);
            const T ge2 = decay(gem, cfe, live), gi2 = decay(gim, cfi, live);
            adex_f(P.c, p, V + d3V * P.dt, W + d3w * P.dt, ge2, gi2, d4V, d4w);
            const T s6 = P.dt * (T)0.166666666667f;
            V = V + (d1V + (T)2 * (d2V + d3V) + d4V) * s6;
            W = W + (d1w + (T)2 * (d2w + d3w) + d4w) * s6;
            ge = ge2;
            gi = gi2;
            fired = V > (T)0;
            if (fired) {
                V = P.c.e_reset;
                W = W + P.c.b;
            }
        }
        v_out[(size_t)(k - 1) * stride + col] = V;
        if (spk_out) spk_out[(size_t)(k - 1) * stride + col] = fired ? 1 : 0;
    }
}

/* adjoint of one neuron; grad_v[(k-1)*n + i] = dL/dV_k.  With `traj` the
 * states before each step's integration, the spike flags and the refractory
 * counts come from a given run ([steps][4][n], [steps][n], [steps][n])
 * instead of this precision's own forward pass. */
template <typename T> void backward_neuron(const Problem<T> &P, int i, const Par<T> &p, const T *w,
                                           const T *grad_v, T *gp /*[4]*/, T *gw /*[n_syn]*/,
                                           const float *traj = nullptr, const uint8_t *traj_spk = nullptr,
                                           const int32_t *traj_since = nullptr) {
    const bool adex = P.model == 1;
    std::vector<T> y((size_t)P.steps * 4), vtr((size_t)P.steps);
    std::vector<uint8_t> spk((size_t)P.steps);
    std::vector<int> since((size_t)P.steps);
    if (traj) {
        for (int k = 0; k < P.steps; k++) {
            for (int q = 0; q < 4; q++) y[(size_t)k * 4 + q] = (T)traj[((size_t)k * 4 + q) * P.n + i];
            spk[(size_t)k] = traj_spk[(size_t)k * P.n + i];
            since[(size_t)k] = traj_since ? traj_since[(size_t)k * P.n + i] : 1 << 30;
        }
    } else {
        forward_neuron(P, i, p, w, vtr.data(), spk.data(), 1, 0, &y, &since);
    }
    const T h = adex ? P.dt * (T)0.5 : P.dt;
    const T cfe = cf_of(h, p.te), cfi = cf_of(h, p.ti);
    T lV = 0, lW = 0, lge = 0, lgi = 0;
    T lp[2] = {0, 0};   /* gl, el */
    T lcfe = 0, lcfi = 0;
    int e = P.ev_ptr[i + 1] - 1;
    for (int k = P.steps; k >= 1; k--) {
        lV += grad_v[(size_t)(k - 1) * P.n + i];
        const T *yk = &y[(size_t)(k - 1) * 4];
        const T V = yk[0], W = yk[1], ge = yk[2], gi = yk[3];
        const bool fired = spk[(size_t)(k - 1)] != 0;
        /* reset */
        if (fired) {
            if (!adex) lp[1] += lV;
            lV = 0;
        }
        T nV = 0, nW = 0, nge = 0, ngi = 0;
        if (!adex) {
            const int s = since[(size_t)k - 1];
            T frac = s >= (int)P.refr.size() ? (T)1 : (P.refr[s] < 0 ? (T)0 : P.refr[s]);
            const T ld = P.dt * lV;
            nV = lV + ld * frac * (-ge - gi - p.gl) * P.c.inv_c_m;
            nge = ld * frac * (P.c.e_exc - V) * P.c.inv_c_m;
            ngi = ld * frac * (P.c.e_inh - V) * P.c.inv_c_m;
            lp[0] += ld * frac * (p.el - V) * P.c.inv_c_m;
            lp[1] += ld * frac * p.gl * P.c.inv_c_m;
            if (!(ge < (T)1e-9f)) { nge += lge * cfe; lcfe += lge * ge; }
            if (!(gi < (T)1e-9f)) { ngi += lgi * cfi; lcfi += lgi * gi; }
        } else {
            bool live;
            T d1V, d1w, d2V, d2w, d3V, d3w, d4V, d4w;
            adex_f(P.c, p, V, W, ge, gi, d1V, d1w);
            const T gem = decay(ge, cfe, live), gim = decay(gi, cfi, live);
            const T a1V = V + d1V * h, a1W = W + d1w * h;
            adex_f(P.c, p, a1V, a1W, gem, gim, d2V, d2w);
            const T a2V = V + d2V * h, a2W = W + d2w * h;
            adex_f(P.c, p, a2V, a2W, gem, gim, d3V, d3w);
            const T ge2 = decay(gem, cfe, live), gi2 = decay(gim, cfi, live);
            const T a3V = V + d3V * P.dt, a3W = W + d3w * P.dt;
            (void)d4V;
            (void)d4w;
            const T s6 = P.dt * (T)0.166666666667f;
            T l4[2] = {lV * s6, lW * s6};
            T l3[2] = {2 * lV * s6, 2 * lW * s6};
            T l2[2] = {2 * lV * s6, 2 * lW * s6};
            T l1[2] = {lV * s6, lW * s6};
            nV = lV;
            nW = lW;
            T lg2[2] = {lge, lgi};   /* twice-decayed conductances */
            T lgm[2] = {0, 0};       /* once-decayed */
            T lx[4];
            /* d4 = f(a3) */
            lx[0] = lx[1] = lx[2] = lx[3] = 0;
            adex_f_vjp(P.c, p, a3V, a3W, ge2, gi2, l4[0], l4[1], lx, lp);
            nV += lx[0]; nW += lx[1]; l3[0] += P.dt * lx[0]; l3[1] += P.dt * lx[1];
            lg2[0] += lx[2]; lg2[1] += lx[3];
            /* d3 = f(a2) */
            lx[0] = lx[1] = lx[2] = lx[3] = 0;
            adex_f_vjp(P.c, p, a2V, a2W, gem, gim, l3[0], l3[1], lx, lp);
            nV += lx[0]; nW += lx[1]; l2[0] += h * lx[0]; l2[1] += h * lx[1];
            lgm[0] += lx[2]; lgm[1] += lx[3];
            /* d2 = f(a1) */
            lx[0] = lx[1] = lx[2] = lx[3] = 0;
            adex_f_vjp(P.c, p, a1V, a1W, gem, gim, l2[0], l2[1], lx, lp);
            nV += lx[0]; nW += lx[1]; l1[0] += h * lx[0]; l1[1] += h * lx[1];
            lgm[0] += lx[2]; lgm[1] += lx[3];
            /* d1 = f(y) */
            lx[0] = lx[1] = lx[2] = lx[3] = 0;
            adex_f_vjp(P.c, p, V, W, ge, gi, l1[0], l1[1], lx, lp);
            nV += lx[0]; nW += lx[1];
            nge = lx[2];
            ngi = lx[3];
            /* conductance decays: g2 = D(gm), gm = D(g) */
            if (!(gem < (T)1e-9f)) { lgm[0] += lg2[0] * cfe; lcfe += lg2[0] * gem; }
            if (!(gim < (T)1e-9f)) { lgm[1] += lg2[1] * cfi; lcfi += lg2[1] * gim; }
            if (!(ge < (T)1e-9f)) { nge += lgm[0] * cfe; lcfe += lgm[0] * ge; }
            if (!(gi < (T)1e-9f)) { ngi += lgm[1] * cfi; lcfi += lgm[1] * gi; }
        }
        lV = nV;
        lW = nW;
        lge = nge;
        lgi = ngi;
        /* inputs of step k: y = x + w */
        for (; e >= P.ev_ptr[i] && P.ev_step[e] == k; e--) {
            const int s = P.ev_syn[e];
            gw[s] += P.syn_type[(size_t)i * P.n_syn + s] == 0 ? lge : lgi;
        }
    }
    /* initial state V = e_leak */
    lp[1] += lV;
    gp[0] = lp[0];
    gp[1] = lp[1];
    /* cf = exp(-h/tau): dcf/dtau = cf * h / tau^2 */
    gp[2] = lcfe * cfe * h / (p.te * p.te);
    gp[3] = lcfi * cfi * h / (p.ti * p.ti);
}

template <typename T> Problem<T> make_problem(int model, const float *fixed, const float *refr, int refr_len,
                                              int n, int steps, int n_syn, float dt_s,
                                              const int32_t *ev_ptr, const int32_t *ev_step,
                                              const int32_t *ev_syn, const uint8_t *syn_type) {
    Problem<T> P;
    P.model = model;
    P.n = n;
    P.steps = steps;
    P.n_syn = n_syn;
    P.dt = (T)(dt_s * 1000.0f);
    std::memset(&P.c, 0, sizeof(P.c));
    if (model == 0) {
        P.c.e_exc = fixed[0]; P.c.e_inh = fixed[1]; P.c.v_thr = fixed[2]; P.c.inv_c_m = (T)1 / (T)fixed[3];
        P.refr.assign(refr, refr + refr_len);
    } else {
        P.c.a = fixed[0]; P.c.b = fixed[1]; P.c.thr = fixed[2]; P.c.inv_thr = (T)1 / (T)fixed[2];
        P.c.v_thr = fixed[3]; P.c.inv_tau_w = (T)1 / (T)fixed[4]; P.c.e_exc = fixed[5]; P.c.e_inh = fixed[6];
        P.c.e_reset = fixed[7]; P.c.inv_c_m = (T)1 / (T)fixed[8];
        P.c.v_cap = (T)fixed[3] + (T)6 * (T)fixed[2];
    }
    P.ev_ptr = ev_ptr;
    P.ev_step = ev_step;
    P.ev_syn = ev_syn;
    P.syn_type = syn_type;
    return P;
}

}  // namespace

extern "C" {

/* Forward in double (fp64 != 0) or float: V trace [steps][n] as double. */
void orc_fit_forward(int model, const float *fixed, const float *refr, int refr_len, int n, int steps,
                     int n_syn, float dt_s, const int32_t *ev_ptr, const int32_t *ev_step,
                     const int32_t *ev_syn, const uint8_t *syn_type, const double *par /*[4][n]*/,
                     const double *w /*[n][n_syn]*/, int fp64, double *v_out, uint8_t *spk_out) {
    if (fp64) {
        Problem<double> P = make_problem<double>(model, fixed, refr, refr_len, n, steps, n_syn, dt_s, ev_ptr,
                                                 ev_step, ev_syn, syn_type);
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < n; i++) {
            Par<double> p{par[i], par[n + i], par[2 * n + i], par[3 * n + i]};
            forward_neuron(P, i, p, w + (size_t)i * n_syn, v_out, spk_out, (size_t)n, (size_t)i,
                           (std::vector<double> *)nullptr, (std::vector<int> *)nullptr);
        }
    } else {
        Problem<float> P = make_problem<float>(model, fixed, refr, refr_len, n, steps, n_syn, dt_s, ev_ptr,
                                               ev_step, ev_syn, syn_type);
        std::vector<float> vf((size_t)steps * n);
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < n; i++) {
            Par<float> p{(float)par[i], (float)par[n + i], (float)par[2 * n + i], (float)par[3 * n + i]};
            std::vector<float> wf(w + (size_t)i * n_syn, w + (size_t)(i + 1) * n_syn);
            forward_neuron(P, i, p, wf.data(), vf.data(), spk_out, (size_t)n, (size_t)i,
                           (std::vector<float> *)nullptr, (std::vector<int> *)nullptr);
        }
        for (size_t q = 0; q < vf.size(); q++) v_out[q] = vf[q];
    }
}

/* Float forward that also records the states before each step's
 * integration ([steps][4][n]) and the refractory counts ([steps][n]). */
void orc_fit_trajectory(int model, const float *fixed, const float *refr, int refr_len, int n, int steps,
                        int n_syn, float dt_s, const int32_t *ev_ptr, const int32_t *ev_step,
                        const int32_t *ev_syn, const uint8_t *syn_type, const double *par, const double *w,
                        float *traj, uint8_t *spk, int32_t *since_out) {
    Problem<float> P = make_problem<float>(model, fixed, refr, refr_len, n, steps, n_syn, dt_s, ev_ptr,
                                           ev_step, ev_syn, syn_type);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; i++) {
        Par<float> p{(float)par[i], (float)par[n + i], (float)par[2 * n + i], (float)par[3 * n + i]};
        std::vector<float> wf(w + (size_t)i * n_syn, w + (size_t)(i + 1) * n_syn);
        std::vector<float> y((size_t)steps * 4), v((size_t)steps);
        std::vector<uint8_t> sp((size_t)steps);
        std::vector<int> since((size_t)steps);
        forward_neuron(P, i, p, wf.data(), v.data(), sp.data(), 1, 0, &y, &since);
        for (int k = 0; k < steps; k++) {
            for (int q = 0; q < 4; q++) traj[((size_t)k * 4 + q) * n + i] = y[(size_t)k * 4 + q];
            spk[(size_t)k * n + i] = sp[(size_t)k];
            since_out[(size_t)k * n + i] = since[(size_t)k];
        }
    }
}

/* FP64 adjoint along a given trajectory (orc_fit_trajectory's layout). */
void orc_fit_backward_along(int model, const float *fixed, const float *refr, int refr_len, int n, int steps,
                            int n_syn, float dt_s, const int32_t *ev_ptr, const int32_t *ev_step,
                            const int32_t *ev_syn, const uint8_t *syn_type, const double *par,
                            const double *w, const double *grad_v, const float *traj, const uint8_t *spk,
                            const int32_t *since, double *grad_par, double *grad_w) {
    Problem<double> P = make_problem<double>(model, fixed, refr, refr_len, n, steps, n_syn, dt_s, ev_ptr,
                                             ev_step, ev_syn, syn_type);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; i++) {
        Par<double> p{par[i], par[n + i], par[2 * n + i], par[3 * n + i]};
        double gp[4];
        std::vector<double> gw((size_t)n_syn, 0.0);
        backward_neuron(P, i, p, w + (size_t)i * n_syn, grad_v, gp, gw.data(), traj, spk, since);
        for (int q = 0; q < 4; q++) grad_par[(size_t)q * n + i] = gp[q];
        for (int s = 0; s < n_syn; s++) grad_w[(size_t)i * n_syn + s] = gw[(size_t)s];
    }
}

/* FP64 adjoint: gradients of sum_k grad_v[k] . V_k. */
void orc_fit_backward(int model, const float *fixed, const float *refr, int refr_len, int n, int steps,
                      int n_syn, float dt_s, const int32_t *ev_ptr, const int32_t *ev_step,
                      const int32_t *ev_syn, const uint8_t *syn_type, const double *par,
                      const double *w, const double *grad_v, double *grad_par /*[4][n]*/,
                      double *grad_w /*[n][n_syn]*/) {
    Problem<double> P = make_problem<double>(model, fixed, refr, refr_len, n, steps, n_syn, dt_s, ev_ptr,
                                             ev_step, ev_syn, syn_type);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; i++) {
        Par<double> p{par[i], par[n + i], par[2 * n + i], par[3 * n + i]};
        double gp[4];
        std::vector<double> gw((size_t)n_syn, 0.0);
        backward_neuron(P, i, p, w + (size_t)i * n_syn, grad_v, gp, gw.data());
        for (int q = 0; q < 4; q++) grad_par[(size_t)q * n + i] = gp[q];
        for (int s = 0; s < n_syn; s++) grad_w[(size_t)i * n_syn + s] = gw[(size_t)s];
    }
}

}  // extern "C"
