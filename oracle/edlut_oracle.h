/*
 * edlut_oracle.h -- CPU restatement of EDLUT's time-driven neuron update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in edlut_amd/ is compared against.  Only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * path never calls it and has no CPU fallback.
 *
 * Every routine restates the reference's CPU algorithm (cited file:line,
 * relative to the reference's repository root) with the reference's own
 * floating-point association, look-up tables and integer indexing, so that a
 * float build of this library reproduces the reference bit for bit.  The same
 * code instantiated with double arithmetic and exact exponentials is the FP64
 * oracle used to state each model's discretisation error.
 *
 * Plain C ABI so that Python (ctypes) and C++ test drivers can call it.
 */
#ifndef EDLUT_ORACLE_H_
#define EDLUT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model and integrator identifiers (same numbering as include/edlut_amd.h). */
enum { ORC_LIF = 0, ORC_ADEX = 1, ORC_HH = 2, ORC_GRANULE = 3, ORC_PURKINJE = 4, ORC_IZHIKEVICH = 5,
       ORC_LIF_IS = 6 };
enum { ORC_EULER = 0, ORC_RK2 = 1, ORC_RK4 = 2, ORC_BDF = 3,
       /* GPU Bifixed methods: n_steps sub-steps of the base method per update */
       ORC_BIFIXED_EULER = 4, ORC_BIFIXED_RK2 = 5, ORC_BIFIXED_RK4 = 6 };

/* ---- look-up tables ---------------------------------------------------- */
/* ExponentialTable::LookUpTable (16384 floats). */
void orc_exp_table(float *out);
/* ExponentialTable::GetResult(x) for each x. */
void orc_exp_lut(const float *x, float *out, int n);
/* g_nmda_inf table (256 floats) for the given reversal potentials. */
void orc_nmda_table(float e_exc, float e_inh, float *out);

/* HH channel-rate table (262144 x {alpha_m, beta_m, alpha_h, beta_h,
 * alpha_n, beta_n} on [-90, 50] mV) for the given v_thr. */
void orc_hh_table(float v_thr, float *out);

/* Granule channel table (1024 x 26 on [-100, 50] mV) and Purkinje channel
 * table (1024 x {alpha_ca, inv_tau_ca, alpha_M, inv_tau_M} on [-100, 35] mV). */
void orc_granule_table(float *out);
void orc_purkinje_table(float *out);
/* The correctly rounded float logarithm the granule restatement uses. */
void orc_log_cr(const float *x, float *out, int n);

/* ---- RandomGenerator / Poisson device ----------------------------------- */
/* Seeds handed to n successive RandomGenerator objects after srand(seed). */
void orc_lcg_seeds(uint32_t seed, int n, uint32_t *out);
/* n successive frand() values of one generator starting from state x0. */
void orc_lcg_frand(uint32_t x0, int n, float *out);

/* ---- one population ----------------------------------------------------- */
/* fp64 != 0 builds the FP64 oracle (double arithmetic, exact exp).
 * bdf_order is the BDF order, or n_steps (>= 1) for the Bifixed methods. */
void *orc_pop_create(int model, int integrator, int bdf_order, const float *params,
                     int n_params, float dt_s, int n_neurons, uint32_t syn_mask, int fp64);
void orc_pop_destroy(void *pop);
int orc_pop_nvars(void *pop);
/* State variable `var` of every neuron, widened to double. */
void orc_pop_get_state(void *pop, int var, double *out);
void orc_pop_set_state(void *pop, int var, const double *in);
/* ProcessInputSpike: state[target][n_diff + type] += w, in call order. */
void orc_pop_add_inputs(void *pop, const int32_t *target, const uint8_t *type,
                        const float *w, int n);
/* Total external current (pA) of each neuron (type 3 input). */
void orc_pop_set_ext_current(void *pop, const float *current);
/* Integrate one step; writes ascending spiking indices; returns their count. */
int orc_pop_update(void *pop, int64_t step, int32_t *spikes);

/* ---- network with canonical spike delivery ------------------------------ */
void *orc_net_create(void *pop, int max_delay);
void orc_net_destroy(void *net);
/* Recurrent source: the population itself.  row_ptr has n*max_delay+1
 * entries; synapses of source s with delay d occupy
 * [row_ptr[s*max_delay+d-1], row_ptr[s*max_delay+d]) in ascending target
 * order; tgt_type packs target | type << 30. */
int orc_net_add_source_self(void *net, const int64_t *row_ptr, const uint32_t *tgt_type,
                            const float *w);
/* Poisson generator device (one RandomGenerator per source). */
int orc_net_add_source_poisson(void *net, int n_src, const uint32_t *lcg_state,
                               float rate_hz, const int64_t *row_ptr,
                               const uint32_t *tgt_type, const float *w);
/* Counter-based binomial external drive into the excitatory conductance. */
void orc_net_set_external(void *net, int n_thr, const uint32_t *thresholds, float w_ext,
                          uint64_t seed);
/* Run n_steps updates starting at update index first_step (>= 1).  Writes
 * (step, index) of every spike, ascending within a step; returns the spike
 * count (or -1 if cap is exceeded). */
int64_t orc_net_run(void *net, int64_t first_step, int n_steps, int32_t *spike_step,
                    int32_t *spike_idx, int64_t cap);

/* Additive STDP (STDPWeightChange) on every synapse of a source (index in
 * the order sources were added): traces decay by ExponentialTable factors,
 * weights clamped to [0, max_weight].  Times in seconds. */
int orc_net_set_stdp(void *net, int source, float max_ltp, float tau_ltp, float max_ltd,
                     float tau_ltd, float max_weight);
/* Current weights and STDP state of a source's synapses (row order). */
int orc_net_get_synapses(void *net, int source, float *w, float *pre, float *post, double *last);

/* Record every delivered recurrent/device event (step, target, type, w) in
 * delivery order, for the ordering tests. */
void orc_net_set_trace(void *net, int on);
int64_t orc_net_get_trace(void *net, int32_t *step, int32_t *tgt, uint8_t *type, float *w,
                          int64_t cap);

/* ---- counter-based drive (shared definition with the HIP path) ----------- */
uint32_t orc_drive_hash(uint64_t seed, int64_t step, uint32_t neuron);

/* ---- several populations ------------------------------------------------ */
/* Projection from `source`'s population onto `net`'s population (CSR over the
 * source's neurons, same layout as the other sources); same max_delay. */
int orc_net_add_link(void *net, void *source, const int64_t *row_ptr, const uint32_t *tgt_type,
                     const float *w);
/* Spike train: update k (1..n_steps) emits idx[step_ptr[k-1] .. step_ptr[k]). */
int orc_net_add_source_train(void *net, int n_src, int64_t n_steps, const int64_t *step_ptr,
                             const int32_t *idx, const int64_t *row_ptr, const uint32_t *tgt_type,
                             const float *w);
/* Per-generator frequencies (Hz) of Poisson source `source`. */
int orc_net_set_poisson_rates(void *net, int source, const float *rates);
/* Lockstep run: update k of every network, in the order given, before update
 * k + 1 of any; spikes go to each network's log (orc_net_take_log). */
void orc_circuit_run(void **nets, int n_nets, int64_t first_step, int n_steps);
int64_t orc_net_take_log(void *net, int32_t *step, int32_t *idx, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
