/*
 * edlut_amd.h -- C ABI of the MI355X (gfx950) time-driven neuron-update path.
 *
 * The library replaces the device side of EDLUT's GPU neuron-model plug-ins
 * (TimeDrivenNeuronModel_GPU_C_INTERFACE and its subclasses) and their
 * integration-method plug-ins.  Plain pointers and sizes only; every call
 * returns 0 on success or a negative EDLUT_E* code, with a message available
 * from edlut_last_error() (the reference prints and exits instead,
 * Kernel/include/cudaError.h:14-22; the C++ host layer turns codes into
 * exceptions).
 *
 * Citations are relative to the reference repository root.
 */
#ifndef EDLUT_AMD_H_
#define EDLUT_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- identifiers -------------------------------------------------------- */
/* Neuron models; registry names as in Kernel/src/neuron_model/NeuronModelFactory.cu:66-91. */
enum edlut_model {
    EDLUT_LIF = 0,      /* "LIFTimeDrivenModel_GPU"  (LIFTimeDrivenModel_GPU_C_INTERFACE.cu)  */
    EDLUT_ADEX = 1,     /* "AdExTimeDrivenModel_GPU" (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:568-570) */
    EDLUT_HH = 2,       /* "HHTimeDrivenModel_GPU" */
    EDLUT_GRANULE = 3,  /* "EgidioGranuleCell_TimeDriven_GPU" (EgidioGranuleCell_TimeDriven_GPU_C_INTERFACE.cu) */
    EDLUT_PURKINJE = 4, /* "TimeDrivenPurkinjeCell_GPU" (TimeDrivenPurkinjeCell_GPU_C_INTERFACE.cu) */
    EDLUT_IZHIKEVICH = 5, /* "IzhikevichTimeDrivenModel_GPU" (IzhikevichTimeDrivenModel_GPU_C_INTERFACE.cu) */
    EDLUT_LIF_IS = 6     /* "LIFTimeDrivenModel_IS_GPU" (LIFTimeDrivenModel_IS_GPU_C_INTERFACE.cu), SI units */
};
/* Integration methods; names of IntegrationMethodFactory_GPU_C_INTERFACE.cu:40-145.
 * The Bifixed methods (Bifixed_{Euler,RK2,RK4}_GPU2.cuh) take n_steps steps of
 * the base method, of step/n_steps each, in every update. */
enum edlut_integrator {
    EDLUT_EULER = 0, EDLUT_RK2 = 1, EDLUT_RK4 = 2, EDLUT_BDF = 3,
    EDLUT_BIFIXED_EULER = 4, EDLUT_BIFIXED_RK2 = 5, EDLUT_BIFIXED_RK4 = 6
};
/* Synapse-type bits (CheckSynapseType, AdExTimeDrivenModel_GPU_C_INTERFACE.cu:239-290). */
enum edlut_synapse { EDLUT_SYN_EXC = 1, EDLUT_SYN_INH = 2, EDLUT_SYN_NMDA = 4, EDLUT_SYN_EXT_I = 8 };

enum edlut_error {
    EDLUT_OK = 0,
    EDLUT_E_ARG = -1,     /* invalid argument (EDLUTException in the reference) */
    EDLUT_E_HIP = -2,     /* HIP runtime error (HANDLE_ERROR in the reference)   */
    EDLUT_E_CAP = -3,     /* a device buffer's capacity was exceeded             */
    EDLUT_E_STATE = -4    /* call out of order (e.g. network already running)    */
};

typedef struct edlut_population edlut_population;
typedef struct edlut_network edlut_network;
typedef struct edlut_circuit edlut_circuit;

/* Message of the last failing call on this thread. */
const char *edlut_last_error(void);

/* ---- population (one neuron-model plug-in instance) --------------------- */
/*
 * Replaces M_GPU_C_INTERFACE construction + SetParameters + InitializeStates
 * (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:48-57,146-167) and
 * VectorNeuronState_GPU_C_INTERFACE::InitializeStatesGPU
 * (VectorNeuronState_GPU_C_INTERFACE.cu:66-131).  params holds the model's
 * float parameters in the order of its GetDefaultParameters() map:
 *   LIF : e_exc e_inh e_leak v_thr c_m tau_exc tau_inh tau_ref g_leak tau_nmda
 *   AdEx: a b thr_slo_fac v_thr tau_w e_exc e_inh e_reset e_leak g_leak c_m
 *         tau_exc tau_inh tau_nmda
 *   HH  : e_exc e_inh e_leak g_leak c_m v_thr tau_exc tau_inh tau_nmda g_na g_kd
 *         e_na e_k
 *   granule (EgidioGranuleCell_TimeDriven_GPU): none (n_params 0, params may
 *         be null); its constants are fixed (EgidioGranuleCell_TimeDriven.cpp:26-58)
 *   Purkinje (TimeDrivenPurkinjeCell_GPU): e_exc e_inh v_thr e_leak tau_exc
 *         tau_inh tau_nmda tau_ref (TimeDrivenPurkinjeCell.cpp:423-434)
 *   Izhikevich (IzhikevichTimeDrivenModel_GPU): a b c d e_exc e_inh c_m tau_exc
 *         tau_inh tau_nmda (IzhikevichTimeDrivenModel.cpp:389-401)
 *   LIF_IS (LIFTimeDrivenModel_IS_GPU): the LIF order in SI units (V, F, s, S)
 *         (LIFTimeDrivenModel_IS.cpp:392-404)
 * bdf_order is the integrator's integer parameter: the BDF method's order
 * (1..6, "bdf_order", BDFn_GPU_C_INTERFACE.cuh:275-288), or the Bifixed
 * methods' sub-steps per update (>= 1, "n_steps",
 * BifixedStep_GPU_C_INTERFACE.cuh:144-158); ignored by Euler, RK2 and RK4.
 * dt_s is the integration method's "step" parameter (seconds).  State lives on
 * `device` as structure-of-arrays: variable v of neuron i at v*stride+i.
 */
int edlut_population_create(int model, int integrator, int bdf_order, const float *params,
                            int n_params, float dt_s, int n_neurons, int device,
                            edlut_population **out);
int edlut_population_destroy(edlut_population *pop);

/*
 * Grouped population: n_groups independent groups (e.g. the networks of a
 * parameter sweep) of one model and integrator in one index space, so one
 * launch per kernel advances all of them.  params holds n_groups parameter
 * vectors (n_params each, order as above); groups may differ in every
 * parameter except e_exc/e_inh (shared g_nmda_inf table) and, for LIF,
 * tau_ref (shared refractory table).  Group g starts at start[g] (a multiple
 * of 256; edlut_population_groups) and owns sizes[g] neurons; the padding
 * neurons between groups never update or spike.  AdEx and LIF under Euler,
 * RK2 and RK4, with excitatory/inhibitory synapses and the external drive.
 * No counterpart in the reference, which builds one network per simulation.
 */
int edlut_population_create_groups(int model, int integrator, const float *params, int n_params,
                                   float dt_s, int n_groups, const int32_t *sizes, int device,
                                   edlut_population **out);
/* Group layout: count, and (if non-null) first neuron and size of each group.
 * A plain population is one group starting at 0. */
int edlut_population_groups(const edlut_population *pop, int *n_groups, int32_t *start,
                            int32_t *size);

/* Number of state variables and the SoA row stride (>= n_neurons). */
int edlut_population_layout(const edlut_population *pop, int *n_vars, int *n_diff,
                            int64_t *stride);

/* Enable synapse types (bit mask of edlut_synapse); replaces
 * SetEnabledSynapsis_GPU2 (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:232-262). */
int edlut_population_enable_synapses(edlut_population *pop, uint32_t mask);

/*
 * Queue input spikes for the next update, in the host's delivery order.
 * Replaces ProcessInputSpike (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:69-73):
 * before the next update, state[ndiff+type][target] += w is applied once per
 * event, in the order given, like the CPU model's direct increments
 * (AdExTimeDrivenModel.cpp:74-79).  type is 0 (exc), 1 (inh) or 2 (nmda).
 */
int edlut_population_push_inputs(edlut_population *pop, const int32_t *target,
                                 const uint8_t *type, const float *weight, int n);

/* Total external current (pA) of every neuron; replaces ProcessInputCurrent
 * (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:75-82).  Enables EXT_I. */
int edlut_population_set_ext_current(edlut_population *pop, const float *current);

/*
 * Advance every neuron by one step (update index `step`, counted from 1) and
 * return the number of spikes.  Replaces UpdateState
 * (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:90-112); synchronous like it.
 * Spiking neuron indices are compacted on the device in ascending order and
 * read with edlut_population_get_spikes (replaces the host bool[] scan of
 * TimeDrivenInternalSpike.cpp:132-160).
 */
int edlut_population_update(edlut_population *pop, int64_t step, int *n_spikes);
int edlut_population_get_spikes(const edlut_population *pop, int32_t *idx, int cap);

/*
 * Allocation-free update.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same step as
 * edlut_population_push_inputs + edlut_population_update, with the input
 * events already in device memory (target, type, weight arrays of n_events
 * entries, delivered in array order) and the spikes written to device memory
 * (spk_idx: room for every neuron, ascending; spk_cnt: one int32).
 * Asynchronous on `stream` (null: the population's stream); calls no
 * hipMalloc, hipFree or host synchronisation.  The input fold's event buffer
 * is a library-wide device array of EDLUT_NOALLOC_MAX_EVENTS events (16 MiB,
 * held by each GPU the library runs on until the process ends): n_events
 * above it is refused (EDLUT_E_CAP).  Every call on one GPU shares that
 * array, so calls that overlap on one GPU race; serialise them.  An event
 * whose target is out of range or whose synapse type is not enabled is
 * dropped and reported (EDLUT_E_ARG) by the population's next synchronous
 * call.  Refused while events queued by edlut_population_push_inputs wait.
 */
#define EDLUT_NOALLOC_MAX_EVENTS (1 << 20)
int edlut_population_update_noalloc(edlut_population *pop, int64_t step, const int32_t *target,
                                    const uint8_t *type, const float *weight, int n_events,
                                    int32_t *spk_idx, int32_t *spk_cnt, void *stream);

/* Copy one state variable to / from host memory (n_neurons floats); replaces
 * the monitored-state copies (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:99-103). */
int edlut_population_get_state(const edlut_population *pop, int var, float *out);
int edlut_population_set_state(edlut_population *pop, int var, const float *in);

/* HH under BDF: run the whole BDF step (predictor, Newton iterations,
 * Jacobian, Gauss-Jordan inverse) in FP64, state kept in float between
 * steps.  Must precede the first update.  No counterpart flag in the
 * reference, whose BDF is float throughout (BDFn.h:288-423). */
int edlut_population_set_newton_fp64(edlut_population *pop, int on);

/* Device pointer of the SoA state block (n_vars rows of `stride` floats). */
int edlut_population_device_state(const edlut_population *pop, float **state);

/* Checksums (FNV-1a over the bytes) of the look-up tables the kernels use:
 * [0] ExponentialTable, [1] g_nmda_inf table, [2] refractory table. */
int edlut_population_lut_checksums(const edlut_population *pop, uint64_t *out3);

/* ---- network mode: device-resident spike delivery ----------------------- */
/*
 * Spikes stay on the device: each source keeps a ring of compacted spike
 * lists, and before update k every event of a spike emitted at update
 * s = k-1-d through a synapse of delay d is folded into its target's
 * conductance, in the order the reference's event queue delivers them.  The
 * queue's operations are replayed on the device after every update
 * (eventq.hpp): one update-event stream per model at its own step, spike
 * groups of up to 1,024 sources per delay queued as they fill, input spikes
 * chaining one propagated spike per delay, a stop event per run, all in one
 * binary heap whose equal keys leave in its layout's order
 * (Kernel/src/simulation/EventQueue.cpp:121-198,
 * Kernel/src/spike/TimeDrivenInternalSpike.cpp:66-104,163-178,
 * Kernel/src/spike/PropagatedSpikeGroup.cpp:66-108,
 * Kernel/src/simulation/Simulation.cpp:322-373); within a group, source index
 * then target index (Kernel/src/spike/Network.cpp:45-68).  A population's
 * sources are one model each unless edlut_network_set_source_model joins
 * them; Poisson devices sample every 0.1 ms and need a 0.1 ms population
 * step.  Grouped populations run in the canonical order (ascending delay,
 * then source, then source index, then target index).
 *
 * Synapse arrays use a per-(source, delay) row pointer: synapses of source s
 * with delay d (1..max_delay steps) occupy [row_ptr[s*max_delay+d-1],
 * row_ptr[s*max_delay+d]), ascending target; tgt_type packs
 * target | type << 30.
 */
int edlut_network_create(edlut_population *pop, int max_delay, edlut_network **out);
int edlut_network_destroy(edlut_network *net);
/* The population itself as a spike source (recurrent synapses). */
int edlut_network_add_recurrent(edlut_network *net, const int64_t *row_ptr,
                                const uint32_t *tgt_type, const float *weight);
/* A Poisson generator device (PoissonGeneratorDeviceVector.cpp:56-70) of
 * n_src sources with their RandomGenerator states (RandomGenerator.h:61-107). */
int edlut_network_add_poisson(edlut_network *net, int n_src, const uint32_t *lcg_state,
                              float rate_hz, const int64_t *row_ptr, const uint32_t *tgt_type,
                              const float *weight);
/*
 * Per-generator frequencies (Hz) of Poisson source `source` (n_src values):
 * PoissonGeneratorDeviceVector keeps one frequency per generator
 * (PoissonGeneratorDeviceVector.cpp:56-70,95-99); a generator fires when
 * frequency * step > frand() and draws nothing when its frequency is not
 * positive.
 */
int edlut_network_set_poisson_rates(edlut_network *net, int source, const float *rates_hz);

/*
 * Spike-train source of n_src input neurons: the spikes that the reference's
 * InputSpikeNeuronModel layers receive from FileInputSpikeDriver
 * (FileInputSpikeDriver.cpp:45-85) or Simulation_API::AddExternalSpikeActivity
 * (Simulation_API.cpp), placed by the host on update steps.  Update k
 * (1..n_steps) emits idx[step_ptr[k-1] .. step_ptr[k]), ascending and without
 * repeats; its events reach the population like a Poisson device's, one
 * update of lookahead plus the synapse's delay later.  Later updates emit
 * nothing.
 */
int edlut_network_add_spike_train(edlut_network *net, int n_src, int64_t n_steps,
                                  const int64_t *step_ptr, const int32_t *idx,
                                  const int64_t *row_ptr, const uint32_t *tgt_type,
                                  const float *weight);
/* Replaces the schedule of spike-train source `source` (same layout, steps
 * counted from update 1; updates already run are not replayed).  The
 * reference inserts external spikes into its event queue between runs
 * (Simulation_API::AddExternalSpikeActivity). */
int edlut_network_set_spike_train(edlut_network *net, int source, int64_t n_steps,
                                  const int64_t *step_ptr, const int32_t *idx);

/*
 * Projection of the population of network `from` onto the population of
 * `net` (another neuron model, e.g. granule -> Purkinje): CSR rows over the
 * source population's neurons, the same layout as edlut_network_add_recurrent.
 * Replaces the reference's inter-population Interconnections delivered by
 * PropagatedSpikeGroup (PropagatedSpikeGroup.cpp:66-108) for GPU->GPU
 * synapses: the spikes come from `from`'s device ring.  Both networks must
 * share the device, the step and max_delay, and run in one circuit.  Events
 * of one update reach the target by delay, then source (in the order the
 * sources were added: recurrent, Poisson devices, projections), then
 * spiking source index, then synapse.
 */
int edlut_network_add_projection(edlut_network *net, edlut_network *from, const int64_t *row_ptr,
                                 const uint32_t *tgt_type, const float *w);

/* Event-queue order.  Sources (Poisson devices, spike trains) of any network
 * with the same key are one model, as the layers of one neuron type are one
 * model in the reference (Network.cpp:1016-1110); rank orders the models'
 * first update events as the network file orders its types
 * (Simulation::InitSimulation, Simulation.cpp:228-268).  Defaults: every
 * source its own model, populations first in circuit order. */
int edlut_network_set_source_model(edlut_network *net, int source, int key, int rank);
int edlut_network_set_model_rank(edlut_network *net, int rank);
/* on = 0: canonical order instead of the event queue's (no counterpart in
 * the reference; for comparisons). */
int edlut_network_set_event_order(edlut_network *net, int on);
/* Stop time (s) of the next run, between its last update's time and the
 * next update's (Simulation_API::RunSimulation's end time, whose stop event
 * shares the queue; Simulation.cpp:322-373).  Default: the last update's
 * time; -2: no stop event (the run and the next one are one reference run). */
int edlut_network_set_stop_time(edlut_network *net, double t);
/* The replayed event queue of a network that runs alone (of a circuit:
 * edlut_circuit_event_queue) as its heap stores it: event times and keys
 * (priority << 28 | kind << 24 | model << 16 | delay; kind 0 update event,
 * 1 spike group, 2 input spike, 3 propagated spike (delay: the index of the
 * neuron's delay), 4 stop), *size events (an error when cap is smaller).
 * No counterpart in the reference, whose queue is EventQueue's heap; for
 * checking the replay against it. */
int edlut_network_event_queue(edlut_network *net, double *t, uint32_t *key, int64_t cap, int64_t *size);
int edlut_circuit_event_queue(edlut_circuit *c, double *t, uint32_t *key, int64_t cap, int64_t *size);
/* Times (s) of a spike train's input spikes, aligned with the entries of its
 * schedule, their load order (order[j] ranks entry j; null: schedule order)
 * and when they are queued (load_first[j] = 0: when the simulation starts,
 * before the models' first update events, as FileInputSpikeDriver's spikes
 * (Simulation.cpp:228-268); k: by the run starting at update k, as
 * Simulation_API::AddExternalSpikeActivity's; null: all 0): the InputSpike
 * events of the reference (FileInputSpikeDriver.cpp:45-85, InputSpike.cpp).
 * Without it an input spike sits at its update's time. */
int edlut_network_set_spike_train_times(edlut_network *net, int source, int64_t n_spikes,
                                        const double *time, const int64_t *order, const int64_t *load_first);

/*
 * Circuit: networks advanced in lockstep on one stream, update k of every
 * member (in the order given) before update k + 1 of any; every projection
 * must come from a member.  While a network belongs to a circuit it is run
 * only through the circuit (its state, spike log and synapses stay readable).
 * The reference advances its populations through one event queue
 * (TimeEventUpdateNeurons.cpp:45-79, one update event per model).
 */
int edlut_circuit_create(edlut_network *const *nets, int n_nets, edlut_circuit **out);
int edlut_circuit_run(edlut_circuit *c, int64_t first_step, int n_steps);
/* Stop time (s) of the circuit's next run (see edlut_network_set_stop_time). */
int edlut_circuit_set_stop_time(edlut_circuit *c, double t);
int edlut_circuit_synchronize(edlut_circuit *c);
int edlut_circuit_destroy(edlut_circuit *c);

/* Counter-based external drive: per neuron and step, k events of weight w_ext
 * into the excitatory conductance, k = #{j : hash(seed,step,i) < thr[j]},
 * folded after the interval's recurrent events.  In a grouped population
 * neuron i of group g uses seed + g and its index within the group, so each
 * group is driven as the same network built on its own would be. */
int edlut_network_set_external(edlut_network *net, int n_thr, const uint32_t *thresholds,
                               float w_ext, uint64_t seed);
/* Enqueue n_steps updates starting at update index first_step (asynchronous;
 * spikes of the population are appended to the device spike log).  Whole
 * chunks of 16 steps are replayed from a captured hipGraph whose kernels
 * read the update index from device memory; the remainder (and every step
 * of a timed run) is launched directly.  Replaces the per-step TimeEventUpdateNeurons / UpdateState /
 * TimeDrivenInternalSpike / PropagatedSpikeGroup cycle for device-resident
 * networks (TimeEventUpdateNeurons.cpp:45-79). */
int edlut_network_run(edlut_network *net, int64_t first_step, int n_steps);
/* Graph replay on (default) or off (every step launched directly). */
int edlut_network_set_graph(edlut_network *net, int on);
int edlut_network_synchronize(edlut_network *net);
/* Spike log (step, index) pairs, ascending within a step; returns the count. */
int edlut_network_spike_log(edlut_network *net, int32_t *step, int32_t *idx, int64_t cap,
                            int64_t *n_out);
/* Empties the population's spike log and the logs of its Poisson sources
 * (the output drivers' GetBufferedSpikes empties its buffer on each read,
 * Simulation_API.cpp:173-204); runs continue unaffected. */
int edlut_network_clear_log(edlut_network *net);
/* Spike log of Poisson device `source` (cap entries, 0 to drop it): the
 * generators that fired and when, for output of a PoissonGeneratorDeviceVector
 * layer (the reference hands its spikes to the output drivers like any
 * neuron's, TimeDrivenInternalSpike.cpp:66-104).  Read with
 * edlut_network_source_spike_log (step, index), E_CAP when it overflowed. */
int edlut_network_set_source_log(edlut_network *net, int source, int64_t cap);
int edlut_network_source_spike_log(edlut_network *net, int source, int32_t *step, int32_t *idx,
                                   int64_t cap, int64_t *n_out);
/* Spike log capacity in entries (default 2^26); 0 keeps only the count.
 * Clears the log. */
int edlut_network_set_log_capacity(edlut_network *net, int64_t cap);
/* Spikes logged (or counted) since the log was last cleared. */
int edlut_network_spike_count(edlut_network *net, int64_t *count);
/* Time the update kernels of the next runs' first max_steps steps (0
 * switches timing off): every workgroup of a timed update kernel stores its
 * start and end on the device wall clock (graph replays included); a
 * launch's duration is the span from its first start to its last end. */
int edlut_network_set_timing(edlut_network *net, int64_t max_steps);
/* Total milliseconds of the timed update-kernel launches of the last run and
 * their count. */
int edlut_network_update_timing(edlut_network *net, double *ms, int64_t *launches);
/* The timed update kernels' launch durations as a kernel trace measures them
 * (the dispatch and the end-of-kernel cache write-back included): summed
 * from each launch's first workgroup start to the first workgroup start of
 * the compaction launch that follows it on the stream.  Plain networks. */
int edlut_network_update_dispatch_timing(edlut_network *net, double *ms, int64_t *launches);

/*
 * Synthetic recurrent wiring generated on the device (replaces building and
 * uploading a host edge list for large synthetic networks; no counterpart in
 * the reference, whose networks come from files, Network.cpp:445-640).
 * Every neuron receives in_degree synapses from sources drawn uniformly
 * within its own group, delays uniform in [1, max_delay], weights
 * lognormal(w_mu, w_sigma); the first round(exc_fraction * size) neurons of a
 * group are excitatory sources, the rest inhibitory.  Group g is generated
 * from seed + g and local indices, so it equals the wiring of a plain
 * population of its size generated with that seed.
 */
int edlut_network_generate_recurrent(edlut_network *net, int in_degree, float exc_fraction,
                                     float w_mu, float w_sigma, uint64_t seed);
/* Copy a source's synapse arrays to the host (row layout above).  Sizes are
 * returned in n_rows (row_ptr holds n_rows+1 entries) and n_syn; null arrays
 * are skipped, so a first call with nulls sizes the buffers. */
int edlut_network_get_synapses(edlut_network *net, int source, int64_t *row_ptr,
                               uint32_t *tgt_type, float *weight, int64_t *n_rows,
                               int64_t *n_syn);
/*
 * Additive STDP on every synapse of a source (source index = order of the
 * add/generate calls), replacing STDPWeightChange / STDPState
 * (STDPWeightChange.cpp:43-76, STDPState.cpp:74-97) for device-resident
 * synapses.  A delivered spike first adds the synapse's current weight to
 * the target, then both traces decay by ExponentialTable factors to its
 * arrival time, pre += 1 and w -= max_ltd * post; a target's spike at
 * update k decays the traces of its plastic inputs to t_k, post += 1 and
 * w += max_ltp * pre.  Weights stay in [0, max_weight]
 * (Interconnection.h:400-409).  Times in seconds: update k at
 * t_k = t_{k-1} + double(float dt), arrival = t_s + the delay rounded by
 * Network::RoundPropagationDelay.  Weights are read back with
 * edlut_network_get_synapses.
 */
int edlut_network_set_stdp(edlut_network *net, int source, float max_ltp, float tau_ltp,
                           float max_ltd, float tau_ltd, float max_weight);
/* Traces and last update times of a plastic source's synapses (row order). */
int edlut_network_get_stdp(edlut_network *net, int source, float *pre, float *post, double *last);
/* Back to the initial state: neuron state as after construction (the
 * external-current row, an input, is kept), Poisson
 * generators at their initial states, plastic weights and traces as when
 * STDP was switched on, empty spike rings and logs (its Poisson sources'
 * logs included); the next run starts at update 1. */
int edlut_network_reset(edlut_network *net);
/* Launch shape: neurons per thread of the update kernel (4; 2 or 4 for a
 * grouped population), lanes per spike and workgroups of the ring fold
 * (lanes 8, 16, 32 or 64).  Results do not depend on the launch shape. */
int edlut_network_set_launch(edlut_network *net, int items, int fold_lanes, int fold_blocks);
int edlut_network_get_launch(const edlut_network *net, int *items, int *fold_lanes,
                             int *fold_blocks);
/* Autotuner: times trial_steps steps of every launch shape on this network
 * and keeps the fastest (returned, with its time per step); resets the
 * network before each trial and after the search. */
int edlut_network_autotune(edlut_network *net, int trial_steps, int *items, int *fold_lanes,
                           int *fold_blocks, double *ms_per_step);

/* ---- trace fitting: forward simulation and discrete adjoint ------------- */
/*
 * No counterpart in the reference (it has no gradient path).  n independent
 * neurons of one model, LIF under Euler or AdEx under RK4, each with n_syn
 * input synapses driven by fixed spike trains, run `steps` steps of
 * dt_s.  params holds the model's parameter vector (order of
 * edlut_population_create); g_leak, e_leak, tau_exc and tau_inh are taken
 * per neuron from the tensors instead.  The forward pass writes the V trace
 * and keeps the state every `checkpoint` steps; the backward pass takes
 * dL/dV for every step and writes dL/d(g_leak, e_leak, tau_exc, tau_inh)
 * per neuron and dL/dweight per synapse.  It is the reverse-mode derivative
 * of the discrete forward map, with the forward run's spike pattern held
 * fixed; it replays each checkpointed segment before sweeping it backwards.
 * The AdEx exponential term uses expf (not the look-up table, whose
 * derivative vanishes) and decay factors are exp(-h/tau) of each neuron.
 */
typedef struct edlut_fit edlut_fit;
typedef struct edlut_fit_tensors {
    const float *g_leak, *e_leak, *tau_exc, *tau_inh; /* [n] device */
    const float *weight;                              /* [n][n_syn] device */
    const uint8_t *syn_type;                          /* [n][n_syn]: 0 exc, 1 inh */
    const int32_t *ev_ptr;                            /* [n+1]: events of neuron i */
    const int32_t *ev_step;                           /* update index (1..steps) the event precedes */
    const int32_t *ev_syn;                            /* synapse of the event (0..n_syn-1) */
    float *v_trace;                                   /* [steps][n]: V after each update */
    uint8_t *spikes;                                  /* [steps][n]: 1 where a spike reset V (optional) */
    float *traj;                                      /* [steps][4][n]: V, w, g_exc, g_inh before each
                                                         step's integration (optional) */
    int32_t *traj_since;                              /* [steps][n]: steps since the last spike (with traj) */
    const float *grad_v;                              /* [steps][n]: dL/dV (backward) */
    float *grad_par;                                  /* [4][n] (backward output) */
    float *grad_w;                                    /* [n][n_syn] (backward output) */
} edlut_fit_tensors;
int edlut_fit_create(int model, const float *params, int n_params, float dt_s, int n_neurons,
                     int steps, int n_syn, int checkpoint, int device, edlut_fit **out);
int edlut_fit_destroy(edlut_fit *fit);
/* Asynchronous on `stream` (a hipStream_t; null = the default stream). */
int edlut_fit_forward(edlut_fit *fit, const edlut_fit_tensors *t, void *stream);
/* Needs the checkpoints of a forward pass with the same tensors. */
int edlut_fit_backward(edlut_fit *fit, const edlut_fit_tensors *t, void *stream);
/* LIF refractory fraction by steps since the last spike (host helper; the
 * double clock of LIFTimeDrivenModel.cpp:189-201). */
int edlut_refractory_table(float dt_s, float tau_ref, float *out, int cap, int *len);

/*
 * Host helper: sort an edge list into the row layout above, in the order of
 * Network::FindOutConnections / qsort_inters (Kernel/src/spike/Network.cpp:
 * 45-123: source, then delay, then target).  delay is in steps (1..max_delay);
 * row_ptr has n_src*max_delay+1 entries; tgt_type and w_out have n_edges.
 */
int edlut_build_out_csr(int64_t n_src, int64_t n_tgt, int max_delay, int64_t n_edges,
                        const int32_t *src, const int32_t *tgt, const int32_t *delay,
                        const uint8_t *type, const float *w, int64_t *row_ptr,
                        uint32_t *tgt_type, float *w_out);

/* ---- misc --------------------------------------------------------------- */
/* Name of the compiled device architecture, e.g. "gfx950". */
const char *edlut_device_arch(void);
/* Counter-based hash used by the external drive (host copy, same definition). */
uint32_t edlut_drive_hash(uint64_t seed, int64_t step, uint32_t neuron);

#ifdef __cplusplus
}
#endif
#endif /* EDLUT_AMD_H_ */
