"""Device-resident network mode: spike sources, synapses and delivery on the GPU.

A `Network` couples one GPU population (a TimeDrivenNeuronModel_GPU after
InitializeStates) with its spike sources: the population itself (recurrent
synapses), Poisson generator devices (PoissonGeneratorDeviceVector with one
RandomGenerator per source) and a counter-based external drive.  Spikes never
leave the device during a run; they are logged on the device and read back at
the end.  Delivery follows the order of the reference's event queue (see
include/edlut_amd.h, network mode).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _capi as C
from ._capi import EDLUTException


def build_out_csr(n_src: int, n_tgt: int, max_delay: int, src, tgt, delay, syn_type, weight):
    """Sort an edge list into the per-(source, delay) row layout, in the order of
    Network::FindOutConnections (Kernel/src/spike/Network.cpp:45-123).

    Returns (row_ptr int64[n_src*max_delay+1], tgt_type uint32[E], w float32[E]).
    """
    src = np.ascontiguousarray(src, np.int32)
    tgt = np.ascontiguousarray(tgt, np.int32)
    delay = np.ascontiguousarray(delay, np.int32)
    syn_type = np.ascontiguousarray(syn_type, np.uint8)
    weight = np.ascontiguousarray(weight, np.float32)
    e = len(src)
    if not (len(tgt) == len(delay) == len(syn_type) == len(weight) == e):
        raise EDLUTException(C.EDLUT_E_ARG, "edge arrays differ in length")
    row_ptr = np.zeros(n_src * max_delay + 1, np.int64)
    tt = np.zeros(max(e, 1), np.uint32)
    w = np.zeros(max(e, 1), np.float32)
    C.check(
        C.lib().edlut_build_out_csr(
            n_src, n_tgt, max_delay, e, C.ptr(src), C.ptr(tgt), C.ptr(delay), C.ptr(syn_type),
            C.ptr(weight), C.ptr(row_ptr), C.ptr(tt), C.ptr(w)
        )
    )
    return row_ptr, tt[:e], w[:e]


def binomial_thresholds(n_syn: int, p: float, max_terms: int = 16) -> np.ndarray:
    """Cumulative thresholds for the external drive: thr[j-1] = round(P(K >= j) * 2^32)
    for K ~ Binomial(n_syn, p), so that k = #{j : u < thr[j]} for a uniform
    32-bit u has that distribution (terms below 2^-32 are dropped)."""
    thr = []
    pmf = [math.comb(n_syn, k) * p**k * (1 - p) ** (n_syn - k) for k in range(n_syn + 1)]
    for j in range(1, min(n_syn, max_terms) + 1):
        tail = sum(pmf[j:])
        v = int(round(tail * 2.0**32))
        if v <= 0:
            break
        thr.append(min(v, 2**32 - 1))
    return np.array(thr, np.uint32)


class Network:
    def __init__(self, model, max_delay: int):
        self.model = model
        self.max_delay = int(max_delay)
        h = ctypes.c_void_p()
        C.check(C.lib().edlut_network_create(model.handle, self.max_delay, C.byref(h)))
        self._h = h
        self._keep = []
        self.next_step = 1

    def close(self):
        # the population owns its network; destroying the model frees it too
        if self._h is not None and self.model._pop is not None:
            C.lib().edlut_network_destroy(self._h)
        self._h = None

    def add_recurrent(self, row_ptr, tgt_type, weight):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        tgt_type = np.ascontiguousarray(tgt_type, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        C.check(C.lib().edlut_network_add_recurrent(self._h, C.ptr(row_ptr), C.ptr(tgt_type), C.ptr(weight)))

    def add_poisson(self, lcg_state, rate_hz, row_ptr, tgt_type, weight):
        lcg_state = np.ascontiguousarray(lcg_state, np.uint32)
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        tgt_type = np.ascontiguousarray(tgt_type, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        C.check(
            C.lib().edlut_network_add_poisson(
                self._h, len(lcg_state), C.ptr(lcg_state), float(rate_hz), C.ptr(row_ptr),
                C.ptr(tgt_type), C.ptr(weight)
            )
        )

    def set_source_model(self, source: int, key: int, rank: int = -1):
        """Event-queue order: sources of any network with the same key are one
        model (one layer type of the reference); `rank` orders the models'
        first update events (the network file's type order)."""
        C.check(C.lib().edlut_network_set_source_model(self._h, int(source), int(key), int(rank)))

    def set_model_rank(self, rank: int):
        """Rank of this population among the models (default: its position)."""
        C.check(C.lib().edlut_network_set_model_rank(self._h, int(rank)))

    def set_event_order(self, on: bool):
        """Deliver in the reference event queue's order (default) or in the
        canonical order (delay, source, spiking index, synapse)."""
        C.check(C.lib().edlut_network_set_event_order(self._h, int(bool(on))))

    def set_stop_time(self, t: float):
        """Stop time (s) of the next run (default: its last update's time)."""
        C.check(C.lib().edlut_network_set_stop_time(self._h, float(t)))

    def event_queue(self):
        """The replayed event queue's heap as stored: (times, keys), key =
        priority << 28 | kind << 24 | model << 16 | delay."""
        return _queue(C.lib().edlut_network_event_queue, self._h)

    def set_spike_train_times(self, source: int, times, order=None, load_first=None):
        """Times (s) of a spike train's input spikes, aligned with its
        schedule's entries, their load order (order[j] ranks entry j) and the
        first update of the run that queues each (0: the simulation's start)."""
        t = np.ascontiguousarray(times, np.float64)
        o = None if order is None else np.ascontiguousarray(order, np.int64)
        lf = None if load_first is None else np.ascontiguousarray(load_first, np.int64)
        C.check(C.lib().edlut_network_set_spike_train_times(self._h, int(source), len(t),
                                                            C.ptr(t) if len(t) else None,
                                                            C.ptr(o) if o is not None and len(o) else None,
                                                            C.ptr(lf) if lf is not None and len(lf) else None))

    def set_poisson_rates(self, source: int, rates_hz):
        """One frequency per generator of Poisson source `source`
        (PoissonGeneratorDeviceVector)."""
        r = np.ascontiguousarray(rates_hz, np.float32)
        C.check(C.lib().edlut_network_set_poisson_rates(self._h, int(source), C.ptr(r)))

    def add_spike_train(self, n_src: int, step_ptr, idx, row_ptr, tgt_type, weight):
        """Input neurons that fire at given updates: update k (1-based) emits
        idx[step_ptr[k-1]:step_ptr[k]] (ascending, no repeats).  The events
        of a spike at update s through a synapse of delay d are consumed by
        update s + d + 1, as for every other source."""
        step_ptr = np.ascontiguousarray(step_ptr, np.int64)
        idx = np.ascontiguousarray(idx, np.int32)
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        tgt_type = np.ascontiguousarray(tgt_type, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        n_steps = max(len(step_ptr) - 1, 0)
        C.check(
            C.lib().edlut_network_add_spike_train(
                self._h, int(n_src), n_steps, C.ptr(step_ptr) if n_steps else None,
                C.ptr(idx) if len(idx) else None, C.ptr(row_ptr), C.ptr(tgt_type), C.ptr(weight)
            )
        )

    def set_spike_train(self, source: int, step_ptr, idx):
        """Replaces the schedule of spike-train source `source`."""
        step_ptr = np.ascontiguousarray(step_ptr, np.int64)
        idx = np.ascontiguousarray(idx, np.int32)
        n_steps = max(len(step_ptr) - 1, 0)
        C.check(C.lib().edlut_network_set_spike_train(self._h, int(source), n_steps,
                                                      C.ptr(step_ptr) if n_steps else None,
                                                      C.ptr(idx) if len(idx) else None))

    def add_projection(self, source: "Network", row_ptr, tgt_type, weight):
        """Synapses from `source`'s population onto this network's population
        (rows over the source neurons).  Networks with projections run in a
        Circuit with their sources."""
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        tgt_type = np.ascontiguousarray(tgt_type, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        C.check(C.lib().edlut_network_add_projection(self._h, source._h, C.ptr(row_ptr), C.ptr(tgt_type),
                                                     C.ptr(weight)))
        self._keep.append(source)   # the source network outlives this one

    def generate_recurrent(self, in_degree: int, exc_fraction: float, w_mu: float, w_sigma: float,
                           seed: int):
        """Fixed in-degree random wiring built on the device (per group for a
        grouped population; see include/edlut_amd.h)."""
        C.check(C.lib().edlut_network_generate_recurrent(self._h, int(in_degree), float(exc_fraction),
                                                         float(w_mu), float(w_sigma), int(seed)))

    def synapses(self, source: int = 0):
        """(row_ptr, tgt_type, weight) of a source, copied from the device."""
        rows = ctypes.c_int64(0)
        nsyn = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_get_synapses(self._h, int(source), None, None, None,
                                                   C.byref(rows), C.byref(nsyn)))
        rp = np.zeros(rows.value + 1, np.int64)
        tt = np.zeros(max(nsyn.value, 1), np.uint32)
        w = np.zeros(max(nsyn.value, 1), np.float32)
        C.check(C.lib().edlut_network_get_synapses(self._h, int(source), C.ptr(rp), C.ptr(tt), C.ptr(w),
                                                   C.byref(rows), C.byref(nsyn)))
        return rp, tt[: nsyn.value], w[: nsyn.value]

    def set_stdp(self, source: int, max_ltp: float, tau_ltp: float, max_ltd: float, tau_ltd: float,
                 max_weight: float):
        """Additive STDP (STDPWeightChange parameters max_LTP, tau_LTP, max_LTD,
        tau_LTD, in seconds) on every synapse of a source."""
        C.check(C.lib().edlut_network_set_stdp(self._h, int(source), float(max_ltp), float(tau_ltp),
                                               float(max_ltd), float(tau_ltd), float(max_weight)))

    def stdp_state(self, source: int):
        """(pre, post, last) of a plastic source's synapses, in row order."""
        _, tt, _ = self.synapses(source)
        m = len(tt)
        pre = np.zeros(max(m, 1), np.float32)
        post = np.zeros(max(m, 1), np.float32)
        last = np.zeros(max(m, 1), np.float64)
        C.check(C.lib().edlut_network_get_stdp(self._h, int(source), C.ptr(pre), C.ptr(post), C.ptr(last)))
        return pre[:m], post[:m], last[:m]

    def reset(self):
        """Back to the initial state; the next run starts at update 1."""
        C.check(C.lib().edlut_network_reset(self._h))
        self.next_step = 1

    def set_launch(self, items: int, fold_lanes: int, fold_blocks: int):
        C.check(C.lib().edlut_network_set_launch(self._h, int(items), int(fold_lanes), int(fold_blocks)))

    def launch(self):
        """(items per thread, fold lanes, fold workgroups) of the current launch shape."""
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        C.check(C.lib().edlut_network_get_launch(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, trial_steps: int = 64):
        """Pick the fastest launch shape by timing trial runs (the network is
        reset afterwards).  Returns (items, lanes, blocks, ms_per_step)."""
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        ms = ctypes.c_double(0)
        C.check(C.lib().edlut_network_autotune(self._h, int(trial_steps), C.byref(a), C.byref(b),
                                               C.byref(c), C.byref(ms)))
        self.next_step = 1
        return a.value, b.value, c.value, ms.value

    def autotune_settled(self, settle_steps: int, trial_steps: int, items=(2, 4), lanes=(16, 32, 64),
                         blocks=(64, 256, 1024, 4096)):
        """Launch-shape search in the network's settled regime: `settle_steps`
        steps first, then every shape times `trial_steps` steps continuing the
        same run (no reset between trials, so each trial sees the activity the
        network has reached).  The network is reset afterwards.  Returns
        (items, lanes, blocks, ms_per_step)."""
        import time

        self.run(settle_steps)
        best = None
        for it in items:
            for ln in lanes:
                for bl in blocks:
                    try:
                        self.set_launch(it, ln, bl)
                    except C.EDLUTException:
                        continue
                    self.run(16)
                    t0 = time.perf_counter()
                    self.run(trial_steps)
                    ms = (time.perf_counter() - t0) * 1e3 / trial_steps
                    if best is None or ms < best[3]:
                        best = (it, ln, bl, ms)
        self.set_launch(*best[:3])
        self.reset()
        return best

    def set_external(self, thresholds, w_ext, seed):
        thr = np.ascontiguousarray(thresholds, np.uint32)
        C.check(C.lib().edlut_network_set_external(self._h, len(thr), C.ptr(thr), float(w_ext), int(seed)))

    def run(self, n_steps: int, first_step: int | None = None, sync: bool = True):
        if first_step is None:
            first_step = self.next_step
        C.check(C.lib().edlut_network_run(self._h, int(first_step), int(n_steps)))
        self.next_step = first_step + n_steps
        if sync:
            self.synchronize()

    def set_graph(self, on: bool):
        """Replay captured 16-step hipGraphs (default) or launch every kernel."""
        C.check(C.lib().edlut_network_set_graph(self._h, int(bool(on))))

    def synchronize(self):
        C.check(C.lib().edlut_network_synchronize(self._h))

    def spikes(self):
        """(step, index) arrays of every logged spike."""
        cnt = ctypes.c_int64(0)
        rc = C.lib().edlut_network_spike_log(self._h, None, None, 0, C.byref(cnt))
        if rc not in (0, C.EDLUT_E_CAP):
            C.check(rc)
        m = cnt.value
        st = np.zeros(max(m, 1), np.int32)
        ix = np.zeros(max(m, 1), np.int32)
        n = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_spike_log(self._h, C.ptr(st), C.ptr(ix), m, C.byref(n)))
        return st[: n.value].copy(), ix[: n.value].copy()

    def set_source_log(self, source: int, cap: int):
        """Keep a spike log of Poisson source `source` (cap entries; 0 drops it)."""
        C.check(C.lib().edlut_network_set_source_log(self._h, int(source), int(cap)))

    def source_spikes(self, source: int):
        """(step, generator index) arrays of Poisson source `source`'s logged spikes."""
        cnt = ctypes.c_int64(0)
        rc = C.lib().edlut_network_source_spike_log(self._h, int(source), None, None, 0, C.byref(cnt))
        if rc not in (0, C.EDLUT_E_CAP):
            C.check(rc)
        m = cnt.value
        st = np.zeros(max(m, 1), np.int32)
        ix = np.zeros(max(m, 1), np.int32)
        n = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_source_spike_log(self._h, int(source), C.ptr(st), C.ptr(ix), m,
                                                        C.byref(n)))
        return st[: n.value].copy(), ix[: n.value].copy()

    def set_log_capacity(self, cap: int):
        """Spike log capacity in entries; 0 keeps only the spike count."""
        C.check(C.lib().edlut_network_set_log_capacity(self._h, int(cap)))

    def spike_count(self) -> int:
        c = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_spike_count(self._h, C.byref(c)))
        return c.value

    def clear_log(self):
        C.check(C.lib().edlut_network_clear_log(self._h))

    def set_timing(self, max_steps: int):
        """Time the update kernels of the first max_steps steps of the next runs
        (per-workgroup device-clock stamps; graph replays included)."""
        C.check(C.lib().edlut_network_set_timing(self._h, int(max_steps)))

    def update_dispatch_timing(self):
        """(total ms, launches) of the timed update kernels from each launch's
        first workgroup start to the next compaction launch's (the duration a
        kernel trace reports for back-to-back dispatches)."""
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_update_dispatch_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def update_timing(self):
        """(total ms, launches) of the timed update kernels of the last run
        (device-clock span from each launch's first workgroup start to its last
        workgroup end)."""
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        C.check(C.lib().edlut_network_update_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def from_spec(spec, device: int = 0, params: dict | None = None):
    """Build the GPU population and its network for a workloads.NetworkSpec.

    Returns (model, network)."""
    from .models import CreateNeuronModel, integration_method

    model = CreateNeuronModel(spec.model, dict(spec.params))
    if params:
        model.SetParameters(params)
    model.SetParameters({"int_meth": integration_method(spec.integrator, float(np.float32(1e-4)))})
    model.InitializeStates(spec.n, 0, device)
    net = Network(model, spec.max_delay)
    if spec.rec is not None:
        s, t, d, y, w = spec.rec
        rp, tt, ww = build_out_csr(spec.n, spec.n, spec.max_delay, s, t, d, y, w)
        net.add_recurrent(rp, tt, ww)
    for states, rate, (s, t, d, y, w) in spec.poisson:
        rp, tt, ww = build_out_csr(len(states), spec.n, spec.max_delay, s, t, d, y, w)
        net.add_poisson(states, rate, rp, tt, ww)
    if spec.ext is not None:
        thr, w_ext, seed = spec.ext
        net.set_external(thr, w_ext, seed)
    return model, net


def cerebellum_from_spec(spec, device: int = 0):
    """GPU circuit of a workloads.CerebellumSpec: granule cells (BDF2) fed
    by the mossy-fibre Poisson devices, Golgi cells (LIF, Euler) and
    Purkinje cells (HH under BDF2 with the FP64 Newton step, or the
    Purkinje-cell model under RK4) fed by projections from the granule
    cells, additive STDP on granule -> Purkinje.

    Returns (models, networks, circuit) in the order granule, Golgi, Purkinje."""
    from .models import CreateNeuronModel, integration_method

    dt = float(np.float32(1e-4))
    D = spec.max_delay
    hh = spec.pc_model == "HHTimeDrivenModel_GPU"
    gr = CreateNeuronModel("EgidioGranuleCell_TimeDriven_GPU", {"int_meth": integration_method("BDF", dt, 2)})
    go = CreateNeuronModel("LIFTimeDrivenModel_GPU", {"int_meth": integration_method("Euler", dt)})
    pc = CreateNeuronModel(spec.pc_model, {"int_meth": integration_method("BDF" if hh else "RK4", dt, 2)})
    for m, n in ((gr, spec.n_grc), (go, spec.n_goc), (pc, spec.n_pc)):
        m.CheckSynapseType(0)
        m.InitializeStates(n, 0, device)
    if hh:
        pc.SetNewtonFP64(True)
    ng, no, npc = Network(gr, D), Network(go, D), Network(pc, D)
    ng.add_poisson(spec.mf_states, 0.0, *build_out_csr(spec.n_mf, spec.n_grc, D, *spec.mf_grc))
    ng.set_poisson_rates(0, spec.mf_rates)
    no.add_projection(ng, *build_out_csr(spec.n_grc, spec.n_goc, D, *spec.grc_goc))
    npc.add_projection(ng, *build_out_csr(spec.n_grc, spec.n_pc, D, *spec.grc_pc))
    npc.set_stdp(0, *spec.stdp)
    circuit = Circuit([ng, no, npc])
    return (gr, go, pc), (ng, no, npc), circuit


def _queue(fn, h):
    size = ctypes.c_int64(0)
    fn(h, None, None, 0, C.byref(size))   # the size (an error when the queue is not empty)
    n = size.value
    t = np.zeros(max(n, 1), np.float64)
    k = np.zeros(max(n, 1), np.uint32)
    C.check(fn(h, C.ptr(t), C.ptr(k), n, C.byref(size)))
    return t[:n], k[:n]


class Circuit:
    """Networks advanced in lockstep on one stream: update k of every member,
    in the order given, before update k + 1 of any (edlut_circuit_*)."""

    def __init__(self, nets):
        self.nets = list(nets)
        arr = (ctypes.c_void_p * len(self.nets))(*[n._h for n in self.nets])
        h = ctypes.c_void_p()
        C.check(C.lib().edlut_circuit_create(arr, len(self.nets), C.byref(h)))
        self._h = h
        self.next_step = 1

    def set_stop_time(self, t: float):
        """Stop time (s) of the next run (default: its last update's time)."""
        C.check(C.lib().edlut_circuit_set_stop_time(self._h, float(t)))

    def event_queue(self):
        """The circuit's replayed event queue (see Network.event_queue)."""
        return _queue(C.lib().edlut_circuit_event_queue, self._h)

    def run(self, n_steps: int, first_step: int | None = None, sync: bool = True):
        if first_step is None:
            first_step = self.next_step
        C.check(C.lib().edlut_circuit_run(self._h, int(first_step), int(n_steps)))
        self.next_step = first_step + n_steps
        for n in self.nets:
            n.next_step = self.next_step
        if sync:
            C.check(C.lib().edlut_circuit_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None:
            C.lib().edlut_circuit_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()
