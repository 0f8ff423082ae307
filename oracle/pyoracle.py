"""ctypes interface of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke entry point and the CPU
baseline of bench.py to check and time the HIP implementation; never by the
product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DUMP = os.path.join(HERE, "_ref", "ref_dump")
REF_EVENTQUEUE = os.path.join(HERE, "_ref", "ref_eventqueue")

MODELS = {"LIFTimeDrivenModel_GPU": 0, "AdExTimeDrivenModel_GPU": 1, "HHTimeDrivenModel_GPU": 2,
          "EgidioGranuleCell_TimeDriven_GPU": 3, "TimeDrivenPurkinjeCell_GPU": 4,
          "IzhikevichTimeDrivenModel_GPU": 5, "LIFTimeDrivenModel_IS_GPU": 6}
INTEGRATORS = {"Euler": 0, "RK2": 1, "RK4": 2, "BDF": 3, "Bifixed_Euler": 4, "Bifixed_RK2": 5, "Bifixed_RK4": 6}

P = ctypes.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def build_ref(reference_kernel: str) -> bool:
    """Compile oracle/_ref from the reference's own sources (needs the checkout)."""
    if not os.path.isdir(reference_kernel):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={reference_kernel}"], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = ctypes.CDLL(LIB)
        l.orc_pop_create.restype = P
        l.orc_pop_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
        l.orc_pop_destroy.argtypes = [P]
        l.orc_pop_nvars.argtypes = [P]
        l.orc_pop_get_state.argtypes = [P, ctypes.c_int, P]
        l.orc_pop_set_state.argtypes = [P, ctypes.c_int, P]
        l.orc_pop_add_inputs.argtypes = [P, P, P, P, ctypes.c_int]
        l.orc_pop_set_ext_current.argtypes = [P, P]
        l.orc_pop_update.argtypes = [P, ctypes.c_int64, P]
        l.orc_set_threads.argtypes = [ctypes.c_int]
        l.orc_set_glibc_logf.argtypes = [ctypes.c_int]
        l.orc_set_order.argtypes = [ctypes.c_int]
        l.orc_net_set_source_model.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        l.orc_net_set_train_times.argtypes = [P, ctypes.c_int, ctypes.c_int64, P, P, P]
        l.orc_net_set_train.argtypes = [P, ctypes.c_int, ctypes.c_int64, P, P]
        l.orc_net_set_stop_time.argtypes = [P, ctypes.c_double]
        l.orc_net_set_model_rank.argtypes = [P, ctypes.c_int]
        l.orc_net_event_queue.restype = ctypes.c_int64
        l.orc_net_event_queue.argtypes = [P, P, P, P, P, P, ctypes.c_int64]
        l.orc_net_take_poisson_log.restype = ctypes.c_int64
        l.orc_net_take_poisson_log.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int64]
        l.orc_heap_replay.restype = ctypes.c_int64
        l.orc_heap_replay.argtypes = [ctypes.c_int64, P, P, P, P]
        l.orc_exp_table.argtypes = [P]
        l.orc_exp_lut.argtypes = [P, P, ctypes.c_int]
        l.orc_nmda_table.argtypes = [ctypes.c_float, ctypes.c_float, P]
        l.orc_hh_table.argtypes = [ctypes.c_float, P]
        l.orc_granule_table.argtypes = [P]
        l.orc_purkinje_table.argtypes = [P]
        l.orc_log_cr.argtypes = [P, P, ctypes.c_int]
        l.orc_lcg_seeds.argtypes = [ctypes.c_uint32, ctypes.c_int, P]
        l.orc_lcg_frand.argtypes = [ctypes.c_uint32, ctypes.c_int, P]
        l.orc_drive_hash.restype = ctypes.c_uint32
        l.orc_drive_hash.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_uint32]
        l.orc_net_create.restype = P
        l.orc_net_create.argtypes = [P, ctypes.c_int]
        l.orc_net_destroy.argtypes = [P]
        l.orc_net_add_source_self.argtypes = [P, P, P, P]
        l.orc_net_add_source_poisson.argtypes = [P, ctypes.c_int, P, ctypes.c_float, P, P, P]
        l.orc_net_set_external.argtypes = [P, ctypes.c_int, P, ctypes.c_float, ctypes.c_uint64]
        l.orc_net_set_trace.argtypes = [P, ctypes.c_int]
        fit = [ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
               P, P, P, P, P, P]
        l.orc_fit_forward.argtypes = fit + [ctypes.c_int, P, P]
        l.orc_fit_backward.argtypes = fit + [P, P, P]
        l.orc_fit_trajectory.argtypes = fit + [P, P, P]
        l.orc_fit_backward_along.argtypes = fit + [P, P, P, P, P, P]
        l.orc_net_set_stdp.argtypes = [P, ctypes.c_int] + [ctypes.c_float] * 5
        l.orc_net_get_synapses.argtypes = [P, ctypes.c_int, P, P, P, P]
        l.orc_net_get_trace.restype = ctypes.c_int64
        l.orc_net_get_trace.argtypes = [P, P, P, P, P, ctypes.c_int64]
        l.orc_net_add_link.argtypes = [P, P, P, P, P]
        l.orc_net_set_poisson_rates.argtypes = [P, ctypes.c_int, P]
        l.orc_net_add_source_train.argtypes = [P, ctypes.c_int, ctypes.c_int64, P, P, P, P, P]
        l.orc_circuit_run.argtypes = [P, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
        l.orc_net_take_log.restype = ctypes.c_int64
        l.orc_net_take_log.argtypes = [P, P, P, ctypes.c_int64]
        l.orc_net_run.restype = ctypes.c_int64
        l.orc_net_run.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, P, ctypes.c_int64]
        _lib = l
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def set_glibc_logf(on: bool) -> None:
    """The granule's Nernst logarithm through the C library's logf (as the
    reference) instead of the correctly rounded one (default)."""
    lib().orc_set_glibc_logf(int(bool(on)))


def set_threads(n: int) -> None:
    """OpenMP threads of the population updates (1: serial)."""
    lib().orc_set_threads(int(n))


def set_order(mode: int) -> None:
    """Delivery order of network runs: 1 the reference's event queue
    (default), 0 the canonical order (delay, source, spiking index, synapse)."""
    lib().orc_set_order(int(mode))


def heap_replay(op, t, prio) -> np.ndarray:
    """Ids in removal order of a sequence of heap operations (op 1 inserts
    (t, prio) with id = number of earlier inserts, op 0 removes the top)."""
    op = np.ascontiguousarray(op, np.int8)
    t = np.ascontiguousarray(t, np.float64)
    prio = np.ascontiguousarray(prio, np.int32)
    out = np.zeros(max(int((op == 0).sum()), 1), np.int32)
    r = lib().orc_heap_replay(len(op), _p(op), _p(t), _p(prio), _p(out))
    return out[:r].copy()


def exp_table() -> np.ndarray:
    t = np.zeros(16384, np.float32)
    lib().orc_exp_table(_p(t))
    return t


def exp_lut(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_exp_lut(_p(x), _p(out), len(x))
    return out


def nmda_table(e_exc: float, e_inh: float) -> np.ndarray:
    t = np.zeros(256, np.float32)
    lib().orc_nmda_table(e_exc, e_inh, _p(t))
    return t


def hh_table(v_thr: float) -> np.ndarray:
    t = np.zeros((262144, 6), np.float32)
    lib().orc_hh_table(v_thr, _p(t))
    return t


def granule_table() -> np.ndarray:
    t = np.zeros((1024, 26), np.float32)
    lib().orc_granule_table(_p(t))
    return t


def purkinje_table() -> np.ndarray:
    t = np.zeros((1024, 4), np.float32)
    lib().orc_purkinje_table(_p(t))
    return t


def log_cr(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_log_cr(_p(x), _p(out), len(x))
    return out


def lcg_seeds(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_lcg_seeds(seed, n, _p(out))
    return out


def lcg_frand(x0: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().orc_lcg_frand(x0, n, _p(out))
    return out


def drive_hash(seed: int, step: int, neuron: int) -> int:
    return int(lib().orc_drive_hash(seed, step, neuron))


class Population:
    """One population of the oracle (float restatement, or FP64 with fp64=True)."""

    def __init__(self, model: str, integrator: str, params, n: int, dt_s=np.float32(1e-4),
                 mask: int = 3, fp64=False, bdf_order: int = 1, n_steps: int = 2):
        # Bifixed methods take n_steps (sub-steps per update) where BDF takes its order
        p = np.ascontiguousarray(params, np.float32)
        self.n = n
        iparam = n_steps if integrator.startswith("Bifixed") else bdf_order
        self.h = lib().orc_pop_create(MODELS[model], INTEGRATORS[integrator], iparam, _p(p), len(p),
                                      float(np.float32(dt_s)), n, mask, int(fp64))
        if not self.h:
            raise ValueError("oracle does not implement this model")
        self.nv = lib().orc_pop_nvars(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pop_destroy(self.h)
            self.h = None

    def state(self, var: int) -> np.ndarray:
        out = np.zeros(self.n, np.float64)
        lib().orc_pop_get_state(self.h, var, _p(out))
        return out

    def set_state(self, var: int, values) -> None:
        v = np.ascontiguousarray(values, np.float64)
        lib().orc_pop_set_state(self.h, var, _p(v))

    def add_inputs(self, tgt, typ, w) -> None:
        t = np.ascontiguousarray(tgt, np.int32)
        y = np.ascontiguousarray(typ, np.uint8)
        ww = np.ascontiguousarray(w, np.float32)
        lib().orc_pop_add_inputs(self.h, _p(t), _p(y), _p(ww), len(t))

    def set_ext_current(self, cur) -> None:
        c = np.ascontiguousarray(cur, np.float32)
        lib().orc_pop_set_ext_current(self.h, _p(c))

    def update(self, step: int) -> np.ndarray:
        out = np.zeros(self.n, np.int32)
        k = lib().orc_pop_update(self.h, step, _p(out))
        return out[:k].copy()


class Net:
    def __init__(self, pop: Population, max_delay: int):
        self.pop = pop
        self.h = lib().orc_net_create(pop.h, max_delay)
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_net_destroy(self.h)
            self.h = None

    def add_self(self, row_ptr, tgt_type, w):
        a = [np.ascontiguousarray(row_ptr, np.int64), np.ascontiguousarray(tgt_type, np.uint32),
             np.ascontiguousarray(w, np.float32)]
        lib().orc_net_add_source_self(self.h, *[_p(x) for x in a])

    def add_poisson(self, states, rate, row_ptr, tgt_type, w):
        s = np.ascontiguousarray(states, np.uint32)
        a = [np.ascontiguousarray(row_ptr, np.int64), np.ascontiguousarray(tgt_type, np.uint32),
             np.ascontiguousarray(w, np.float32)]
        lib().orc_net_add_source_poisson(self.h, len(s), _p(s), float(rate), *[_p(x) for x in a])

    def add_train(self, n_src, step_ptr, idx, row_ptr, tgt_type, w):
        """Input neurons firing at given updates: update k emits
        idx[step_ptr[k-1]:step_ptr[k]]."""
        sp = np.ascontiguousarray(step_ptr, np.int64)
        ix = np.ascontiguousarray(idx, np.int32)
        if len(sp) == 0:
            sp = np.zeros(1, np.int64)
        a = [np.ascontiguousarray(row_ptr, np.int64), np.ascontiguousarray(tgt_type, np.uint32),
             np.ascontiguousarray(w, np.float32)]
        lib().orc_net_add_source_train(self.h, int(n_src), len(sp) - 1, _p(sp), _p(ix) if len(ix) else None,
                                       *[_p(x) for x in a])

    def set_source_model(self, source: int, key: int, rank: int = -1):
        """Sources with the same key, across networks, are one model (one
        layer type of the reference); rank orders the models' first updates."""
        if lib().orc_net_set_source_model(self.h, int(source), int(key), int(rank)) != 0:
            raise ValueError("bad source or simulation already started")

    def set_train_times(self, source: int, times=None, order=None, load_first=None):
        """Times (s), load ranks and first update of the loading run (0: when
        the simulation starts) of a spike train's input spikes, aligned with
        its schedule's entries (None: the update's time, schedule order, 0)."""
        arrs = [None if x is None else np.ascontiguousarray(x, dt) for x, dt in
                ((times, np.float64), (order, np.int64), (load_first, np.int64))]
        m = next((len(a) for a in arrs if a is not None), 0)
        self._keep += [a for a in arrs if a is not None]
        if lib().orc_net_set_train_times(self.h, int(source), m, *[_p(a) if a is not None else None for a in arrs]) != 0:
            raise ValueError("not a spike-train source, or the times do not match its schedule")

    def set_train(self, source: int, step_ptr, idx):
        """Replaces a spike train's schedule (times, ranks and load runs reset)."""
        sp = np.ascontiguousarray(step_ptr, np.int64)
        ix = np.ascontiguousarray(idx, np.int32)
        self._keep += [sp, ix]
        if lib().orc_net_set_train(self.h, int(source), len(sp) - 1, _p(sp), _p(ix) if len(ix) else None) != 0:
            raise ValueError("not a spike-train source")

    def set_model_rank(self, rank: int):
        """Rank of this network's population among the models (update-event
        order at start; default: its position in the circuit)."""
        if lib().orc_net_set_model_rank(self.h, int(rank)) != 0:
            raise ValueError("simulation already started")

    def event_queue(self):
        """The simulation's event heap as stored: (time, priority, kind,
        model, delay) arrays (kind 0 update, 1 group, 2 input, 3 propagated
        spike, 4 stop)."""
        m = lib().orc_net_event_queue(self.h, None, None, None, None, None, 0)
        out = [np.zeros(m, np.float64)] + [np.zeros(m, np.int32) for _ in range(4)]
        lib().orc_net_event_queue(self.h, *[_p(x) for x in out], m)
        return tuple(out)

    def set_stop_time(self, t: float):
        """Stop time (s) of the next run (default: its last update's time;
        -2: no stop event, the next run continues this one)."""
        lib().orc_net_set_stop_time(self.h, float(t))

    def take_poisson_log(self, source: int):
        """(update, generator) of every spike of a Poisson source since the last call."""
        m = lib().orc_net_take_poisson_log(self.h, int(source), None, None, 0)
        st = np.zeros(max(m, 1), np.int32)
        ix = np.zeros(max(m, 1), np.int32)
        lib().orc_net_take_poisson_log(self.h, int(source), _p(st), _p(ix), m)
        return st[:m].copy(), ix[:m].copy()

    def set_external(self, thr, w_ext, seed):
        t = np.ascontiguousarray(thr, np.uint32)
        lib().orc_net_set_external(self.h, len(t), _p(t), float(w_ext), int(seed))

    def set_stdp(self, source: int, max_ltp, tau_ltp, max_ltd, tau_ltd, max_weight):
        if lib().orc_net_set_stdp(self.h, source, max_ltp, tau_ltp, max_ltd, tau_ltd, max_weight) != 0:
            raise ValueError("bad STDP arguments")

    def synapses(self, source: int):
        """(w, pre, post, last) of every synapse of a source, in row order."""
        m = lib().orc_net_get_synapses(self.h, source, None, None, None, None)
        w = np.zeros(m, np.float32)
        pre = np.zeros(m, np.float32)
        post = np.zeros(m, np.float32)
        last = np.zeros(m, np.float64)
        lib().orc_net_get_synapses(self.h, source, _p(w), _p(pre), _p(post), _p(last))
        return w, pre, post, last

    def set_trace(self, on: bool):
        lib().orc_net_set_trace(self.h, int(on))

    def trace(self):
        n = lib().orc_net_get_trace(self.h, None, None, None, None, 0)
        st = np.zeros(n, np.int32)
        tg = np.zeros(n, np.int32)
        ty = np.zeros(n, np.uint8)
        w = np.zeros(n, np.float32)
        lib().orc_net_get_trace(self.h, _p(st), _p(tg), _p(ty), _p(w), n)
        return st, tg, ty, w

    def add_link(self, source: "Net", row_ptr, tgt_type, w):
        """Projection of `source`'s population onto this one."""
        a = [np.ascontiguousarray(row_ptr, np.int64), np.ascontiguousarray(tgt_type, np.uint32),
             np.ascontiguousarray(w, np.float32)]
        if lib().orc_net_add_link(self.h, source.h, *[_p(x) for x in a]) < 0:
            raise ValueError("networks of a projection must share max_delay")
        self._keep.append(source)

    def set_poisson_rates(self, source: int, rates):
        r = np.ascontiguousarray(rates, np.float32)
        if lib().orc_net_set_poisson_rates(self.h, source, _p(r)) != 0:
            raise ValueError("not a Poisson source")
        self._keep.append(r)

    def take_log(self):
        m = lib().orc_net_take_log(self.h, None, None, 0)
        st = np.zeros(max(m, 1), np.int32)
        ix = np.zeros(max(m, 1), np.int32)
        lib().orc_net_take_log(self.h, _p(st), _p(ix), m)
        return st[:m].copy(), ix[:m].copy()

    def run(self, first_step: int, n_steps: int, cap: int = 1 << 24):
        st = np.zeros(cap, np.int32)
        ix = np.zeros(cap, np.int32)
        n = lib().orc_net_run(self.h, first_step, n_steps, _p(st), _p(ix), cap)
        if n < 0:
            raise RuntimeError("oracle spike buffer too small")
        return st[:n].copy(), ix[:n].copy()


def circuit_run(nets, first_step: int, n_steps: int) -> None:
    """Lockstep run of several networks (update k of each, in order)."""
    arr = (ctypes.c_void_p * len(nets))(*[n.h for n in nets])
    lib().orc_circuit_run(arr, len(nets), first_step, n_steps)


def from_spec(spec, params: np.ndarray, csr_builder, fp64: bool = False):
    """Oracle population + network for a workloads.NetworkSpec.  csr_builder
    sorts edges into rows (the tests check it against an independent sort)."""
    mask = 0
    if spec.rec is not None:
        mask |= int(np.bitwise_or.reduce(1 << np.unique(spec.rec[3]).astype(np.int64)))
    for _, _, e in spec.poisson:
        mask |= int(np.bitwise_or.reduce(1 << np.unique(e[3]).astype(np.int64)))
    if spec.ext is not None:
        mask |= 1
    pop = Population(spec.model, spec.integrator, params, spec.n, mask=mask, fp64=fp64)
    net = Net(pop, spec.max_delay)
    if spec.rec is not None:
        net.add_self(*csr_builder(spec.n, spec.n, spec.max_delay, *spec.rec))
    for states, rate, e in spec.poisson:
        net.add_poisson(states, rate, *csr_builder(len(states), spec.n, spec.max_delay, *e))
    if spec.ext is not None:
        thr, w_ext, seed = spec.ext
        net.set_external(thr, w_ext, seed)
    return pop, net


class FitOracle:
    """FP64 (or float) forward simulation and FP64 adjoint of the trace-fitting
    path, for a workloads.FitSpec-like set of inputs."""

    MODELS = {"LIFTimeDrivenModel_GPU": 0, "AdExTimeDrivenModel_GPU": 1}

    def __init__(self, model, params, n, steps, syn_type, events, refr=None, dt_s=np.float32(1e-4)):
        self.model = self.MODELS[model]
        p = np.asarray(params, np.float32)
        if self.model == 0:   # LIF: e_exc e_inh v_thr c_m
            fixed = p[[0, 1, 3, 4]]
        else:                 # AdEx: a b thr_slo_fac v_thr tau_w e_exc e_inh e_reset c_m
            fixed = p[[0, 1, 2, 3, 4, 5, 6, 7, 10]]
        self.fixed = np.ascontiguousarray(fixed, np.float32)
        self.refr = np.ascontiguousarray(refr if refr is not None else np.zeros(0), np.float32)
        self.n, self.steps = n, steps
        self.syn_type = np.ascontiguousarray(syn_type, np.uint8)
        self.n_syn = self.syn_type.shape[1]
        self.ev = [np.ascontiguousarray(e, np.int32) for e in events]
        self.dt_s = float(np.float32(dt_s))

    def _common(self):
        return [self.model, _p(self.fixed), _p(self.refr) if len(self.refr) else None, len(self.refr),
                self.n, self.steps, self.n_syn, self.dt_s, *[_p(e) for e in self.ev], _p(self.syn_type)]

    @staticmethod
    def _par(gl, el, te, ti):
        return np.ascontiguousarray(np.stack([gl, el, te, ti]).astype(np.float64))

    def forward(self, gl, el, te, ti, w, fp64=True):
        par = self._par(gl, el, te, ti)
        ww = np.ascontiguousarray(w, np.float64)
        v = np.zeros((self.steps, self.n), np.float64)
        spk = np.zeros((self.steps, self.n), np.uint8)
        lib().orc_fit_forward(*self._common(), _p(par), _p(ww), int(fp64), _p(v), _p(spk))
        return v, spk

    def trajectory(self, gl, el, te, ti, w):
        """Float run: states before each step's integration [steps, 4, n],
        spike flags [steps, n], refractory counts [steps, n]."""
        par = self._par(gl, el, te, ti)
        ww = np.ascontiguousarray(w, np.float64)
        traj = np.zeros((self.steps, 4, self.n), np.float32)
        spk = np.zeros((self.steps, self.n), np.uint8)
        since = np.zeros((self.steps, self.n), np.int32)
        lib().orc_fit_trajectory(*self._common(), _p(par), _p(ww), _p(traj), _p(spk), _p(since))
        return traj, spk, since

    def backward_along(self, gl, el, te, ti, w, grad_v, traj, spk, since):
        """FP64 adjoint evaluated along a given (e.g. float) trajectory."""
        par = self._par(gl, el, te, ti)
        ww = np.ascontiguousarray(w, np.float64)
        gv = np.ascontiguousarray(grad_v, np.float64)
        gp = np.zeros((4, self.n), np.float64)
        gw = np.zeros((self.n, self.n_syn), np.float64)
        lib().orc_fit_backward_along(*self._common(), _p(par), _p(ww), _p(gv),
                                     _p(np.ascontiguousarray(traj, np.float32)),
                                     _p(np.ascontiguousarray(spk, np.uint8)),
                                     _p(np.ascontiguousarray(since, np.int32)), _p(gp), _p(gw))
        return gp, gw

    def backward(self, gl, el, te, ti, w, grad_v):
        par = self._par(gl, el, te, ti)
        ww = np.ascontiguousarray(w, np.float64)
        gv = np.ascontiguousarray(grad_v, np.float64)
        gp = np.zeros((4, self.n), np.float64)
        gw = np.zeros((self.n, self.n_syn), np.float64)
        lib().orc_fit_backward(*self._common(), _p(par), _p(ww), _p(gv), _p(gp), _p(gw))
        return gp, gw
