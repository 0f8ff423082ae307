"""Simulation front-end: EDLUT's Simulation_API / pyedlut over device-resident networks.

`Simulation_API` keeps the reference's call sequence and names
(Kernel/src/simulation/Simulation_API.cpp:63-443, Python wrapper
Kernel/python/simulation_wrapper.pyx:23-716): AddNeuronLayer,
AddLearningRule, AddSynapticLayer, Initialize, AddExternalSpikeActivity,
RunSimulation, GetSpikeActivity, GetWeights.  `Simulation_API.from_files`
builds the same simulation from a network file, a weights file and
optionally an input-spike file (Network::ParseNet, FileInputSpikeDriver).

What the reference does with one global neuron index space and an event
queue, the front-end maps onto this library's device-resident networks:

* layers of equal model and parameters form one neuron-model type (one
  population), in order of first appearance, with their neurons in layer
  order (Network::CreateNeuronLayers, Network.cpp:1016-1110);
* each GPU time-driven type is a population with its network; the synapses
  onto it are grouped by source type into the network's sources, in
  ascending source type: the population itself (recurrent), other GPU types
  (projections), PoissonGeneratorDeviceVector types (Poisson devices, with one
  RandomGenerator per generator taking the global seed sequence in type
  order, PoissonGeneratorDeviceVector.cpp:86-100) and InputSpikeNeuronModel
  types (spike trains);
* synapses of one source type with an STDP rule form a source of their own
  (one rule and one max_weight per source), after the static ones;
* delays are rounded like Network::RoundPropagationDelay (Network.cpp:1311-1317)
  and become the number of updates between emission and consumption: an event
  sent at update s with a rounded delay D is consumed by the first update k
  with t_k > t_s + D, where t_k = t_{k-1} + double(float(step))
  (TimeEventUpdateNeurons.cpp:45-79; an update and an arrival at the same time
  run update first, Event.h:37).  The library consumes it at s + d + 1, so
  d = k - s - 1 must be at least 1;
* an input spike at time t is placed on the update s for which its events are
  consumed by s + d + 1 for every delay of its synapses; a spike time that
  would need different updates for different delays, or that precedes the
  first update's consumption window, is rejected.

Every population runs on the GPU; the host only parses, places input
spikes and reads results.  There is no CPU fallback: the HIP library must be
built.
"""
from __future__ import annotations

import numpy as np

from . import _capi as C
from ._capi import EDLUTException
from . import netfile as NF
from .models import REGISTRY, CreateNeuronModel, integration_method
from .network import Circuit, Network, build_out_csr

MIN_PROPAGATION_DELAY = 0.0001   # Network::minpropagationdelay
MAX_SOURCES = 4                  # sources per population in the library
MAX_RAND = np.float32(2147483647)
POISSON_LOG_CAP = 1 << 22        # logged spikes per Poisson layer with output neurons


class RandomGenerator:
    """RandomGenerator (Kernel/include/simulation/RandomGenerator.h:61-107,
    RandomGenerator.cpp:21-27): an LCG per object, seeded from a global
    sequence g -> (765426813 g + 54321) & 0x7fffffff that every construction
    advances."""

    global_seed = 1

    @classmethod
    def srand(cls, seed: int) -> None:
        cls.global_seed = int(seed) & 0xFFFFFFFF

    @classmethod
    def take_seed(cls) -> int:
        s = cls.global_seed
        cls.global_seed = (cls.global_seed * 765426813 + 54321) & 0x7FFFFFFF
        return s

    def __init__(self):
        self.next_element = RandomGenerator.take_seed()

    def frand(self) -> np.float32:
        self.next_element = (self.next_element * 1103515245 + 12345) & 0x7FFFFFFF
        return np.float32(self.next_element) / MAX_RAND


def round_propagation_delay(delay_s) -> np.ndarray:
    """Network::RoundPropagationDelay of float delays (double result)."""
    d = np.asarray(delay_s, np.float32).astype(np.float64)
    m = MIN_PROPAGATION_DELAY
    r = np.floor(d * (1.0 / m) + m * 0.5) * m
    return np.where(r <= 0, m, r)


def update_times(fdt: float, n: int) -> np.ndarray:
    """t_0..t_n with t_k = t_{k-1} + fdt, added one by one (np.add.accumulate
    is sequential)."""
    inc = np.full(n + 1, fdt, np.float64)
    inc[0] = 0.0
    return np.add.accumulate(inc)


def ring_delay(delay_rounded: float, fdt: float) -> int:
    """Updates between emission and consumption, minus one: with
    t_{s+m} - t_s = m * fdt, the first m with m * fdt > D, less one."""
    m = int(np.floor(delay_rounded / fdt))
    while m * fdt <= delay_rounded:
        m += 1
    while m > 1 and (m - 1) * fdt > delay_rounded:
        m -= 1
    return m - 1


def compress_weights(weights):
    """Run-length encoding of consecutive equal float weights, as
    Network::GetCompressedWeights builds it (Network.cpp:1204-1231)."""
    counts, values = [], []
    for w in np.asarray(weights, np.float32).tolist():
        if values and values[-1] == w:
            counts[-1] += 1
        else:
            counts.append(1)
            values.append(w)
    return counts, values


class _Type:
    def __init__(self, model: NF.ModelDescription, index: int):
        self.model = model
        self.index = index
        self.globals: list[int] = []
        self.output = np.zeros(0, bool)    # per local neuron: layer has output_activity
        self.log = np.zeros(0, bool)


def gpu_model_name(model: NF.ModelDescription):
    """The GPU plug-in that runs a layer's model, or None for an input device."""
    name = model.model_name
    if name in (NF.INPUT_SPIKE, NF.POISSON, NF.INPUT_CURRENT):
        return None
    if name in REGISTRY:
        return name
    if name in NF.CPU_TIME_DRIVEN:
        im = model.param_map.get("int_meth", {})
        if str(im.get("name", "")).startswith("Bifixed_"):
            raise EDLUTException(C.EDLUT_E_ARG, f"{name} with {im['name']}: the CPU Bifixed methods are not the "
                                 "GPU ones; use the model's _GPU plug-in name")
        return NF.CPU_TIME_DRIVEN[name]
    raise EDLUTException(C.EDLUT_E_ARG, f"neuron model {name!r} is not available (GPU time-driven models and their "
                         "CPU names, InputSpikeNeuronModel, InputCurrentNeuronModel, PoissonGeneratorDeviceVector)")


def _canonical(model: NF.ModelDescription) -> tuple:
    """The model's parameters with its defaults filled in (float parameters
    as floats, the integration method as integration_method builds it), so
    that layers compare like the reference's NeuronModel::compare."""
    name = model.model_name
    pm = model.param_map
    g = gpu_model_name(model)
    if g is None:
        if name == NF.POISSON:
            return (name, float(np.float32(pm.get("frequency", 10.0))))
        return (name,)
    vals = tuple(float(np.float32(pm.get(k, v))) for k, v in REGISTRY[g].PARAMS)
    im = dict(pm.get("int_meth", integration_method()))
    imc = integration_method(im.get("name", "Euler"), float(np.float32(im.get("step", 1e-3))),
                             int(im.get("bdf_order", 1)), int(im.get("n_steps", 2)))
    return (name, vals, tuple(sorted(imc.items())))


def _same_model(a: NF.ModelDescription, b: NF.ModelDescription) -> bool:
    return _canonical(a) == _canonical(b)


class SimulationPlan:
    """Host-side mapping of a simulation onto populations and sources (no
    device work): built by Simulation_API.Initialize, consumed by its device
    builder and, in tests, by other back ends."""

    def __init__(self):
        self.types: list[_Type] = []
        self.type_of = np.zeros(0, np.int32)      # global neuron -> type
        self.local_of = np.zeros(0, np.int64)     # global neuron -> index in its type
        self.dt = None                            # float(step) of the GPU types
        self.max_delay = 1
        # per GPU type index: list of source dicts, in the order they are added
        self.sources: dict[int, list[dict]] = {}
        self.rules: list[NF.ModelDescription] = []


class Simulation_API:
    """Simulation_API (Simulation_API.cpp) on device-resident networks."""

    def __init__(self, device: int = 0, full_spike_log: bool = False):
        """full_spike_log keeps every neuron's spikes in the device logs for
        the whole simulation (Network.spikes() of each population, for
        monitoring); by default only populations with output neurons keep a
        log, and GetSpikeActivity empties it on each read."""
        self.device = device
        self.full_spike_log = bool(full_spike_log)
        self.neuron_layer_list: list[NF.NeuronLayerDescription] = []
        self.learning_rule_list: list[NF.ModelDescription] = []
        self.synaptic_layers: list[dict] = []
        self.num_neuron_created = 0
        self.num_synapses_created = 0
        self.initialized = False
        self.simulation_properties = {"num_threads": 1, "num_simulation_queues": 1,
                                      "resolution": -1.0}
        self._cur_changes: list[tuple] = []   # pending current-synapse changes (k, t, seq, type, j, value)
        self._cur_value = {}         # per population: current of each current synapse (pA)
        self._cur_seq = 0
        self.current_log: list[tuple] = []    # (update k, population type, total currents) applied
        self._input_new_from = 0     # input spikes from here on are not uploaded yet
        self._trains_dirty = False
        self._input_times = np.zeros(0, np.float64)
        self._input_cells = np.zeros(0, np.int64)
        self._input_load = np.zeros(0, np.int64)   # per input spike: first update of the run that queues it
        self._file_inputs = False    # AddExternalSpikeActivity is loading an input-spike file
        self.run_log: list[tuple] = []   # device runs: (first update, updates, stop time or -2)
        self.plan: SimulationPlan | None = None
        self.models = {}
        self.nets = {}
        self.circuit = None
        self.current_step = 0       # last update run
        self._in_cursor = -np.inf   # input spikes up to this time were returned
        self._log_read = {}         # per type: spikes of its log already read
        self._syn_pos = None        # per synapse: (type, source slot, position) for GetWeights

    # ---- construction (before Initialize) -------------------------------
    def AddNeuronLayer(self, num_neurons: int = 1, model_name: str = NF.INPUT_SPIKE, param_dict=None,
                       log_activity: bool = False, output_activity: bool = False) -> list:
        """Simulation_API.cpp:63-92; returns the layer's global indices."""
        self._before_init("AddNeuronLayer")
        if num_neurons <= 0:
            raise EDLUTException(C.EDLUT_E_ARG, "a layer needs at least one neuron")
        params = dict(param_dict or {})
        params.setdefault("name", model_name)
        self.neuron_layer_list.append(NF.NeuronLayerDescription(
            int(num_neurons), NF.ModelDescription(model_name, params), bool(output_activity), bool(log_activity)))
        first = self.num_neuron_created
        self.num_neuron_created += int(num_neurons)
        return list(range(first, self.num_neuron_created))

    def AddLearningRule(self, model_name: str, param_dict=None) -> int:
        """Simulation_API.cpp:94-110; only "STDP" (STDPWeightChange) is built."""
        self._before_init("AddLearningRule")
        if model_name != "STDP":
            raise EDLUTException(C.EDLUT_E_ARG, f"learning rule {model_name!r} is not available (only \"STDP\")")
        p = {"max_LTP": 0.010, "tau_LTP": 0.100, "max_LTD": 0.010, "tau_LTD": 0.100}
        p.update(param_dict or {})
        if not (p["tau_LTP"] > 0 and p["tau_LTD"] > 0):
            raise EDLUTException(C.EDLUT_E_ARG, "the STDP time constants must be positive")
        self.learning_rule_list.append(NF.ModelDescription("STDP", p))
        return len(self.learning_rule_list) - 1

    def AddSynapticLayer(self, source_list, target_list, param_dict=None) -> list:
        """Simulation_API.cpp:112-135 with the per-synapse defaults of
        Network::CreateConnections (Network.cpp:790-900): delay 0.001 s, type 0,
        weight 1, max_weight 1, no learning rule; each parameter a scalar or a
        list."""
        self._before_init("AddSynapticLayer")
        src = np.asarray(source_list, np.int64).ravel()
        tgt = np.asarray(target_list, np.int64).ravel()
        if len(src) != len(tgt):
            raise EDLUTException(C.EDLUT_E_ARG, "source and target lists differ in length")
        n = len(src)
        p = dict(param_dict or {})

        def get(key, default, dtype):
            v = p.pop(key, default)
            a = np.asarray(v, dtype)
            return np.full(n, a, dtype) if a.ndim == 0 else a.ravel()[:n].astype(dtype)

        layer = dict(source=src.astype(np.int32), target=tgt.astype(np.int32),
                     delay=get("delay", 0.001, np.float32), type=get("type", 0, np.int32),
                     weight=get("weight", 1.0, np.float32), max_weight=get("max_weight", 1.0, np.float32),
                     wchange=get("wchange", -1, np.int32), trigger_wchange=get("trigger_wchange", -1, np.int32))
        if p:
            raise EDLUTException(C.EDLUT_E_ARG, f"unknown synapse parameters {sorted(p)}")
        self.synaptic_layers.append(layer)
        first = self.num_synapses_created
        self.num_synapses_created += n
        return list(range(first, self.num_synapses_created))

    def SetSimulationParameters(self, params_dict) -> None:
        """Simulation_API.cpp:405-426: only known keys are accepted."""
        self._before_init("SetSimulationParameters")
        for k, v in dict(params_dict).items():
            if k not in self.simulation_properties:
                raise EDLUTException(C.EDLUT_E_ARG, f"invalid simulation parameter {k!r}")
            self.simulation_properties[k] = v

    def SetRandomGeneratorSeed(self, seed: int) -> None:
        """Simulation_API.cpp:579-581."""
        RandomGenerator.srand(seed)

    @staticmethod
    def GetAvailableNeuronModels() -> list:
        return sorted(REGISTRY) + [NF.INPUT_SPIKE, NF.INPUT_CURRENT, NF.POISSON]

    @staticmethod
    def GetNeuronModelDefParams(model_name: str) -> dict:
        if model_name in REGISTRY:
            return REGISTRY[model_name].GetDefaultParameters()
        if model_name == NF.POISSON:
            return {"frequency": 10.0, "name": NF.POISSON}
        if model_name in (NF.INPUT_SPIKE, NF.INPUT_CURRENT):
            return {"name": model_name}
        raise EDLUTException(C.EDLUT_E_ARG, f"Neuron model \"{model_name}\" does not exist")

    @staticmethod
    def GetLearningRuleDefParams(model_name: str) -> dict:
        if model_name != "STDP":
            raise EDLUTException(C.EDLUT_E_ARG, f"learning rule {model_name!r} is not available")
        return {"max_LTP": 0.010, "tau_LTP": 0.100, "max_LTD": 0.010, "tau_LTD": 0.100}

    @staticmethod
    def GetIntegrationMethodDefParams(model_name: str) -> dict:
        if model_name.startswith("Bifixed_"):
            return integration_method(model_name, 1e-3, n_steps=2)
        return integration_method(model_name, 1e-3)

    @classmethod
    def from_files(cls, netfile: str, wfile: str, input_spike_file: str | None = None,
                   base_dir: str | None = None, device: int = 0,
                   input_current_file: str | None = None, full_spike_log: bool = False) -> "Simulation_API":
        """A simulation built from a network file and a weights file
        (Network::ParseNet; the weights file's random weights take one
        RandomGenerator), with the input spikes of a FileInputSpikeDriver file
        queued for the first run.  Model files are read relative to base_dir
        (default: the working directory, as the reference's fopen)."""
        rng = RandomGenerator()
        layers, rules, syn = NF.parse_net(netfile, wfile, base_dir=base_dir, rng=rng)
        sim = cls(device, full_spike_log)
        for L in layers:
            sim.AddNeuronLayer(L.num_neurons, L.neuron_model.model_name, L.neuron_model.param_map,
                               L.log_activity, L.output_activity)
        sim.learning_rule_list = list(rules)
        sim.synaptic_layers.append(dict(source=syn.source, target=syn.target, delay=syn.delay, type=syn.type,
                                        weight=syn.weight, max_weight=syn.max_weight, wchange=syn.wchange,
                                        trigger_wchange=syn.trigger_wchange))
        sim.num_synapses_created = len(syn.source)
        sim.Initialize()
        if input_spike_file is not None:
            t, c = NF.parse_input_spikes(input_spike_file, sim.num_neuron_created)
            sim._file_inputs = True
            try:
                sim.AddExternalSpikeActivity(t, c)
            finally:
                sim._file_inputs = False
        if input_current_file is not None:
            t, c, v = NF.parse_input_currents(input_current_file, sim.num_neuron_created)
            sim.AddExternalCurrentActivity(t, c, v)
        return sim

    # ---- initialisation ---------------------------------------------------
    def _before_init(self, what):
        if self.initialized:
            raise EDLUTException(C.EDLUT_E_STATE, f"{what} must be called before Initialize")

    def _after_init(self, what):
        if not self.initialized:
            raise EDLUTException(C.EDLUT_E_STATE, f"{what} must be called after Initialize")

    def build_plan(self) -> SimulationPlan:
        """Maps layers, rules and synapses onto populations and sources (host
        logic only)."""
        plan = SimulationPlan()
        plan.rules = list(self.learning_rule_list)
        n_total = self.num_neuron_created
        plan.type_of = np.zeros(n_total, np.int32)
        plan.local_of = np.zeros(n_total, np.int64)
        outs, logs = [], []
        g = 0
        for L in self.neuron_layer_list:
            m = L.neuron_model
            gpu_model_name(m)
            t = next((T for T in plan.types if _same_model(T.model, m)), None)
            if t is None:
                t = _Type(m, len(plan.types))
                plan.types.append(t)
                outs.append([])
                logs.append([])
            for _ in range(L.num_neurons):
                plan.type_of[g] = t.index
                plan.local_of[g] = len(t.globals)
                t.globals.append(g)
                outs[t.index].append(L.output_activity)
                logs[t.index].append(L.log_activity)
                g += 1
        for t in plan.types:
            t.output = np.array(outs[t.index], bool)
            t.log = np.array(logs[t.index], bool)
        gpu = [t for t in plan.types if gpu_model_name(t.model) is not None]
        steps = {float(np.float32(t.model.param_map.get("int_meth", integration_method())["step"]))
                 for t in gpu}
        if len(steps) > 1:
            raise EDLUTException(C.EDLUT_E_ARG, "every GPU neuron model must use the same integration step "
                                 f"(found {sorted(steps)})")
        plan.dt = steps.pop() if steps else float(np.float32(1e-3))
        fdt = float(np.float32(plan.dt))
        # PoissonGeneratorDeviceVector samples on its own 0.1 ms clock
        # (PoissonGeneratorDeviceVector.cpp:26,137); the device draws one
        # sample per population update, so the two clocks must agree
        if any(t.model.model_name == NF.POISSON for t in plan.types) and fdt != float(np.float32(1e-4)):
            raise EDLUTException(C.EDLUT_E_ARG, "Poisson generator layers sample every 0.1 ms: the GPU neuron "
                                 f"models must use an integration step of 0.0001 s (found {fdt:g})")

        cat = {k: np.concatenate([L[k] for L in self.synaptic_layers]) if self.synaptic_layers else
               np.zeros(0, np.int32) for k in ("source", "target", "delay", "type", "weight", "max_weight",
                                                "wchange", "trigger_wchange")}
        nsyn = len(cat["source"])
        for k in ("source", "target"):
            if nsyn and (cat[k].min() < 0 or cat[k].max() >= n_total):
                raise EDLUTException(C.EDLUT_E_ARG, f"a synapse {k} is not a neuron of the network")
        if nsyn and (cat["trigger_wchange"] >= 0).any():
            raise EDLUTException(C.EDLUT_E_ARG, "trigger learning rules are not available")
        if nsyn and (cat["wchange"] >= len(plan.rules)).any():
            raise EDLUTException(C.EDLUT_E_ARG, "a synapse names a learning rule that does not exist")
        stype = plan.type_of[cat["source"]] if nsyn else np.zeros(0, np.int32)
        ttype = plan.type_of[cat["target"]] if nsyn else np.zeros(0, np.int32)
        for q in np.nonzero(np.isin(ttype, [t.index for t in plan.types if t not in gpu]))[0][:1]:
            raise EDLUTException(C.EDLUT_E_ARG, f"synapse {q}: its target is an input device")
        # CheckSynapseType (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:239-290):
        # types 0 (AMPA), 1 (GABA) and 2 (NMDA) take spikes, type 3 currents
        if nsyn and ((cat["type"] < 0) | (cat["type"] > 3)).any():
            raise EDLUTException(C.EDLUT_E_ARG, "synapse types are 0 (AMPA), 1 (GABA), 2 (NMDA) and 3 (current)")
        cur_src = np.array([plan.types[t].model.model_name == NF.INPUT_CURRENT for t in stype], bool)
        is_cur = cat["type"] == 3 if nsyn else np.zeros(0, bool)
        if nsyn and (is_cur != cur_src).any():
            q = int(np.nonzero(is_cur != cur_src)[0][0])
            raise EDLUTException(C.EDLUT_E_ARG, f"synapse {q}: type 3 takes currents from InputCurrentNeuronModel "
                                 "neurons, types 0-2 take spikes")
        dround = round_propagation_delay(cat["delay"])
        dsteps = np.zeros(nsyn, np.int32)
        for D in np.unique(dround[~is_cur]):
            d = ring_delay(float(D), fdt)
            if d < 1:
                raise EDLUTException(C.EDLUT_E_ARG, f"a delay of {D:g} s is consumed by the update right after the "
                                     f"spike at step {fdt:g} s; the device networks need delays of at least two "
                                     "updates' spacing minus the step")
            dsteps[dround == D] = d
        plan.max_delay = max(int(dsteps.max()), 1) if nsyn else 1
        if plan.max_delay > 256:
            raise EDLUTException(C.EDLUT_E_ARG, "delays above 256 updates are not available")
        plan.syn_dsteps = dsteps
        plan.syn_dround = dround
        plan.syn = cat
        plan.syn_stype = stype
        plan.syn_ttype = ttype
        self._syn_pos = np.full((nsyn, 3), -1, np.int64)
        # current synapses stay on the host: per target population, in the
        # order they were created, which is each target's CurrentSynapseModel
        # subindex order (SetSubindexType in CheckSynapseType)
        plan.current = {}
        for t in gpu:
            cq = np.nonzero((ttype == t.index) & is_cur)[0]
            if len(cq):
                plan.current[t.index] = dict(syn=cq, src=cat["source"][cq].astype(np.int64),
                                             tgt=plan.local_of[cat["target"][cq]], delay=dround[cq])
        for t in gpu:
            srcs = []
            onto = np.nonzero((ttype == t.index) & ~is_cur)[0]
            for st in sorted(set(stype[onto].tolist())):
                sel = onto[stype[onto] == st]
                groups = [("static", sel[cat["wchange"][sel] < 0])]
                plastic = sel[cat["wchange"][sel] >= 0]
                for key in sorted({(int(cat["wchange"][q]), float(cat["max_weight"][q])) for q in plastic}):
                    groups.append((key, plastic[(cat["wchange"][plastic] == key[0]) &
                                                (cat["max_weight"][plastic] == np.float32(key[1]))]))
                s_model = plan.types[st].model.model_name
                kind = ("self" if st == t.index else "link" if gpu_model_name(plan.types[st].model)
                        else "poisson" if s_model == NF.POISSON else "train")
                for key, q in groups:
                    if len(q) == 0:
                        continue
                    if key != "static" and kind == "train":
                        raise EDLUTException(C.EDLUT_E_ARG, "STDP on synapses from input-spike layers is not "
                                             "available")
                    n_src = len(plan.types[st].globals)
                    order_src = plan.local_of[cat["source"][q]]
                    order_tgt = plan.local_of[cat["target"][q]]
                    srcs.append(dict(kind=kind, src_type=st, n_src=n_src, syn=q,
                                     edges=(order_src.astype(np.int32), order_tgt.astype(np.int32),
                                            dsteps[q].astype(np.int32), cat["type"][q].astype(np.uint8),
                                            cat["weight"][q].astype(np.float32)),
                                     stdp=None if key == "static" else
                                     (plan.rules[key[0]].param_map, np.float32(key[1]))))
            if sum(1 for s in srcs if s["kind"] == "self") > 1:
                raise EDLUTException(C.EDLUT_E_ARG, "recurrent synapses of one population must share one "
                                     "learning rule (or none)")
            if len(srcs) > MAX_SOURCES:
                raise EDLUTException(C.EDLUT_E_CAP, f"population of {t.model.model_name} receives from "
                                     f"{len(srcs)} sources; at most {MAX_SOURCES} are available")
            plan.sources[t.index] = srcs
        # Poisson generators: one RandomGenerator each, in type order
        # (Network::InitializeStates, PoissonGeneratorDeviceVector::InitializeStates)
        plan.poisson_states = {}
        for t in plan.types:
            if t.model.model_name == NF.POISSON:
                plan.poisson_states[t.index] = np.array([RandomGenerator.take_seed() for _ in t.globals],
                                                        np.uint32)
        return plan

    def Initialize(self) -> None:
        """Simulation_API.cpp:352-403: builds the populations, networks and the
        circuit on the device."""
        self._before_init("Initialize")
        plan = self.build_plan()
        self.plan = plan
        gpu = [t for t in plan.types if gpu_model_name(t.model) is not None]
        for t in gpu:
            params = {k: v for k, v in t.model.param_map.items() if k != "name"}
            m = CreateNeuronModel(gpu_model_name(t.model), params)
            for s in plan.sources.get(t.index, []):
                for ty in np.unique(s["edges"][3]):
                    m.CheckSynapseType(int(ty))
            if t.index in plan.current:
                m.CheckSynapseType(3, source_emits_spikes=False)
            m.InitializeStates(len(t.globals), 0, self.device)
            self.models[t.index] = m
            self.nets[t.index] = Network(m, plan.max_delay)
        for t in gpu:
            net = self.nets[t.index]
            # event-queue order: models are the layer types, ranked in type order
            net.set_model_rank(t.index)
            for slot, s in enumerate(plan.sources.get(t.index, [])):
                rp, tt, ww = build_out_csr(s["n_src"], len(t.globals), plan.max_delay, *s["edges"])
                s["csr"] = (rp, tt, ww)
                if s["kind"] == "self":
                    net.add_recurrent(rp, tt, ww)
                elif s["kind"] == "link":
                    net.add_projection(self.nets[s["src_type"]], rp, tt, ww)
                elif s["kind"] == "poisson":
                    st = plan.types[s["src_type"]]
                    net.add_poisson(plan.poisson_states[s["src_type"]], 0.0, rp, tt, ww)
                    rates = np.full(s["n_src"], np.float32(st.model.param_map.get("frequency", 10.0)), np.float32)
                    net.set_poisson_rates(slot, rates)
                else:
                    net.add_spike_train(s["n_src"], np.zeros(1, np.int64), np.zeros(0, np.int32), rp, tt, ww)
                if s["kind"] in ("poisson", "train"):
                    net.set_source_model(slot, s["src_type"], s["src_type"])
                if s["stdp"] is not None:
                    p, wmax = s["stdp"]
                    net.set_stdp(slot, float(p["max_LTP"]), float(p["tau_LTP"]), float(p["max_LTD"]),
                                 float(p["tau_LTD"]), float(wmax))
                s["slot"] = slot
                # where each synapse of the layer lists sits in the CSR
                # (build_out_csr orders by source, delay, then target; a
                # stable sort reproduces it)
                e = s["edges"]
                order = np.lexsort((e[1], e[2], e[0]))
                pos = np.empty(len(order), np.int64)
                pos[order] = np.arange(len(order))
                self._syn_pos[s["syn"]] = np.stack([np.full(len(pos), t.index), np.full(len(pos), slot), pos], 1)
        for tidx, cur in plan.current.items():
            self._cur_value[tidx] = np.zeros(len(cur["syn"]), np.float32)
        # Poisson layers with output neurons: one of their device copies logs
        # the generators' spikes; a layer with no synapse onto a GPU
        # population gets a copy without synapses, on the first population's
        # network, that only logs
        self._poisson_log = {}
        for t in plan.types:
            if t.model.model_name != NF.POISSON or not t.output.any():
                continue
            for tidx, srcs in plan.sources.items():
                hit = [s2 for s2 in srcs if s2["kind"] == "poisson" and s2["src_type"] == t.index]
                if hit:
                    break
            else:
                if not gpu:
                    raise EDLUTException(C.EDLUT_E_ARG, "Poisson generator layers need a GPU neuron population "
                                         "to run beside")
                tidx = gpu[0].index
                srcs = plan.sources.setdefault(tidx, [])
                if len(srcs) >= MAX_SOURCES:
                    raise EDLUTException(C.EDLUT_E_CAP, f"population of {gpu[0].model.model_name} has no source "
                                         "slot left for logging a Poisson output layer")
                n_src = len(t.globals)
                empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                         np.zeros(0, np.uint8), np.zeros(0, np.float32))
                rp, tt, ww = build_out_csr(n_src, len(gpu[0].globals), plan.max_delay, *empty)
                net = self.nets[tidx]
                slot = len(srcs)
                net.add_poisson(plan.poisson_states[t.index], 0.0, rp, tt, ww)
                net.set_poisson_rates(slot, np.full(n_src, np.float32(t.model.param_map.get("frequency", 10.0)),
                                                    np.float32))
                net.set_source_model(slot, t.index, t.index)
                hit = [dict(kind="poisson", src_type=t.index, n_src=n_src, syn=np.zeros(0, np.int64), edges=empty,
                            stdp=None, csr=(rp, tt, ww), slot=slot)]
                srcs.append(hit[0])
            self.nets[tidx].set_source_log(hit[0]["slot"], POISSON_LOG_CAP)
            self._poisson_log[t.index] = (tidx, hit[0]["slot"])
            self._log_read[("poisson", t.index)] = 0
        # spike logs: only populations with output neurons keep one
        if not self.full_spike_log:
            for t in gpu:
                if not t.output.any():
                    self.nets[t.index].set_log_capacity(0)
        if gpu:
            self.circuit = Circuit([self.nets[t.index] for t in gpu])
        self._log_read.update({t.index: 0 for t in gpu})
        self.initialized = True

    # ---- run --------------------------------------------------------------
    def AddExternalSpikeActivity(self, event_time_list, neuron_index_list) -> None:
        """Simulation_API.cpp:137-153: spikes of InputSpikeNeuronModel neurons,
        delivered from the next RunSimulation on (times before the current
        simulation time are rejected)."""
        self._after_init("AddExternalSpikeActivity")
        t = np.asarray(event_time_list, np.float64).ravel()
        c = np.asarray(neuron_index_list, np.int64).ravel()
        if len(t) != len(c):
            raise EDLUTException(C.EDLUT_E_ARG, "time and neuron lists differ in length")
        if len(c) and (c.min() < 0 or c.max() >= self.num_neuron_created):
            raise EDLUTException(C.EDLUT_E_ARG, "input spike neuron index out of range")
        for q in range(len(c)):
            if self.plan.types[self.plan.type_of[c[q]]].model.model_name != NF.INPUT_SPIKE:
                raise EDLUTException(C.EDLUT_E_ARG, f"neuron {c[q]} is not an InputSpikeNeuronModel neuron")
        self._input_times = np.concatenate([self._input_times, t])
        self._input_cells = np.concatenate([self._input_cells, c])
        # queued by the next run (Simulation_API::AddExternalSpikeActivity
        # inserts them into the event queue between runs); a file's spikes
        # are queued when the simulation starts (FileInputSpikeDriver)
        self._input_load = np.concatenate([self._input_load, np.full(len(t), self._load_update(), np.int64)])
        self._trains_dirty = True

    def _load_update(self) -> int:
        return 0 if self._file_inputs else self.current_step + 1

    def AddExternalCurrentActivity(self, event_time_list, neuron_index_list, current_value_list) -> None:
        """Simulation_API.cpp:155-171: currents (pA) of InputCurrentNeuronModel
        neurons from the given times on.  Each reaches the populations through
        its current synapses (type 3) after their delay: from the first update
        after the arrival, the synapse carries the new current, and a target's
        external current is the float sum of its current synapses' currents in
        the order they were created (CurrentSynapseModel::GetTotalCurrent,
        CurrentSynapseModel.cpp:76-82; ProcessInputCurrent,
        LIFTimeDrivenModel_GPU_C_INTERFACE.cu:72-79)."""
        self._after_init("AddExternalCurrentActivity")
        t = np.asarray(event_time_list, np.float64).ravel()
        c = np.asarray(neuron_index_list, np.int64).ravel()
        v = np.asarray(current_value_list, np.float32).ravel()
        if not (len(t) == len(c) == len(v)):
            raise EDLUTException(C.EDLUT_E_ARG, "time, neuron and current lists differ in length")
        if len(c) and (c.min() < 0 or c.max() >= self.num_neuron_created):
            raise EDLUTException(C.EDLUT_E_ARG, "input current neuron index out of range")
        plan = self.plan
        for q in range(len(c)):
            if plan.types[plan.type_of[c[q]]].model.model_name != NF.INPUT_CURRENT:
                raise EDLUTException(C.EDLUT_E_ARG, f"neuron {c[q]} is not an InputCurrentNeuronModel neuron")
        if len(t) == 0:
            return
        fdt = float(np.float32(plan.dt))
        dmax = max((float(cur["delay"].max()) for cur in plan.current.values()), default=0.0)
        grid = update_times(fdt, int(np.ceil((t.max() + dmax) / fdt)) + 8)
        new = []
        for q in range(len(t)):
            for tidx, cur in plan.current.items():
                for j in np.nonzero(cur["src"] == c[q])[0]:
                    k = int(np.searchsorted(grid, t[q] + float(cur["delay"][j]), side="right"))
                    if k <= self.current_step:
                        raise EDLUTException(C.EDLUT_E_ARG, f"input current at {t[q]!r} s of neuron {c[q]} reaches "
                                             "its target in updates already simulated")
                    new.append((k, float(t[q]), self._cur_seq, tidx, int(j), np.float32(v[q])))
                    self._cur_seq += 1
        self._cur_changes += new

    def _apply_currents(self, k: int, changes) -> None:
        touched = set()
        for (_, _, _, tidx, j, value) in changes:
            self._cur_value[tidx][j] = value
            touched.add(tidx)
        for tidx in sorted(touched):
            cur = self.plan.current[tidx]
            total = np.zeros(len(self.plan.types[tidx].globals), np.float32)
            vals = self._cur_value[tidx]
            for j in range(len(cur["syn"])):
                total[cur["tgt"][j]] = np.float32(total[cur["tgt"][j]] + vals[j])
            self.models[tidx].ProcessInputCurrent(total)
            self.current_log.append((k, tidx, total.copy()))

    def place_input_spikes(self, times, cells):
        """Update step of each input spike (see the module notes): 0 for a
        spike whose neuron has no synapses; raises for spikes that cannot be
        placed."""
        plan = self.plan
        fdt = float(np.float32(plan.dt))
        times = np.asarray(times, np.float64)
        cells = np.asarray(cells, np.int64)
        steps = np.zeros(len(times), np.int64)
        if len(times) == 0:
            return steps
        kmax = int(np.ceil((times.max() + plan.max_delay * fdt + 1e-3) / fdt)) + plan.max_delay + 8
        grid = update_times(fdt, kmax)
        syn = plan.syn
        delays_of = {}
        for q in range(len(times)):
            g = int(cells[q])
            if g not in delays_of:
                out = np.nonzero(syn["source"] == g)[0]
                delays_of[g] = sorted({(float(plan.syn_dround[e]), int(plan.syn_dsteps[e])) for e in out})
            if not delays_of[g]:
                continue
            cand = {int(np.searchsorted(grid, times[q] + D, side="right")) - d - 1 for D, d in delays_of[g]}
            if len(cand) != 1:
                raise EDLUTException(C.EDLUT_E_ARG, f"input spike at {times[q]!r} s of neuron {g}: its delays "
                                     "reach the population at inconsistent updates (too close to an update time)")
            s = cand.pop()
            if s < 1:
                raise EDLUTException(C.EDLUT_E_ARG, f"input spike at {times[q]!r} s of neuron {g} comes before "
                                     "the first update")
            steps[q] = s
        return steps

    def _upload_trains(self):
        plan = self.plan
        steps = self.place_input_spikes(self._input_times, self._input_cells)
        new = steps[self._input_new_from:]
        if ((new >= 1) & (new <= self.current_step)).any():
            q = self._input_new_from + int(np.nonzero((new >= 1) & (new <= self.current_step))[0][0])
            raise EDLUTException(C.EDLUT_E_ARG, f"input spike at {self._input_times[q]!r} s falls in updates "
                                 "already simulated")
        for tidx, srcs in plan.sources.items():
            for s in srcs:
                if s["kind"] != "train":
                    continue
                ptr, idx, times, rank, load = self.train_schedule(s["src_type"], steps)
                self.nets[tidx].set_spike_train(s["slot"], ptr, idx)
                # the InputSpike events: times, load order (the order the
                # spikes were given in) and the run that queues them
                self.nets[tidx].set_spike_train_times(s["slot"], times, rank, load)
        self._input_new_from = len(self._input_times)
        self._trains_dirty = False

    def train_schedule(self, src_type: int, steps):
        """Spike-train schedule of input layer `src_type` (updates from 1,
        ascending neuron within an update) and, per entry, the spike's time
        and load rank (its index among the spikes given)."""
        plan = self.plan
        q = np.nonzero((plan.type_of[self._input_cells] == src_type) & (steps >= 1))[0]
        loc = plan.local_of[self._input_cells[q]]
        stp = steps[q]
        order = np.lexsort((loc, stp))
        q, loc, stp = q[order], loc[order], stp[order]
        if len(q) > 1 and ((stp[1:] == stp[:-1]) & (loc[1:] == loc[:-1])).any():
            raise EDLUTException(C.EDLUT_E_ARG, "two spikes of one input neuron fall on the same update")
        n_steps = int(stp.max()) if len(stp) else 0
        ptr = np.zeros(n_steps + 1, np.int64)
        np.add.at(ptr, stp, 1)
        return (np.cumsum(ptr), loc.astype(np.int32), self._input_times[q].astype(np.float64), q.astype(np.int64),
                self._input_load[q].astype(np.int64))

    def RunSimulation(self, end_simulation_time: float) -> None:
        """Simulation_API.cpp:428-442: every update with t_k <= end (an update
        runs before the stop event of the same time, Event.h:37)."""
        self._after_init("RunSimulation")
        fdt = float(np.float32(self.plan.dt))
        est = int(np.floor(end_simulation_time / fdt)) + 2
        grid = update_times(fdt, max(est, 1))
        k_end = int(np.searchsorted(grid, end_simulation_time, side="right")) - 1
        if self._trains_dirty:
            self._upload_trains()
        # current changes split the run: updates before k, then the new
        # currents, then update k onward
        due = sorted(x for x in self._cur_changes if x[0] <= k_end)
        self._cur_changes = [x for x in self._cur_changes if x[0] > k_end]
        # one reference run: the device runs it in pieces (before each current
        # change) that carry no stop event, and the last piece stops at `end`
        # (its stop event shares the event queue, Simulation.cpp:322-373)
        for k in sorted({x[0] for x in due}):
            if k - 1 > self.current_step and self.circuit is not None:
                self._run_piece(k - 1 - self.current_step, -2.0)
                self.current_step = k - 1
            self._apply_currents(k, [x for x in due if x[0] == k])
        n = k_end - self.current_step
        if self.circuit is not None:
            self._run_piece(max(n, 0), float(end_simulation_time))
        if n > 0:
            self.current_step = k_end

    def _run_piece(self, n: int, stop: float) -> None:
        self.circuit.set_stop_time(stop)
        self.circuit.run(n, first_step=self.current_step + 1)
        self.run_log.append((self.current_step + 1, n, stop))

    def update_time(self, k: int) -> float:
        return float(update_times(float(np.float32(self.plan.dt)), k)[k])

    def GetSpikeActivity(self):
        """Simulation_API.cpp:173-204: (times, neuron indices) of the output
        neurons' spikes since the last call, in time order.  Time-driven
        neurons fire at their update time; input neurons at their input
        times."""
        self._after_init("GetSpikeActivity")
        plan = self.plan
        fdt = float(np.float32(plan.dt))
        grid = update_times(fdt, max(self.current_step, 1))
        t_now = grid[self.current_step] if self.current_step else 0.0
        rows = []
        for tidx, net in self.nets.items():
            T = plan.types[tidx]
            if not T.output.any():
                continue
            st, ix = net.spikes()
            new = slice(self._log_read[tidx], len(st))
            self._log_read[tidx] = len(st)
            st, ix = st[new], ix[new]
            keep = T.output[ix]
            g = np.asarray(T.globals, np.int64)[ix[keep]]
            rows.append((grid[st[keep]], np.full(len(g), tidx), ix[keep], g))
        # Poisson layers with output neurons: their generators' spikes
        for ptype, (tidx, slot) in self._poisson_log.items():
            st, ix = self.nets[tidx].source_spikes(slot)
            key = ("poisson", ptype)
            new = slice(self._log_read[key], len(st))
            self._log_read[key] = len(st)
            st, ix = st[new], ix[new]
            T = plan.types[ptype]
            keep = T.output[ix]
            rows.append((grid[st[keep]], np.full(int(keep.sum()), ptype), ix[keep],
                         np.asarray(T.globals, np.int64)[ix[keep]]))
        if not self.full_spike_log:
            # the logs are drained, as the output drivers' buffers
            # (GetBufferedSpikes empties its buffer on each read)
            for tidx, net in self.nets.items():
                if plan.types[tidx].output.any() or any(v[0] == tidx for v in self._poisson_log.values()):
                    net.clear_log()
            for key in self._log_read:
                self._log_read[key] = 0
        # input neurons with output activity: their spikes up to the current time
        if len(self._input_times):
            prev = self._in_cursor
            sel = (self._input_times > prev) & (self._input_times <= t_now)
            for q in np.nonzero(sel)[0]:
                T = plan.types[plan.type_of[self._input_cells[q]]]
                loc = plan.local_of[self._input_cells[q]]
                if T.output[loc]:
                    rows.append((np.array([self._input_times[q]]), np.array([T.index]), np.array([loc]),
                                 np.array([self._input_cells[q]])))
            self._in_cursor = t_now
        if not rows:
            return [], []
        t = np.concatenate([r[0] for r in rows])
        ty = np.concatenate([r[1] for r in rows])
        lo = np.concatenate([r[2] for r in rows])
        g = np.concatenate([r[3] for r in rows])
        order = np.lexsort((lo, ty, t))
        return t[order].tolist(), g[order].tolist()

    def GetWeights(self) -> list:
        """Simulation_API.cpp:326-337: every synapse's weight, in creation order."""
        self._after_init("GetWeights")
        n = len(self._syn_pos)
        w = np.zeros(n, np.float32)
        cache = {}
        for q in range(n):
            tidx, slot, pos = self._syn_pos[q]
            if tidx < 0:
                w[q] = self.plan.syn["weight"][q]      # current synapse (host side)
                continue
            key = (int(tidx), int(slot))
            if key not in cache:
                cache[key] = self.nets[key[0]].synapses(key[1])[2]
            w[q] = cache[key][pos]
        return w.tolist()

    def GetCompressedWeights(self):
        """Simulation_API.cpp:313-324 / Network::GetCompressedWeights
        (Network.cpp:1204-1231): runs of equal consecutive weights as
        (counts, weights)."""
        return compress_weights(self.GetWeights())

    def GetSelectedWeights(self, indexes) -> list:
        w = self.GetWeights()
        return [w[i] for i in indexes]

    def close(self):
        if self.circuit is not None:
            self.circuit.close()
            self.circuit = None
