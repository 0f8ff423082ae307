"""Readers of EDLUT's text input files: network, weights, neuron models,
learning rules and input spikes.

The formats and their checks follow the reference's parsers:

* network file: Network::ParseNet (Kernel/src/spike/Network.cpp:422-645);
* weights file: the second half of ParseNet (Network.cpp:647-706);
* neuron-model files ``<ident>.cfg``: NeuronModelFactory::ParseNeuronModel
  (Kernel/src/neuron_model/NeuronModelFactory.cu:280-291) and each GPU model's
  ParseNeuronModel (e.g. LIFTimeDrivenModel_GPU_C_INTERFACE.cu:412-495), with
  the integration method read by
  TimeDrivenNeuronModel_GPU_C_INTERFACE::ParseIntegrationMethod
  (TimeDrivenNeuronModel_GPU_C_INTERFACE.cuh:384-394) and the method's own
  reader (IntegrationMethod_GPU_C_INTERFACE.cuh:230-239,
  BDFn_GPU_C_INTERFACE.cuh:314-333, BifixedStep_GPU_C_INTERFACE.cuh:184-199);
* Poisson generator devices: PoissonGeneratorDeviceVector::ParseNeuronModel
  (PoissonGeneratorDeviceVector.cpp:176-200);
* learning rules: STDPWeightChange::ParseLearningRule (STDPWeightChange.cpp:77-98);
* input spikes: FileInputSpikeDriver::LoadInputs (FileInputSpikeDriver.cpp:45-85).

Files are read through `_CFile`, which restates the stdio behaviour these
parsers depend on: `skip_comments` / `is_end_line`
(Kernel/src/simulation/Utils.cpp:18-56; a comment starts at '/' and runs to the
end of the line), and the conversions ``%i`` (base from the prefix: 0x hex, 0
octal), ``%d``, ``%li``, ``%f`` (to float) and ``%64s``.  Line numbers in errors
count the newlines these helpers consume, as the reference's do.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

from ._capi import EDLUT_E_ARG, EDLUTException

MAXIDSIZE = 64


class EDLUTFileException(EDLUTException):
    """A parse error in an input file (EDLUTFileException: task, file, line)."""

    def __init__(self, msg: str, filename: str, line: int):
        super().__init__(EDLUT_E_ARG, f"{filename}:{line}: {msg}")
        self.filename = filename
        self.line = line


_INT_I = re.compile(r"[+-]?(0[xX][0-9a-fA-F]+|0[0-7]*|[1-9][0-9]*)")
_INT_D = re.compile(r"[+-]?[0-9]+")
_FLOAT = re.compile(r"[+-]?(([0-9]+\.?[0-9]*|\.[0-9]+)([eE][+-]?[0-9]+)?|inf(inity)?|nan)", re.IGNORECASE)
_SPACE = " \t\n\v\f\r"


class _CFile:
    """A text file read like the reference reads it with fgetc/fscanf."""

    def __init__(self, path: str):
        self.path = path
        try:
            with open(path, "rb") as fh:
                self.s = fh.read().decode("latin-1")
        except OSError:
            raise EDLUTFileException("the file cannot be opened", path, 0) from None
        self.i = 0
        self.line = 1

    # Utils.cpp:18-26: spaces and newlines (newlines counted)
    def _skip_spaces(self):
        while self.i < len(self.s) and self.s[self.i] in " \n":
            if self.s[self.i] == "\n":
                self.line += 1
            self.i += 1
        return self.s[self.i] if self.i < len(self.s) else None

    def skip_comments(self):
        """Utils.cpp:28-39."""
        while self._skip_spaces() == "/":
            while self.i < len(self.s) and self.s[self.i] != "\n":
                self.i += 1
            if self.i < len(self.s):
                self.i += 1
                self.line += 1

    def is_end_line(self) -> bool:
        """Utils.cpp:41-56: after spaces, a newline, end of file or a comment."""
        while self.i < len(self.s) and self.s[self.i] == " ":
            self.i += 1
        return self.i >= len(self.s) or self.s[self.i] in "\n/"

    def _ws(self):
        # every numeric/string conversion of fscanf skips white space first
        while self.i < len(self.s) and self.s[self.i] in _SPACE:
            self.i += 1

    def _match(self, rx):
        self._ws()
        m = rx.match(self.s, self.i)
        if not m:
            return None
        self.i = m.end()
        return m.group(0)

    def scan_i(self):
        """fscanf "%i" / "%li": the base follows the prefix."""
        t = self._match(_INT_I)
        if t is None:
            return None
        neg = t.startswith("-")
        u = t.lstrip("+-")
        if u[:2] in ("0x", "0X"):
            v = int(u, 16)
        elif len(u) > 1 and u[0] == "0":
            v = int(u, 8)
        else:
            v = int(u)
        return -v if neg else v

    def scan_d(self):
        t = self._match(_INT_D)
        return None if t is None else int(t)

    def scan_f(self):
        """fscanf "%f" into a float."""
        t = self._match(_FLOAT)
        return None if t is None else np.float32(float(t))

    def scan_s(self, maxlen: int = MAXIDSIZE):
        """fscanf "%64s": up to maxlen non-space characters."""
        self._ws()
        j = self.i
        while j < len(self.s) and self.s[j] not in _SPACE and j - self.i < maxlen:
            j += 1
        if j == self.i:
            return None
        t = self.s[self.i:j]
        self.i = j
        return t

    def scan_t_d(self):
        """fscanf " t%d": white space, a literal 't', then a decimal.  On a
        mismatch the white space (and a matched 't') stay consumed."""
        self._ws()
        if self.i < len(self.s) and self.s[self.i] == "t":
            self.i += 1
            return self.scan_d()
        return None

    def fail(self, msg: str):
        raise EDLUTFileException(msg, self.path, self.line)


# ---------------------------------------------------------------------------
# descriptions (ModelDescription, NeuronLayerDescription, SynapticLayerDescription)
# ---------------------------------------------------------------------------
@dataclass
class ModelDescription:
    model_name: str
    param_map: dict = field(default_factory=dict)


@dataclass
class NeuronLayerDescription:
    num_neurons: int
    neuron_model: ModelDescription
    output_activity: bool = False
    log_activity: bool = False


@dataclass
class SynapticLayerDescription:
    """One synaptic layer: parallel per-synapse arrays (Network.cpp:530-640)."""

    source: np.ndarray
    target: np.ndarray
    delay: np.ndarray          # float seconds, before RoundPropagationDelay
    type: np.ndarray
    weight: np.ndarray
    max_weight: np.ndarray
    wchange: np.ndarray        # learning rule index, -1 for none
    trigger_wchange: np.ndarray


# CPU time-driven models whose file format (NeuronModelFactory.cu:134-230,
# e.g. LIFTimeDrivenModel.cpp ParseNeuronModel) equals their GPU plug-in's:
# a network naming one runs it on the GPU plug-in, whose arithmetic is the
# CPU model's (the CPU Bifixed methods are a different algorithm and are
# refused under these names).
CPU_TIME_DRIVEN = {
    "LIFTimeDrivenModel": "LIFTimeDrivenModel_GPU",
    "AdExTimeDrivenModel": "AdExTimeDrivenModel_GPU",
    "HHTimeDrivenModel": "HHTimeDrivenModel_GPU",
    "IzhikevichTimeDrivenModel": "IzhikevichTimeDrivenModel_GPU",
    "LIFTimeDrivenModel_IS": "LIFTimeDrivenModel_IS_GPU",
    "EgidioGranuleCell_TimeDriven": "EgidioGranuleCell_TimeDriven_GPU",
    "TimeDrivenPurkinjeCell": "TimeDrivenPurkinjeCell_GPU",
}

# Input devices of the reference handled by the front-end.
INPUT_SPIKE = "InputSpikeNeuronModel"
INPUT_CURRENT = "InputCurrentNeuronModel"
POISSON = "PoissonGeneratorDeviceVector"


def parse_integration_method(f: _CFile) -> dict:
    """TimeDrivenNeuronModel_GPU_C_INTERFACE::ParseIntegrationMethod and the
    method's reader: "Euler"/"RK2"/"RK4" step; "BDF" order (1..6) then step;
    "Bifixed_*" step then n_steps (>= 1)."""
    from .models import _INTEGRATORS

    f.skip_comments()
    name = f.scan_s()
    if name is None:
        f.fail("the integration method cannot be read")
    if name not in _INTEGRATORS:
        f.fail(f"integration method {name!r} does not exist")
    out = {"name": name}
    if name == "BDF":
        f.skip_comments()
        order = f.scan_d()
        if order is None:
            f.fail("the BDF order cannot be read")
        if order < 1 or order > 6:
            f.fail("the BDF order must be between 1 and 6")
        out["bdf_order"] = order
    f.skip_comments()
    step = f.scan_f()
    if step is None:
        f.fail("the integration step cannot be read")
    out["step"] = float(step)
    if name.startswith("Bifixed_"):
        f.skip_comments()
        n = f.scan_d()
        if n is None:
            f.fail("the number of Bifixed sub-steps cannot be read")
        if n <= 0:
            f.fail("the number of Bifixed sub-steps must be positive")
        out["n_steps"] = n
    elif name != "BDF":
        out["bdf_order"] = 1
    return out


def parse_neuron_model(model_name: str, ident: str, base_dir: str | None = None) -> ModelDescription:
    """NeuronModelFactory::ParseNeuronModel: reads ``ident + '.cfg'`` (relative
    to base_dir, else the working directory, as fopen does) with the model's
    reader.  GPU models: their float parameters in GetDefaultParameters()
    order, each after skip_comments, then the integration method."""
    from .models import REGISTRY as MODEL_CLASSES

    if model_name == INPUT_SPIKE or model_name == INPUT_CURRENT:
        # InputSpikeNeuronModel.cpp:69-75: no file is read
        return ModelDescription(model_name, {"name": model_name})
    path = ident + ".cfg"
    if base_dir is not None and not os.path.isabs(path):
        path = os.path.join(base_dir, path)
    if model_name == POISSON:
        f = _CFile(path)
        f.skip_comments()
        v = f.scan_f()
        if v is None:
            f.fail("the Poisson generator frequency cannot be read")
        return ModelDescription(model_name, {"name": model_name, "frequency": float(v)})
    cls = MODEL_CLASSES.get(CPU_TIME_DRIVEN.get(model_name, model_name))
    if cls is None:
        raise EDLUTException(EDLUT_E_ARG, f"Neuron model \"{model_name}\" does not exist")
    f = _CFile(path)
    params = {"name": model_name}
    for pname, _ in cls.PARAMS:
        f.skip_comments()
        v = f.scan_f()
        if v is None:
            f.fail(f"parameter {pname!r} cannot be read")
        params[pname] = float(v)
    if "c_m" in params and not params["c_m"] > 0.0:
        f.fail("c_m must be positive")
    params["int_meth"] = parse_integration_method(f)
    return ModelDescription(model_name, params)


def parse_learning_rule(name: str, f: _CFile) -> ModelDescription:
    """LearningRuleFactory::ParseLearningRule for the rule built here:
    "STDP" max_LTP tau_LTP max_LTD tau_LTD (STDPWeightChange.cpp:77-98)."""
    if name != "STDP":
        f.fail(f"learning rule {name!r} is not available (only \"STDP\")")
    vals = []
    for _ in range(4):
        v = f.scan_f()
        if v is None:
            f.fail("the STDP parameters cannot be read")
        vals.append(float(v))
    if vals[1] <= 0 or vals[3] <= 0:
        f.fail("the STDP time constants must be positive")
    return ModelDescription("STDP", {"max_LTP": vals[0], "tau_LTP": vals[1], "max_LTD": vals[2],
                                     "tau_LTD": vals[3]})


def parse_net(netfile: str, wfile: str, base_dir: str | None = None, rng=None):
    """Network::ParseNet.  Returns (neuron_layer_list, learning_rule_list,
    synaptic_layer).  Weights below zero in the weights file draw
    frand() * max_weight from `rng` (a RandomGenerator, taking the next
    global seed; see simulation.GlobalSeed), above max_weight are clamped."""
    f = _CFile(netfile)
    f.skip_comments()
    n_types = f.scan_i()
    if n_types is None or n_types <= 0:
        f.fail("the number of neuron types cannot be read")
    f.skip_comments()
    n_neurons = f.scan_i()
    if n_neurons is None or n_neurons <= 0:
        f.fail("the number of neurons cannot be read")
    layers = []
    tind = 0
    while tind < n_neurons:
        f.skip_comments()
        nn = f.scan_i()
        if nn is None or nn <= 0:
            f.fail("the neuron layer cannot be read")
        ident_type = f.scan_s()
        ident = f.scan_s()
        outn = f.scan_i()
        monit = f.scan_i()
        if ident_type is None or ident is None or outn is None or monit is None:
            f.fail("the neuron layer cannot be read")
        if tind + nn > n_neurons:
            f.fail("the layers hold more neurons than declared")
        try:
            nmodel = parse_neuron_model(ident_type, ident, base_dir)
        except EDLUTFileException:
            raise
        except EDLUTException as e:
            f.fail(str(e))
        layers.append(NeuronLayerDescription(nn, nmodel, bool(outn), bool(monit)))
        tind += nn

    f.skip_comments()
    n_rules = f.scan_i()
    if n_rules is None or n_rules < 0:
        f.fail("the number of learning rules cannot be read")
    rules = []
    for _ in range(n_rules):
        f.skip_comments()
        name = f.scan_s()
        if name is None:
            f.fail("the learning rule cannot be read")
        rules.append(parse_learning_rule(name, f))

    f.skip_comments()
    n_inters = f.scan_i()
    if n_inters is None or n_inters <= 0:
        f.fail("the number of synapses cannot be read")
    src = np.zeros(n_inters, np.int32)
    tgt = np.zeros(n_inters, np.int32)
    delay = np.full(n_inters, np.float32(0.001), np.float32)
    typ = np.zeros(n_inters, np.int32)
    weight = np.ones(n_inters, np.float32)
    max_weight = np.ones(n_inters, np.float32)
    trig = np.full(n_inters, -1, np.int32)
    wch = np.full(n_inters, -1, np.int32)
    iind = 0
    while iind < n_inters:
        f.skip_comments()
        vals = [f.scan_i(), f.scan_i(), f.scan_i(), f.scan_i(), f.scan_i(), f.scan_f(), f.scan_f(),
                f.scan_i(), f.scan_f()]
        if any(v is None for v in vals):
            f.fail("the synapse line cannot be read")
        source, nsources, target, ntargets, nreps, d0, dinc, stype, maxw = vals
        r1 = f.scan_t_d()
        trigger1 = r1 is not None
        if not trigger1:
            r1 = f.scan_d()
            if r1 is None:
                f.fail("the first learning rule index cannot be read")
        if r1 >= len(rules):
            f.fail("the first learning rule index does not exist")
        r2, trigger2 = -1, False
        if r1 >= 0 and not f.is_end_line():
            r2 = f.scan_t_d()
            trigger2 = r2 is not None
            if not trigger2:
                r2 = f.scan_d()
                if r2 is None:
                    f.fail("the second learning rule index cannot be read")
            if r2 >= len(rules):
                f.fail("the second learning rule index does not exist")
        count = nsources * ntargets * nreps
        if iind + count > n_inters:
            f.fail("the synapse lines hold more synapses than declared")
        if not trigger1 and (r2 >= 0 and not trigger2):
            f.fail("a synapse cannot have two non-trigger learning rules")
        if trigger1 and trigger2:
            f.fail("a synapse cannot have two trigger learning rules")
        if count <= 0:
            f.fail("a synapse line must create at least one synapse")
        r = np.arange(nreps).repeat(nsources * ntargets)
        si = np.tile(np.arange(nsources).repeat(ntargets), nreps)
        ti = np.tile(np.arange(ntargets), nsources * nreps)
        sl = slice(iind, iind + count)
        src[sl] = source + r * nsources + si
        tgt[sl] = target + r * ntargets + ti
        # delay + delayinc * tind in float (Network.cpp:623)
        delay[sl] = np.float32(d0) + np.float32(dinc) * ti.astype(np.float32)
        typ[sl] = stype
        max_weight[sl] = maxw
        for rule, trigger in ((r1, trigger1), (r2, trigger2)):
            if rule >= 0:
                (trig if trigger else wch)[sl] = rule
        iind += count

    # weights file
    g = _CFile(wfile)
    g.skip_comments()
    conn = 0
    while conn < n_inters:
        g.skip_comments()
        nw = g.scan_i()
        wv = g.scan_f()
        if nw is None or wv is None:
            g.fail("the weights cannot be read")
        if nw < 0 or nw + conn > n_inters:
            g.fail("the weights file holds more weights than synapses")
        if nw == 0:
            nw = n_inters - conn
        for q in range(conn, conn + nw):
            if wv < 0.0:
                if rng is None:
                    g.fail("random weights need a random generator")
                weight[q] = np.float32(float(rng.frand()) * float(max_weight[q]))
            elif wv > max_weight[q]:
                weight[q] = max_weight[q]
            else:
                weight[q] = wv
        conn += nw

    syn = SynapticLayerDescription(src, tgt, delay, typ, weight, max_weight, wch, trig)
    return layers, rules, syn


def parse_input_spikes(path: str, n_neurons: int | None = None):
    """FileInputSpikeDriver::LoadInputs: a count, then lines
    ``time nspikes interval first_neuron nneurons``; spike j of a line fires
    at float(time + j * interval) on each of its neurons.  Returns
    (times float64, neuron indices int64) in file order."""
    f = _CFile(path)
    f.skip_comments()
    ninputs = f.scan_i()
    if ninputs is None:
        f.fail("the number of input spikes cannot be read")
    times, cells = [], []
    i = 0
    while i < ninputs:
        f.skip_comments()
        t = f.scan_f()
        ns = f.scan_i()
        iv = f.scan_f()
        nn = f.scan_i()
        nr = f.scan_i()
        if t is None or ns is None or iv is None or nn is None or nr is None:
            f.fail("fewer input spikes than declared")
        if nn < 0 or (n_neurons is not None and nn + nr > n_neurons):
            f.fail("input spike neuron index out of range")
        i += ns * nr
        if i > ninputs:
            f.fail("more input spikes than declared")
        for it in range(ns):
            tt = np.float32(t + np.float32(it) * iv)
            for r in range(nr):
                times.append(float(tt))
                cells.append(nn + r)
    return np.array(times, np.float64), np.array(cells, np.int64)


def parse_input_currents(path: str, n_neurons: int | None = None):
    """FileInputCurrentDriver::LoadInputs (FileInputCurrentDriver.cpp:45-90): a
    count, then lines ``time ncurrents interval first_neuron nneurons
    current``; change j of a line happens at float(time + j * interval) on
    each of its neurons.  Returns (times float64, neuron indices int64,
    currents float32) in file order."""
    f = _CFile(path)
    f.skip_comments()
    ninputs = f.scan_i()
    if ninputs is None:
        f.fail("the number of input currents cannot be read")
    times, cells, vals = [], [], []
    i = 0
    while i < ninputs:
        f.skip_comments()
        t = f.scan_f()
        ns = f.scan_i()
        iv = f.scan_f()
        nn = f.scan_i()
        nr = f.scan_i()
        cur = f.scan_f()
        if t is None or ns is None or iv is None or nn is None or nr is None or cur is None:
            f.fail("fewer input currents than declared")
        if nn < 0 or (n_neurons is not None and nn + nr > n_neurons):
            f.fail("input current neuron index out of range")
        i += ns * nr
        if i > ninputs:
            f.fail("more input currents than declared")
        for it in range(ns):
            tt = np.float32(t + np.float32(it) * iv)
            for r in range(nr):
                times.append(float(tt))
                cells.append(nn + r)
                vals.append(cur)
    return np.array(times, np.float64), np.array(cells, np.int64), np.array(vals, np.float32)
