"""Host-side neuron-model plug-ins over the HIP library.

Each class mirrors one of EDLUT's GPU plug-ins (TimeDrivenNeuronModel_GPU_C_INTERFACE
subclasses, Kernel/include/neuron_model/TimeDrivenNeuronModel_GPU_C_INTERFACE.cuh:64-423):
same registry name, same parameter names and defaults, same integration-method
description ({"name", "step", "bdf_order"}; {"name", "step", "n_steps"} for the
Bifixed methods), and the same calls
(InitializeStates, CheckSynapseType, ProcessInputSpike, ProcessInputCurrent,
UpdateState).  UpdateState returns False like the reference; the spikes of the
step are read with GetInternalSpikeIndexs() (the reference exposes a bool array
through VectorNeuronState_GPU_C_INTERFACE, TimeDrivenInternalSpike.cpp:132-160;
here the device hands back the compacted ascending index list).
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict

import numpy as np

from . import _capi as C
from ._capi import EDLUTException

_INTEGRATORS = {"Euler": C.EDLUT_EULER, "RK2": C.EDLUT_RK2, "RK4": C.EDLUT_RK4, "BDF": C.EDLUT_BDF,
                "Bifixed_Euler": C.EDLUT_BIFIXED_EULER, "Bifixed_RK2": C.EDLUT_BIFIXED_RK2,
                "Bifixed_RK4": C.EDLUT_BIFIXED_RK4}


def integration_method(name: str = "Euler", step: float = 1e-3, bdf_order: int = 1, n_steps: int = 2) -> dict:
    """Integration-method description (IntegrationMethod.h GetDefaultParameters:
    step 0.001 s; BDFn_GPU_C_INTERFACE.cuh:275-288 for bdf_order; the Bifixed
    methods' n_steps, sub-steps per update, default 2,
    BifixedStep_GPU_C_INTERFACE.cuh:144-171)."""
    if name.startswith("Bifixed_"):
        return {"name": name, "step": float(step), "n_steps": int(n_steps)}
    return {"name": name, "step": float(step), "bdf_order": int(bdf_order)}


def _int_param(m: dict) -> int:
    """The integrator's integer parameter of the C ABI: n_steps or bdf_order."""
    return int(m["n_steps"]) if m["name"].startswith("Bifixed_") else int(m["bdf_order"])


class TimeDrivenNeuronModel_GPU:
    """Common part of the GPU time-driven plug-ins."""

    NAME = ""
    MODEL = -1
    # (name, default) in GetDefaultParameters() order
    PARAMS: tuple = ()
    STATE_VARS: tuple = ()
    N_DIFF = 0
    # synapse types: 0 AMPA, 1 GABA, 2 NMDA take spikes; 3 takes currents
    N_SYNAPSE_TYPES = 4

    def __init__(self, params: dict | None = None):
        self._p = OrderedDict(self.PARAMS)
        self._int = integration_method("Euler", 1e-3)
        self._pop = None
        self._n = 0
        self._mask = 0
        self._spikes = np.zeros(0, np.int32)
        self._dev = 0
        if params:
            self.SetParameters(params)

    # -- registry --------------------------------------------------------
    @classmethod
    def GetName(cls) -> str:
        return cls.NAME

    @classmethod
    def GetDefaultParameters(cls) -> dict:
        d = OrderedDict(cls.PARAMS)
        d["int_meth"] = integration_method("Euler", 1e-3)
        d["name"] = cls.NAME
        return d

    def GetParameters(self) -> dict:
        d = OrderedDict(self._p)
        d["int_meth"] = dict(self._int)
        d["name"] = self.NAME
        return d

    def SetParameters(self, params: dict) -> None:
        """Unknown keys raise EDLUTException like the reference's SetParameters."""
        if self._pop is not None:
            raise EDLUTException(C.EDLUT_E_STATE, "parameters are fixed once the states are initialised")
        for k, v in params.items():
            if k == "name":
                if v != self.NAME:
                    raise EDLUTException(C.EDLUT_E_ARG, f"model name {v!r} != {self.NAME!r}")
            elif k == "int_meth":
                m = dict(v)
                if m.get("name") not in _INTEGRATORS:
                    raise EDLUTException(C.EDLUT_E_ARG, f"unknown integration method {m.get('name')!r}")
                if not float(m.get("step", 0.0)) > 0.0:
                    raise EDLUTException(C.EDLUT_E_ARG, "integration step must be positive")
                if m["name"].startswith("Bifixed_") and int(m.get("n_steps", 2)) <= 0:
                    raise EDLUTException(C.EDLUT_E_ARG, "n_steps must be positive")
                self._int = integration_method(m["name"], m.get("step", 1e-3), m.get("bdf_order", 1),
                                               m.get("n_steps", 2))
            elif k in self._p:
                self._p[k] = float(np.float32(v))
            else:
                raise EDLUTException(C.EDLUT_E_ARG, f"unknown parameter {k!r} for {self.NAME}")

    # -- life cycle ------------------------------------------------------
    def InitializeStates(self, N_neurons: int, OpenMPQueueIndex: int = 0, device: int = 0) -> None:
        if self._pop is not None:
            raise EDLUTException(C.EDLUT_E_STATE, "states already initialised")
        params = np.array(list(self._p.values()), np.float32)
        h = ctypes.c_void_p()
        C.check(
            C.lib().edlut_population_create(
                self.MODEL,
                _INTEGRATORS[self._int["name"]],
                _int_param(self._int),
                C.ptr(params),
                len(params),
                float(np.float32(self._int["step"])),
                int(N_neurons),
                int(device),
                C.byref(h),
            )
        )
        self._pop = h
        self._n = int(N_neurons)
        self._dev = device
        if self._mask:
            C.check(C.lib().edlut_population_enable_synapses(self._pop, self._mask))

    def InitializeGroups(self, sizes, group_params=None, device: int = 0) -> None:
        """Grouped population: len(sizes) independent groups of this model in
        one index space (parameter sweeps).  group_params is a list of dicts
        overriding this model's parameters per group, or an array of shape
        (n_groups, n_params) in GetDefaultParameters() order.  Group g owns
        neurons [start[g], start[g] + sizes[g]) (GetGroups); the state arrays
        cover the whole index space, padding included."""
        if self._pop is not None:
            raise EDLUTException(C.EDLUT_E_STATE, "states already initialised")
        sizes = np.ascontiguousarray(sizes, np.int32)
        ng = len(sizes)
        base = np.array(list(self._p.values()), np.float32)
        if group_params is None:
            params = np.tile(base, (ng, 1))
        elif isinstance(group_params, np.ndarray):
            params = np.ascontiguousarray(group_params, np.float32)
            if params.shape != (ng, len(base)):
                raise EDLUTException(C.EDLUT_E_ARG, f"group parameters must have shape ({ng}, {len(base)})")
        else:
            if len(group_params) != ng:
                raise EDLUTException(C.EDLUT_E_ARG, "one parameter dict per group expected")
            names = list(self._p.keys())
            params = np.tile(base, (ng, 1))
            for g, d in enumerate(group_params):
                for k, v in d.items():
                    if k not in names:
                        raise EDLUTException(C.EDLUT_E_ARG, f"unknown parameter {k!r} for {self.NAME}")
                    params[g, names.index(k)] = np.float32(v)
        params = np.ascontiguousarray(params, np.float32)
        h = ctypes.c_void_p()
        C.check(
            C.lib().edlut_population_create_groups(
                self.MODEL, _INTEGRATORS[self._int["name"]], C.ptr(params), params.shape[1],
                float(np.float32(self._int["step"])), ng, C.ptr(sizes), int(device), C.byref(h)
            )
        )
        self._pop = h
        self._dev = device
        start = np.zeros(ng, np.int32)
        size = np.zeros(ng, np.int32)
        cnt = ctypes.c_int(0)
        C.check(C.lib().edlut_population_groups(self._pop, C.byref(cnt), C.ptr(start), C.ptr(size)))
        self._groups = (start, size)
        self._n = int(start[-1]) + ((int(size[-1]) + 255) // 256) * 256
        if self._mask:
            C.check(C.lib().edlut_population_enable_synapses(self._pop, self._mask))

    def GetGroups(self):
        """(start, size) arrays of the groups (one group for a plain population)."""
        self._need()
        return getattr(self, "_groups", (np.zeros(1, np.int32), np.array([self._n], np.int32)))

    def close(self) -> None:
        if self._pop is not None:
            C.lib().edlut_population_destroy(self._pop)
            self._pop = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _need(self):
        if self._pop is None:
            raise EDLUTException(C.EDLUT_E_STATE, "InitializeStates has not been called")

    @property
    def handle(self):
        self._need()
        return self._pop

    @property
    def size(self) -> int:
        return self._n

    # -- synapses --------------------------------------------------------
    def CheckSynapseType(self, synapse_type: int, source_emits_spikes: bool = True) -> bool:
        """Enable a synapse type (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:239-290):
        types 0-2 need spiking sources, type 3 a current source."""
        t = int(synapse_type)
        if t < 0 or t >= self.N_SYNAPSE_TYPES:
            return False
        if (t < 3) != bool(source_emits_spikes):
            return False
        self._mask |= 1 << t
        if self._pop is not None:
            C.check(C.lib().edlut_population_enable_synapses(self._pop, 1 << t))
        return True

    def ProcessInputSpike(self, target: int, synapse_type: int, weight: float) -> None:
        self.ProcessInputSpikes(np.array([target]), np.array([synapse_type]), np.array([weight]))

    def ProcessInputSpikes(self, targets, types, weights) -> None:
        """Queue input spikes for the next update, in delivery order."""
        self._need()
        t = np.ascontiguousarray(targets, np.int32)
        y = np.ascontiguousarray(types, np.uint8)
        w = np.ascontiguousarray(weights, np.float32)
        if not (len(t) == len(y) == len(w)):
            raise EDLUTException(C.EDLUT_E_ARG, "event arrays differ in length")
        C.check(C.lib().edlut_population_push_inputs(self._pop, C.ptr(t), C.ptr(y), C.ptr(w), len(t)))

    def ProcessInputCurrent(self, total_current) -> None:
        """Total external current (pA) per neuron (the reference sums the
        per-synapse currents with CurrentSynapseModel::GetTotalCurrent)."""
        self._need()
        c = np.ascontiguousarray(total_current, np.float32)
        if c.shape != (self._n,):
            raise EDLUTException(C.EDLUT_E_ARG, "one current per neuron expected")
        self._mask |= C.SYN_EXT_I
        C.check(C.lib().edlut_population_set_ext_current(self._pop, C.ptr(c)))

    # -- stepping --------------------------------------------------------
    def UpdateState(self, index: int = -1, CurrentTime: float = 0.0, step: int | None = None) -> bool:
        """Advance all neurons one step.  `step` is the update index (1 for the
        first update); by default it is derived from CurrentTime / dt."""
        self._need()
        if step is None:
            step = int(round(CurrentTime / self._int["step"]))
        n = ctypes.c_int(0)
        C.check(C.lib().edlut_population_update(self._pop, int(step), C.byref(n)))
        out = np.zeros(max(n.value, 1), np.int32)
        C.check(C.lib().edlut_population_get_spikes(self._pop, C.ptr(out), len(out)))
        self._spikes = out[: n.value].copy()
        return False

    def GetInternalSpikeIndexs(self) -> np.ndarray:
        return self._spikes

    def GetVectorNeuronState(self, var: int | str) -> np.ndarray:
        self._need()
        if isinstance(var, str):
            var = self.STATE_VARS.index(var)
        out = np.zeros(self._n, np.float32)
        C.check(C.lib().edlut_population_get_state(self._pop, int(var), C.ptr(out)))
        return out

    def SetVectorNeuronState(self, var: int | str, values) -> None:
        self._need()
        if isinstance(var, str):
            var = self.STATE_VARS.index(var)
        v = np.ascontiguousarray(values, np.float32)
        if v.shape != (self._n,):
            raise EDLUTException(C.EDLUT_E_ARG, "one value per neuron expected")
        C.check(C.lib().edlut_population_set_state(self._pop, int(var), C.ptr(v)))

    def LutChecksums(self) -> tuple:
        self._need()
        out = np.zeros(3, np.uint64)
        C.check(C.lib().edlut_population_lut_checksums(self._pop, C.ptr(out)))
        return tuple(int(x) for x in out)

    def GetModelSimulationMethod(self) -> str:
        return "TIME_DRIVEN_MODEL_GPU"

    def GetModelType(self) -> str:
        return "NEURAL_LAYER"

    def GetModelOutputActivityType(self) -> str:
        return "OUTPUT_SPIKE"

    def GetModelInputActivityType(self) -> str:
        return "INPUT_SPIKE_AND_CURRENT"


class LIFTimeDrivenModel_GPU(TimeDrivenNeuronModel_GPU):
    """Leaky integrate-and-fire, conductance based
    (Kernel/src/neuron_model/LIFTimeDrivenModel_GPU_C_INTERFACE.cu; defaults of
    LIFTimeDrivenModel.cpp:435-450)."""

    NAME = "LIFTimeDrivenModel_GPU"
    MODEL = C.EDLUT_LIF
    PARAMS = (
        ("e_exc", 0.0),
        ("e_inh", -80.0),
        ("e_leak", -65.0),
        ("v_thr", -50.0),
        ("c_m", 2.0),
        ("tau_exc", 5.0),
        ("tau_inh", 10.0),
        ("tau_ref", 1.0),
        ("g_leak", 0.2),
        ("tau_nmda", 20.0),
    )
    STATE_VARS = ("V_m", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 1


class AdExTimeDrivenModel_GPU(TimeDrivenNeuronModel_GPU):
    """Adaptive exponential integrate-and-fire
    (Kernel/src/neuron_model/AdExTimeDrivenModel_GPU_C_INTERFACE.cu:44-593;
    defaults of AdExTimeDrivenModel.cpp:449-467)."""

    NAME = "AdExTimeDrivenModel_GPU"
    MODEL = C.EDLUT_ADEX
    PARAMS = (
        ("a", 1.0),
        ("b", 9.0),
        ("thr_slo_fac", 2.0),
        ("v_thr", -50.0),
        ("tau_w", 50.0),
        ("e_exc", 0.0),
        ("e_inh", -80.0),
        ("e_reset", -80.0),
        ("e_leak", -65.0),
        ("g_leak", 10.0),
        ("c_m", 110.0),
        ("tau_exc", 5.0),
        ("tau_inh", 10.0),
        ("tau_nmda", 20.0),
    )
    STATE_VARS = ("V_m", "w", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 2


class HHTimeDrivenModel_GPU(TimeDrivenNeuronModel_GPU):
    """Hodgkin-Huxley (Kernel/src/neuron_model/HHTimeDrivenModel_GPU_C_INTERFACE.cu;
    defaults of HHTimeDrivenModel.cpp GetDefaultParameters).  Under BDF the
    Newton step can run in FP64 (SetNewtonFP64)."""

    NAME = "HHTimeDrivenModel_GPU"
    MODEL = C.EDLUT_HH
    PARAMS = (
        ("e_exc", 0.0),
        ("e_inh", -80.0),
        ("e_leak", -65.0),
        ("g_leak", 10.0),
        ("c_m", 120.0),
        ("v_thr", -52.0),
        ("tau_exc", 5.0),
        ("tau_inh", 10.0),
        ("tau_nmda", 20.0),
        ("g_na", 20000.0),
        ("g_kd", 6000.0),
        ("e_na", 50.0),
        ("e_k", -90.0),
    )
    STATE_VARS = ("V_m", "m", "h", "n", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 4

    def SetNewtonFP64(self, on: bool = True) -> None:
        self._need()
        C.check(C.lib().edlut_population_set_newton_fp64(self._pop, int(bool(on))))


class EgidioGranuleCell_TimeDriven_GPU(TimeDrivenNeuronModel_GPU):
    """Cerebellar granule cell with fifteen differential equations
    (Kernel/src/neuron_model/EgidioGranuleCell_TimeDriven_GPU_C_INTERFACE.cu;
    CPU model EgidioGranuleCell_TimeDriven.cpp).  Its constants are fixed: the
    model has no parameters besides the integration method."""

    NAME = "EgidioGranuleCell_TimeDriven_GPU"
    MODEL = C.EDLUT_GRANULE
    PARAMS = ()
    STATE_VARS = ("V_m", "xNa_f", "yNa_f", "xNa_r", "yNa_r", "xNa_p", "xK_V", "xK_A", "yK_A", "xK_IR",
                  "xK_Ca", "xCa", "yCa", "xK_sl", "Ca", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 15


class TimeDrivenPurkinjeCell_GPU(TimeDrivenNeuronModel_GPU):
    """Purkinje cell with calcium and muscarinic currents and a stereotyped
    spike waveform (Kernel/src/neuron_model/TimeDrivenPurkinjeCell_GPU_C_INTERFACE.cu;
    defaults of TimeDrivenPurkinjeCell.cpp:423-434)."""

    NAME = "TimeDrivenPurkinjeCell_GPU"
    MODEL = C.EDLUT_PURKINJE
    PARAMS = (
        ("e_exc", 0.0),
        ("e_inh", -80.0),
        ("v_thr", -35.0),
        ("e_leak", -70.0),
        ("tau_exc", 1.0),
        ("tau_inh", 2.0),
        ("tau_nmda", 20.0),
        ("tau_ref", 1.35),
    )
    STATE_VARS = ("V_m", "Ca", "M", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 3


class IzhikevichTimeDrivenModel_GPU(TimeDrivenNeuronModel_GPU):
    """Izhikevich's simple model with conductance-based synapses
    (Kernel/src/neuron_model/IzhikevichTimeDrivenModel_GPU_C_INTERFACE.cu;
    defaults of IzhikevichTimeDrivenModel.cpp:389-401)."""

    NAME = "IzhikevichTimeDrivenModel_GPU"
    MODEL = C.EDLUT_IZHIKEVICH
    PARAMS = (
        ("a", 0.1),
        ("b", 0.23),
        ("c", -65.0),
        ("d", 0.2),
        ("e_exc", 0.0),
        ("e_inh", -80.0),
        ("c_m", 100.0),
        ("tau_exc", 5.0),
        ("tau_inh", 10.0),
        ("tau_nmda", 20.0),
    )
    STATE_VARS = ("V_m", "u", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 2


class LIFTimeDrivenModel_IS_GPU(TimeDrivenNeuronModel_GPU):
    """Leaky integrate-and-fire in SI units: volts, farads, siemens, seconds
    (Kernel/src/neuron_model/LIFTimeDrivenModel_IS_GPU_C_INTERFACE.cu;
    defaults of LIFTimeDrivenModel_IS.cpp:392-404)."""

    NAME = "LIFTimeDrivenModel_IS_GPU"
    MODEL = C.EDLUT_LIF_IS
    PARAMS = (
        ("e_exc", 0.0),
        ("e_inh", -80.0e-3),
        ("e_leak", -65.0e-3),
        ("v_thr", -50.0e-3),
        ("c_m", 2.0e-12),
        ("tau_exc", 5.0e-3),
        ("tau_inh", 10.0e-3),
        ("tau_ref", 1.0e-3),
        ("g_leak", 0.2e-9),
        ("tau_nmda", 20.0e-3),
    )
    STATE_VARS = ("V_m", "g_exc", "g_inh", "g_nmda", "ext_I")
    N_DIFF = 1


# NeuronModelFactory registry (Kernel/src/neuron_model/NeuronModelFactory.cu:66-91)
REGISTRY = {
    LIFTimeDrivenModel_GPU.NAME: LIFTimeDrivenModel_GPU,
    AdExTimeDrivenModel_GPU.NAME: AdExTimeDrivenModel_GPU,
    HHTimeDrivenModel_GPU.NAME: HHTimeDrivenModel_GPU,
    EgidioGranuleCell_TimeDriven_GPU.NAME: EgidioGranuleCell_TimeDriven_GPU,
    TimeDrivenPurkinjeCell_GPU.NAME: TimeDrivenPurkinjeCell_GPU,
    IzhikevichTimeDrivenModel_GPU.NAME: IzhikevichTimeDrivenModel_GPU,
    LIFTimeDrivenModel_IS_GPU.NAME: LIFTimeDrivenModel_IS_GPU,
}


def CreateNeuronModel(name: str, params: dict | None = None) -> TimeDrivenNeuronModel_GPU:
    if name not in REGISTRY:
        raise EDLUTException(C.EDLUT_E_ARG, f"unknown neuron model {name!r}")
    return REGISTRY[name](params)
