"""ctypes binding of the C ABI declared in include/edlut_amd.h.

The shared library is built in-tree (edlut_amd/libedlut_amd.so, see
edlut_amd/build.py).  There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedlut_amd.so")

# Every symbol include/edlut_amd.h declares (checked by the CPU test suite).
EXPORTS = (
    "edlut_last_error",
    "edlut_population_create",
    "edlut_population_create_groups",
    "edlut_population_groups",
    "edlut_population_destroy",
    "edlut_population_layout",
    "edlut_population_enable_synapses",
    "edlut_population_push_inputs",
    "edlut_population_set_ext_current",
    "edlut_population_update",
    "edlut_population_get_spikes",
    "edlut_population_update_noalloc",
    "edlut_population_get_state",
    "edlut_population_set_state",
    "edlut_population_set_newton_fp64",
    "edlut_population_device_state",
    "edlut_population_lut_checksums",
    "edlut_network_create",
    "edlut_network_destroy",
    "edlut_network_add_recurrent",
    "edlut_network_add_poisson",
    "edlut_network_set_external",
    "edlut_network_set_poisson_rates",
    "edlut_network_add_spike_train",
    "edlut_network_set_spike_train",
    "edlut_network_add_projection",
    "edlut_network_set_source_model",
    "edlut_network_set_model_rank",
    "edlut_network_set_event_order",
    "edlut_network_set_stop_time",
    "edlut_network_set_spike_train_times",
    "edlut_circuit_set_stop_time",
    "edlut_network_event_queue",
    "edlut_circuit_event_queue",
    "edlut_circuit_create",
    "edlut_circuit_run",
    "edlut_circuit_synchronize",
    "edlut_circuit_destroy",
    "edlut_network_run",
    "edlut_network_set_graph",
    "edlut_network_synchronize",
    "edlut_network_spike_log",
    "edlut_network_clear_log",
    "edlut_network_set_source_log",
    "edlut_network_source_spike_log",
    "edlut_network_set_log_capacity",
    "edlut_network_update_dispatch_timing",
    "edlut_network_spike_count",
    "edlut_network_set_timing",
    "edlut_network_update_timing",
    "edlut_network_generate_recurrent",
    "edlut_network_get_synapses",
    "edlut_network_reset",
    "edlut_network_set_stdp",
    "edlut_network_get_stdp",
    "edlut_network_set_launch",
    "edlut_network_get_launch",
    "edlut_network_autotune",
    "edlut_fit_create",
    "edlut_fit_destroy",
    "edlut_fit_forward",
    "edlut_fit_backward",
    "edlut_refractory_table",
    "edlut_device_arch",
    "edlut_drive_hash",
    "edlut_build_out_csr",
)

EDLUT_LIF, EDLUT_ADEX, EDLUT_HH, EDLUT_GRANULE, EDLUT_PURKINJE, EDLUT_IZHIKEVICH, EDLUT_LIF_IS = range(7)
EDLUT_EULER, EDLUT_RK2, EDLUT_RK4, EDLUT_BDF, EDLUT_BIFIXED_EULER, EDLUT_BIFIXED_RK2, EDLUT_BIFIXED_RK4 = range(7)
SYN_EXC, SYN_INH, SYN_NMDA, SYN_EXT_I = 1, 2, 4, 8
EDLUT_E_ARG, EDLUT_E_HIP, EDLUT_E_CAP, EDLUT_E_STATE = -1, -2, -3, -4


class EDLUTException(RuntimeError):
    """Error raised for a failing library call (the reference throws
    EDLUTException for configuration errors, Kernel/src/spike/EDLUTException.cpp)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
F32 = ctypes.c_float


def _sig(lib):
    s = {
        "edlut_last_error": (ctypes.c_char_p, []),
        "edlut_device_arch": (ctypes.c_char_p, []),
        "edlut_drive_hash": (U32, [U64, I64, U32]),
        "edlut_population_create": (I32, [I32, I32, I32, P, I32, F32, I32, I32, P]),
        "edlut_population_create_groups": (I32, [I32, I32, P, I32, F32, I32, P, I32, P]),
        "edlut_population_groups": (I32, [P, P, P, P]),
        "edlut_population_destroy": (I32, [P]),
        "edlut_population_layout": (I32, [P, P, P, P]),
        "edlut_population_enable_synapses": (I32, [P, U32]),
        "edlut_population_push_inputs": (I32, [P, P, P, P, I32]),
        "edlut_population_set_ext_current": (I32, [P, P]),
        "edlut_population_update": (I32, [P, I64, P]),
        "edlut_population_get_spikes": (I32, [P, P, I32]),
        "edlut_population_update_noalloc": (I32, [P, I64, P, P, P, I32, P, P, P]),
        "edlut_population_get_state": (I32, [P, I32, P]),
        "edlut_population_set_state": (I32, [P, I32, P]),
        "edlut_population_set_newton_fp64": (I32, [P, I32]),
        "edlut_population_device_state": (I32, [P, P]),
        "edlut_population_lut_checksums": (I32, [P, P]),
        "edlut_network_create": (I32, [P, I32, P]),
        "edlut_network_destroy": (I32, [P]),
        "edlut_network_add_recurrent": (I32, [P, P, P, P]),
        "edlut_network_add_poisson": (I32, [P, I32, P, F32, P, P, P]),
        "edlut_network_set_external": (I32, [P, I32, P, F32, U64]),
        "edlut_network_set_poisson_rates": (I32, [P, I32, P]),
        "edlut_network_add_spike_train": (I32, [P, I32, I64, P, P, P, P, P]),
        "edlut_network_set_spike_train": (I32, [P, I32, I64, P, P]),
        "edlut_network_add_projection": (I32, [P, P, P, P, P]),
        "edlut_network_set_source_model": (I32, [P, I32, I32, I32]),
        "edlut_network_set_model_rank": (I32, [P, I32]),
        "edlut_network_set_event_order": (I32, [P, I32]),
        "edlut_network_set_stop_time": (I32, [P, ctypes.c_double]),
        "edlut_network_set_spike_train_times": (I32, [P, I32, I64, P, P, P]),
        "edlut_circuit_set_stop_time": (I32, [P, ctypes.c_double]),
        "edlut_network_event_queue": (I32, [P, P, P, I64, P]),
        "edlut_circuit_event_queue": (I32, [P, P, P, I64, P]),
        "edlut_circuit_create": (I32, [P, I32, P]),
        "edlut_circuit_run": (I32, [P, I64, I32]),
        "edlut_circuit_synchronize": (I32, [P]),
        "edlut_circuit_destroy": (I32, [P]),
        "edlut_network_run": (I32, [P, I64, I32]),
        "edlut_network_set_graph": (I32, [P, I32]),
        "edlut_network_synchronize": (I32, [P]),
        "edlut_network_spike_log": (I32, [P, P, P, I64, P]),
        "edlut_network_clear_log": (I32, [P]),
        "edlut_network_set_source_log": (I32, [P, I32, I64]),
        "edlut_network_source_spike_log": (I32, [P, I32, P, P, I64, P]),
        "edlut_network_set_log_capacity": (I32, [P, I64]),
        "edlut_network_update_dispatch_timing": (I32, [P, P, P]),
        "edlut_network_spike_count": (I32, [P, P]),
        "edlut_network_set_timing": (I32, [P, I64]),
        "edlut_network_update_timing": (I32, [P, P, P]),
        "edlut_network_generate_recurrent": (I32, [P, I32, F32, F32, F32, U64]),
        "edlut_network_get_synapses": (I32, [P, I32, P, P, P, P, P]),
        "edlut_network_reset": (I32, [P]),
        "edlut_network_set_stdp": (I32, [P, I32, F32, F32, F32, F32, F32]),
        "edlut_network_get_stdp": (I32, [P, I32, P, P, P]),
        "edlut_network_set_launch": (I32, [P, I32, I32, I32]),
        "edlut_network_get_launch": (I32, [P, P, P, P]),
        "edlut_network_autotune": (I32, [P, I32, P, P, P, P]),
        "edlut_fit_create": (I32, [I32, P, I32, F32, I32, I32, I32, I32, I32, P]),
        "edlut_fit_destroy": (I32, [P]),
        "edlut_fit_forward": (I32, [P, P, P]),
        "edlut_fit_backward": (I32, [P, P, P]),
        "edlut_refractory_table": (I32, [F32, F32, P, I32, P]),
        "edlut_build_out_csr": (I32, [I64, I64, I32, I64, P, P, P, P, P, P, P, P]),
    }
    for name, (res, args) in s.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        # PyTorch ships its own HIP runtime under the same soname.  Loading it
        # first makes this library bind to that one runtime; loaded the other
        # way round, torch finds no GPU once this library has initialised the
        # system runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        _sig(l)
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().edlut_last_error().decode(errors="replace")
        raise EDLUTException(rc, msg)


def ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def byref(x):
    return ctypes.byref(x)
