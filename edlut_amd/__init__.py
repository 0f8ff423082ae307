"""MI355X (gfx950) implementation of EDLUT's time-driven neuron-update path.

The compute path is the HIP library edlut_amd/libedlut_amd.so (C ABI in
include/edlut_amd.h); this package is its host-side Python interface.
"""
from . import _capi
from ._capi import EDLUTException, lib
from .models import (
    AdExTimeDrivenModel_GPU,
    CreateNeuronModel,
    EgidioGranuleCell_TimeDriven_GPU,
    HHTimeDrivenModel_GPU,
    IzhikevichTimeDrivenModel_GPU,
    LIFTimeDrivenModel_GPU,
    LIFTimeDrivenModel_IS_GPU,
    REGISTRY,
    TimeDrivenNeuronModel_GPU,
    TimeDrivenPurkinjeCell_GPU,
    integration_method,
)
from .network import Circuit, Network, binomial_thresholds, build_out_csr, from_spec
from .simulation import RandomGenerator, Simulation_API

__all__ = [
    "EDLUTException",
    "lib",
    "AdExTimeDrivenModel_GPU",
    "LIFTimeDrivenModel_GPU",
    "HHTimeDrivenModel_GPU",
    "EgidioGranuleCell_TimeDriven_GPU",
    "TimeDrivenPurkinjeCell_GPU",
    "IzhikevichTimeDrivenModel_GPU",
    "LIFTimeDrivenModel_IS_GPU",
    "TimeDrivenNeuronModel_GPU",
    "CreateNeuronModel",
    "REGISTRY",
    "integration_method",
    "Network",
    "Circuit",
    "binomial_thresholds",
    "build_out_csr",
    "from_spec",
    "Simulation_API",
    "RandomGenerator",
]
