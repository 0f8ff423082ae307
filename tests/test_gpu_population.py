"""GPU parity of one population (plug-in mode) against the oracle.

Every update's spike list must equal the oracle's exactly, and every state
variable must be bit-identical after the run (the HIP kernels keep the
reference CPU model's float association).  Inputs are pushed in host order
with many events per target, so the ordered fold is exercised.
"""
import zlib

import numpy as np
import pytest

import edlut_amd as E
from tests.helpers import ADEX_DEFAULTS, LIF_DEFAULTS, bits

pytestmark = pytest.mark.gpu

CASES = [
    ("AdExTimeDrivenModel_GPU", "Euler"),
    ("AdExTimeDrivenModel_GPU", "RK2"),
    ("AdExTimeDrivenModel_GPU", "RK4"),
    ("LIFTimeDrivenModel_GPU", "Euler"),
    ("LIFTimeDrivenModel_GPU", "RK2"),
    ("LIFTimeDrivenModel_GPU", "RK4"),
]


def _params(model):
    return ADEX_DEFAULTS if model.startswith("AdEx") else LIF_DEFAULTS


def _make(model, integ, n, types):
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, 1e-4)})
    for t in types:
        assert m.CheckSynapseType(t)
    m.InitializeStates(n, 0, 0)
    return m


@pytest.mark.parametrize("model,integ", CASES)
def test_population_bit_exact(oracle, model, integ):
    n, steps = 3001, 400
    print(f"case {model} {integ} n={n} steps={steps}", flush=True)
    rng = np.random.default_rng(zlib.crc32(f"{model}/{integ}".encode()))
    types = (0, 1, 2)
    m = _make(model, integ, n, types)
    o = oracle.Population(model, integ, _params(model), n, mask=7)
    wmax = 12.0 if model.startswith("AdEx") else 0.6
    total = 0
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, 6000))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        # a hot target receives many inputs in one step
        tgt[: ne // 10] = 7
        typ = rng.choice(np.array([0, 0, 0, 1, 2], np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        m.ProcessInputSpikes(tgt, typ, w)
        o.add_inputs(tgt, typ, w)
        m.UpdateState(step=k)
        ref = o.update(k)
        got = m.GetInternalSpikeIndexs()
        assert np.array_equal(got, ref), f"update {k}: spikes differ"
        total += len(ref)
    assert total > 0, "the test drive produced no spikes"
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"variable {v}"


@pytest.mark.parametrize("model", ["AdExTimeDrivenModel_GPU", "LIFTimeDrivenModel_GPU"])
def test_external_current_and_nan_reset(oracle, model):
    n, steps = 777, 150
    rng = np.random.default_rng(5)
    m = _make(model, "RK4", n, (0,))
    o = oracle.Population(model, "RK4", _params(model), n, mask=1 | 8)
    scale = 400.0 if model.startswith("AdEx") else 12.0
    cur = rng.uniform(-0.2, 1.0, n).astype(np.float32) * scale
    m.ProcessInputCurrent(cur)
    o.set_ext_current(cur)
    for k in range(1, steps + 1):
        if k == 40:
            v = m.GetVectorNeuronState(0)
            v[[3, 500]] = np.nan
            m.SetVectorNeuronState(0, v)
            ov = o.state(0)
            ov[[3, 500]] = np.nan
            o.set_state(0, ov)
        m.UpdateState(step=k)
        assert np.array_equal(m.GetInternalSpikeIndexs(), o.update(k)), f"update {k}"
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"variable {v}"


def test_no_inputs_rest_is_fixed_point(oracle):
    """Without input an AdEx neuron at e_leak stays there (derivative of the
    leak and adaptation terms vanish; the exponential term is below float
    resolution of the table's effect) -- compared with the oracle bit-exactly."""
    n = 300
    m = _make("AdExTimeDrivenModel_GPU", "RK4", n, ())
    o = oracle.Population("AdExTimeDrivenModel_GPU", "RK4", ADEX_DEFAULTS, n, mask=0)
    for k in range(1, 51):
        m.UpdateState(step=k)
        assert len(m.GetInternalSpikeIndexs()) == 0
        o.update(k)
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v)))


def test_errors_are_reported():
    m = E.AdExTimeDrivenModel_GPU({"int_meth": E.integration_method("RK4", 1e-4)})
    m.CheckSynapseType(0)
    m.InitializeStates(100, 0, 0)
    with pytest.raises(E.EDLUTException):
        m.ProcessInputSpikes(np.array([100]), np.array([0]), np.array([1.0]))  # target out of range
    with pytest.raises(E.EDLUTException):
        m.ProcessInputSpikes(np.array([1]), np.array([1]), np.array([1.0]))  # INH not enabled
    with pytest.raises(E.EDLUTException):
        m.UpdateState(step=0)


def test_lut_checksums_match_oracle_tables(oracle):
    """The tables uploaded to the device are the oracle's, byte for byte."""
    def fnv(b: bytes) -> int:
        h = 1469598103934665603
        for x in b:
            h ^= x
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        return h

    m = _make("AdExTimeDrivenModel_GPU", "RK4", 10, (0,))
    sums = m.LutChecksums()
    assert sums[0] == fnv(oracle.exp_table().tobytes())
    assert sums[1] == fnv(oracle.nmda_table(0.0, -80.0).tobytes())
