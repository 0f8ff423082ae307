"""GPU parity of the Egidio granule cell and the Purkinje cell against the oracle.

The granule cell (fifteen differential equations) runs under BDF orders 1-2
at the 0.1 ms step of the cerebellar network, and under Euler, RK2 and RK4 at
a step short enough for explicit methods (its fastest gate has a time
constant of a few microseconds).  The Purkinje cell runs under every
integrator at 0.1 ms; its spikes start a stereotyped waveform timed by the
refractory clock.  Every update's spike list and the final state must be
bit-identical to the oracle.
"""
import zlib

import numpy as np
import pytest

import edlut_amd as E
from tests.helpers import bits, defaults

pytestmark = pytest.mark.gpu

GR = "EgidioGranuleCell_TimeDriven_GPU"
PK = "TimeDrivenPurkinjeCell_GPU"


def _run(oracle, model, integ, order, n, steps, dt, wmax, cur_scale, types=(0, 0, 0, 1, 2), rate=0.5):
    print(f"case {model} {integ} order={order} dt={dt} n={n} steps={steps}", flush=True)
    rng = np.random.default_rng(zlib.crc32(f"{model}/{integ}/{order}/{dt}".encode()))
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, dt, order)})
    mask = 8
    for t in sorted(set(types)):
        assert m.CheckSynapseType(t)
        mask |= 1 << t
    m.InitializeStates(n, 0, 0)
    o = oracle.Population(model, integ, defaults(model), n, dt_s=np.float32(dt), mask=mask, bdf_order=order)
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"initial var {v}"
    cur = rng.uniform(0.0, 1.0, n).astype(np.float32) * cur_scale
    m.ProcessInputCurrent(cur)
    o.set_ext_current(cur)
    counts = np.zeros(n, np.int64)
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, int(n * rate) + 1))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        tgt[: ne // 8] = 3
        typ = rng.choice(np.array(types, np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        m.ProcessInputSpikes(tgt, typ, w)
        o.add_inputs(tgt, typ, w)
        m.UpdateState(step=k)
        ref = o.update(k)
        assert np.array_equal(m.GetInternalSpikeIndexs(), ref), f"update {k}"
        counts[ref] += 1
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"var {v}"
    return counts


@pytest.mark.parametrize("order", [1, 2])
def test_granule_bdf_bit_exact(oracle, order):
    counts = _run(oracle, GR, "BDF", order, 300, 600, 1e-4, wmax=1.0, cur_scale=10.0)
    assert counts.sum() > 50 and (counts > 1).sum() > 10


@pytest.mark.parametrize("integ", ["Euler", "RK2", "RK4"])
def test_granule_explicit_bit_exact(oracle, integ):
    counts = _run(oracle, GR, integ, 1, 515, 4000, 2e-6, wmax=2.0, cur_scale=30.0, rate=0.1)
    assert counts.sum() > 50


@pytest.mark.parametrize("integ,order", [("Euler", 1), ("RK2", 1), ("RK4", 1), ("BDF", 2)])
def test_purkinje_bit_exact(oracle, integ, order):
    counts = _run(oracle, PK, integ, order, 1025, 800, 1e-4, wmax=1.0, cur_scale=300.0)
    assert counts.sum() > 200 and (counts > 1).sum() > 100


def test_granule_synapse_subsets(oracle):
    """The granule kernel takes the enabled synapse types at run time: an
    excitatory-only population equals the oracle's with the same mask."""
    counts = _run(oracle, GR, "BDF", 2, 200, 300, 1e-4, wmax=1.0, cur_scale=5.0, types=(0,))
    assert counts.sum() > 0
