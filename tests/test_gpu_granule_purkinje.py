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


def _run(oracle, model, integ, order, n, steps, dt, wmax, cur_scale, types=(0, 0, 0, 1, 2), rate=0.5,
         nan_at=None):
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
    late = np.zeros(n, np.int64)   # spikes from the NaN update on
    for k in range(1, steps + 1):
        if nan_at and k == nan_at[0]:   # {variable: neurons} set to NaN before update k
            for var, idx in nan_at[1].items():
                v = m.GetVectorNeuronState(var)
                v[idx] = np.nan
                m.SetVectorNeuronState(var, v)
                ov = o.state(var)
                ov[idx] = np.nan
                o.set_state(var, ov)
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
        if nan_at and k >= nan_at[0]:
            late[ref] += 1
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"var {v}"
    if nan_at:
        return counts, late, [m.GetVectorNeuronState(v) for v in range(o.nv)]
    return counts


@pytest.mark.parametrize("order", [1, 2])
def test_granule_bdf_bit_exact(oracle, order):
    counts = _run(oracle, GR, "BDF", order, 300, 600, 1e-4, wmax=1.0, cur_scale=10.0)
    assert counts.sum() > 50 and (counts > 1).sum() > 10


def test_granule_bdf_standing_nan(oracle):
    """A granule cell whose BDF step goes NaN is put back to its initial
    state, but its BDF history keeps the NaN, so every later step is NaN
    again and resets it again: the neuron stays at its initial state and
    never fires.  NaN in V, in a gate and in the calcium concentration, set
    before update 150 of 400."""
    nan = {0: [3, 77], 5: [150], 14: [299]}
    counts, late, state = _run(oracle, GR, "BDF", 2, 300, 400, 1e-4, wmax=1.0, cur_scale=10.0,
                               nan_at=(150, nan))
    stuck = [3, 77, 150, 299]
    live = np.setdiff1d(np.arange(300), stuck)
    assert late[live].sum() > 50 and late[stuck].sum() == 0
    fresh = E.CreateNeuronModel(GR, {"int_meth": E.integration_method("BDF", 1e-4, 2)})
    fresh.InitializeStates(1, 0, 0)
    for v, x in enumerate(state):   # every step ends in the NaN reset
        assert np.array_equal(bits(x[stuck]), np.repeat(bits(fresh.GetVectorNeuronState(v)), 4)), f"var {v}"


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


@pytest.mark.parametrize("n,steps", [(1, 120), (70, 120), (140_000, 40)])
def test_granule_bdf_lane_refill_shapes(oracle, n, steps):
    """The granule BDF kernel splits the population into equal ranges, one
    per wave of at most one workgroup per CU, and lanes take a range's
    neurons as their correctors end.  One neuron (most waves empty), a
    population inside one wave's reach, and 140,000 cells: more 512-neuron
    tiles than CUs, so the ranges cross tiles and the last workgroup writes
    the spike lists of all 274 tiles.  Spike lists of every update and the
    final state against the oracle, under a drive that fires within tens of
    updates."""
    import os
    print(f"case granule BDF2 n={n} steps={steps}", flush=True)
    rng = np.random.default_rng(n)
    m = E.CreateNeuronModel(GR, {"int_meth": E.integration_method("BDF", 1e-4, 2)})
    assert m.CheckSynapseType(0)
    m.InitializeStates(n, 0, 0)
    o = oracle.Population(GR, "BDF", defaults(GR), n, dt_s=np.float32(1e-4), mask=1 | 8, bdf_order=2)
    cur = rng.uniform(0.0, 1.0, n).astype(np.float32) * 200.0
    m.ProcessInputCurrent(cur)
    o.set_ext_current(cur)
    oracle.set_threads(min(16, os.cpu_count() or 1))
    total = 0
    try:
        for k in range(1, steps + 1):
            ne = max(1, n // 4)
            tgt = rng.integers(0, n, ne).astype(np.int32)
            w = rng.uniform(0.0, 1.0, ne).astype(np.float32)
            typ = np.zeros(ne, np.uint8)
            m.ProcessInputSpikes(tgt, typ, w)
            o.add_inputs(tgt, typ, w)
            m.UpdateState(step=k)
            ref = o.update(k)
            assert np.array_equal(m.GetInternalSpikeIndexs(), ref), f"update {k}"
            total += len(ref)
    finally:
        oracle.set_threads(1)
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"var {v}"
    assert total > 0
