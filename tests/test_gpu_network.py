"""GPU parity of the device-resident network mode against the oracle.

The oracle folds every delivered event sequentially in the reference's
delivery order; the GPU must produce the same spike raster (step and index
of every spike) and bit-identical final state.
"""
import numpy as np
import pytest

import edlut_amd as E
from edlut_amd import workloads
from tests.helpers import ADEX_DEFAULTS, LIF_DEFAULTS, bits, np_out_csr

pytestmark = pytest.mark.gpu


def _run_both(oracle, spec, params, steps, chunks=1):
    """The same runs on both sides (each run is one RunSimulationSlot: its
    stop event is part of the event queue)."""
    model, net = E.from_spec(spec)
    pop, onet = oracle.from_spec(spec, params, np_out_csr)
    per = steps // chunks
    ost, oix = [], []
    for c in range(chunks):
        net.run(per)
        a, b = onet.run(1 + c * per, per)
        ost.append(a)
        oix.append(b)
    st, ix = net.spikes()
    return model, net, pop, (st, ix), (np.concatenate(ost), np.concatenate(oix))


def _keys(oracle_queue):
    t, prio, kind, em, d = oracle_queue
    return t, ((prio.astype(np.uint32) << 28) | (kind.astype(np.uint32) << 24) | (em.astype(np.uint32) << 16)
               | d.astype(np.uint32))


def test_event_queue_replay_equals_oracle(oracle):
    """The device's replay of the reference's event queue holds the same heap
    (times, priorities, kinds, models, delays, in heap-array order) as the
    oracle's event loop after every run, through bursts of more than 1,024
    spikes per update (groups split and tied) and runs of uneven lengths."""
    spec = workloads.adex_recurrent(n=6000, steps=700, seed=3, w_ext=1.0)
    print("case event queue replay n=6000", flush=True)
    model, net = E.from_spec(spec)
    pop, onet = oracle.from_spec(spec, ADEX_DEFAULTS, np_out_csr)
    k = 1
    burst = 0
    for n in (1, 37, 100, 16, 200, 3, 250, 93):
        net.run(n)
        ost, oix = onet.run(k, n)
        k += n
        t_d, key_d = net.event_queue()
        t_o, key_o = _keys(onet.event_queue())
        assert len(t_d) == len(t_o), f"queue sizes differ after update {k - 1}"
        assert np.array_equal(t_d.view(np.uint64), t_o.view(np.uint64)), f"event times differ after update {k - 1}"
        assert np.array_equal(key_d, key_o), f"event keys differ after update {k - 1}"
        if len(ost):
            burst = max(burst, int(np.bincount(ost).max()))
    assert burst > 1024, "the workload should burst past one group"
    for v in range(pop.nv):
        assert np.array_equal(bits(model.GetVectorNeuronState(v)), bits(pop.state(v))), f"variable {v}"


def test_lif_poisson_config(oracle):
    """1,000 LIF neurons, 125,000 Poisson sources at 10 Hz, Euler, 1 s."""
    spec = workloads.lif_poisson(n=1000, steps=10000)
    print("case lif_poisson n=1000 steps=10000", flush=True)
    model, net, pop, (st, ix), (ost, oix) = _run_both(oracle, spec, LIF_DEFAULTS, 10000, chunks=4)
    assert len(ost) > 1000, "the workload should spike"
    assert np.array_equal(st, ost) and np.array_equal(ix, oix)
    for v in range(pop.nv):
        assert np.array_equal(bits(model.GetVectorNeuronState(v)), bits(pop.state(v))), f"variable {v}"


@pytest.mark.parametrize("n,steps,w_ext", [(3000, 1500, 1.0), (20000, 3000, 0.7)])
def test_adex_recurrent_scaled(oracle, n, steps, w_ext):
    """AdEx recurrent network (in-degree 200, delays 1-10 steps, external drive), RK4.
    w_ext 1.0 drives the network into its saturated state (dense delivery,
    tens of events per target per step); 0.7 stays sparse for ~2,000 steps
    and then ignites."""
    spec = workloads.adex_recurrent(n=n, steps=steps, seed=3, w_ext=w_ext)
    print(f"case adex_recurrent n={n} steps={steps} w_ext={w_ext}", flush=True)
    model, net, pop, (st, ix), (ost, oix) = _run_both(oracle, spec, ADEX_DEFAULTS, steps, chunks=2)
    assert len(ost) > 100, "the workload should spike"
    assert np.array_equal(st, ost) and np.array_equal(ix, oix)
    for v in range(pop.nv):
        assert np.array_equal(bits(model.GetVectorNeuronState(v)), bits(pop.state(v))), f"variable {v}"


def test_adex_full_size_properties():
    """1,000,000-neuron AdEx network, 300 steps: size-independent properties
    (ascending indices within each step, steps ascending, indices in range,
    reset values of the neurons that fired on the last step)."""
    spec = workloads.adex_recurrent(n=1_000_000, steps=300, seed=2)
    print("case adex_recurrent n=1e6 steps=300", flush=True)
    model, net = E.from_spec(spec)
    net.run(300)
    st, ix = net.spikes()
    assert np.all(np.diff(st) >= 0)
    same = st[1:] == st[:-1]
    assert np.all(np.diff(ix)[same] > 0)
    assert ix.min(initial=0) >= 0 and ix.max(initial=0) < spec.n
    last = ix[st == 300]
    if len(last):
        v = model.GetVectorNeuronState(0)
        assert np.all(v[last] == np.float32(-80.0))


def test_graph_replay_equals_direct_launches():
    """16-step hipGraph replays (update index read from device memory) give the
    same raster and state as launching every kernel directly; runs split at
    odd lengths exercise the graph/remainder boundaries."""
    spec = workloads.adex_recurrent(n=4000, steps=0, seed=5, w_ext=1.0)
    print("case graph vs direct n=4000", flush=True)
    res = []
    for graph in (True, False):
        model, net = E.from_spec(spec)
        net.set_graph(graph)
        for n in (370, 16, 5, 641, 1, 300):
            net.run(n)
        st, ix = net.spikes()
        res.append((st, ix, [model.GetVectorNeuronState(v) for v in range(4)]))
        net.close()
        model.close()
    (a, b, sa), (c, d, sc) = res
    assert len(a) > 100
    assert np.array_equal(a, c) and np.array_equal(b, d)
    for x, y in zip(sa, sc):
        assert np.array_equal(bits(x), bits(y))


@pytest.mark.parametrize("kind", ["adex", "lif"])
def test_two_items_per_thread_launch(oracle, kind):
    """Plain AdEx/LIF networks also run with two neurons per thread (the
    autotuner's other tile shape): same raster and state as the oracle."""
    if kind == "adex":
        spec, params = workloads.adex_recurrent(n=20000, steps=1500, seed=5, w_ext=1.0), ADEX_DEFAULTS
    else:
        spec, params = workloads.lif_poisson(n=1000, steps=3000), LIF_DEFAULTS
    print(f"case two-items {kind}", flush=True)
    model, net = E.from_spec(spec)
    net.set_launch(2, 32, 256)
    net.run(spec_steps := (1500 if kind == "adex" else 3000))
    st, ix = net.spikes()
    pop, onet = oracle.from_spec(spec, params, np_out_csr)
    ost, oix = onet.run(1, spec_steps)
    assert len(ost) > 100
    assert np.array_equal(st, ost) and np.array_equal(ix, oix)
    for v in range(pop.nv):
        assert np.array_equal(bits(model.GetVectorNeuronState(v)), bits(pop.state(v))), f"variable {v}"


def test_reset_and_clear_log_empty_source_logs():
    """reset and clear_log empty the Poisson sources' spike logs with the
    population's: a run after a reset logs the same spikes as the first run,
    with nothing left over from it."""
    spec = workloads.lif_poisson(n=200, steps=0)
    print("case source logs reset", flush=True)
    model, net = E.from_spec(spec)
    net.set_source_log(0, 1 << 16)
    net.run(300)
    first = net.source_spikes(0)
    assert len(first[0]) > 100
    net.reset()
    assert len(net.source_spikes(0)[0]) == 0 and len(net.spikes()[0]) == 0
    net.run(300)
    again = net.source_spikes(0)
    assert np.array_equal(first[0], again[0]) and np.array_equal(first[1], again[1])
    net.clear_log()
    assert len(net.source_spikes(0)[0]) == 0 and len(net.spikes()[0]) == 0
    net.run(50)
    st, _ = net.source_spikes(0)
    assert len(st) > 0 and st.min() == 301
