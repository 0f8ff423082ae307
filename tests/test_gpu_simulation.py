"""GPU parity of the simulation front-end (Simulation_API and the file readers).

A simulation is built through Simulation_API (or from network, weights and
input-spike files), run on the device, and compared with the oracle's
circuit built from the same host mapping: every GPU population's spike
raster and state, and the plastic weights, must be bit-identical.  The
mapping itself (types, delays in updates, input spike placement, seeds) is
pinned on the CPU in tests/test_netfile.py.
"""
import os

import numpy as np
import pytest

from edlut_amd import simulation as S
from edlut_amd.models import REGISTRY
from edlut_amd._capi import EDLUTException
from tests.helpers import bits, np_out_csr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "edlut_input_files")


def _oracle(oracle, sim):
    """The oracle circuit of the same mapping, models ranked and input spikes
    timed and loaded as the front-end gives them to the device."""
    plan = sim.plan
    pops, nets, order = {}, {}, []
    steps = sim.place_input_spikes(sim._input_times, sim._input_cells)
    for t in plan.types:
        name = S.gpu_model_name(t.model)
        if name is None:
            continue
        im = t.model.param_map["int_meth"]
        params = np.array([t.model.param_map.get(k, v) for k, v in REGISTRY[name].PARAMS], np.float32)
        mask = 8 if t.index in plan.current else 0
        for s in plan.sources.get(t.index, []):
            for ty in np.unique(s["edges"][3]):
                mask |= 1 << int(ty)
        pops[t.index] = oracle.Population(name, im["name"], params, len(t.globals), dt_s=np.float32(im["step"]),
                                          mask=mask, bdf_order=im.get("bdf_order", 1), n_steps=im.get("n_steps", 2))
        nets[t.index] = oracle.Net(pops[t.index], plan.max_delay)
        nets[t.index].set_model_rank(t.index)
        order.append(t.index)
    for ti in order:
        net = nets[ti]
        for slot, s in enumerate(plan.sources[ti]):
            csr = np_out_csr(s["n_src"], len(plan.types[ti].globals), plan.max_delay, *s["edges"])
            if s["kind"] == "self":
                net.add_self(*csr)
            elif s["kind"] == "link":
                net.add_link(nets[s["src_type"]], *csr)
            elif s["kind"] == "poisson":
                net.add_poisson(plan.poisson_states[s["src_type"]], 0.0, *csr)
                f = plan.types[s["src_type"]].model.param_map.get("frequency", 10.0)
                net.set_poisson_rates(slot, np.full(s["n_src"], np.float32(f), np.float32))
                net.set_source_model(slot, s["src_type"], s["src_type"])
            else:
                ptr, idx, times, rank, load = sim.train_schedule(s["src_type"], steps)
                net.add_train(s["n_src"], ptr, idx, *csr)
                net.set_source_model(slot, s["src_type"], s["src_type"])
                net.set_train_times(slot, times, rank, load)
            if s["stdp"] is not None:
                p, wmax = s["stdp"]
                net.set_stdp(slot, float(p["max_LTP"]), float(p["tau_LTP"]), float(p["max_LTD"]),
                             float(p["tau_LTD"]), float(wmax))
    return pops, nets, order


def _compare(oracle, sim, n_steps):
    pops, nets, order = _oracle(oracle, sim)
    # the device runs (pieces of RunSimulation, with their stop events) and
    # the external currents the front-end set before them, replayed
    ci = 0
    for first, n, stop in sim.run_log:
        while ci < len(sim.current_log) and sim.current_log[ci][0] <= first:
            _, tidx, total = sim.current_log[ci]
            pops[tidx].set_ext_current(total)
            ci += 1
        nets[order[0]].set_stop_time(stop)
        oracle.circuit_run([nets[i] for i in order], first, n)
    assert sum(n for _, n, _ in sim.run_log) == n_steps
    total = 0
    for ti in order:
        gs, gi = sim.nets[ti].spikes()
        os_, oi = nets[ti].take_log()
        assert np.array_equal(gs, os_) and np.array_equal(gi, oi), f"raster of type {ti} differs"
        total += len(gs)
        m = sim.models[ti]
        for v in range(len(m.STATE_VARS)):
            assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(pops[ti].state(v).astype(np.float32))), \
                f"type {ti} state {v}"
        for slot, s in enumerate(sim.plan.sources[ti]):
            if s["stdp"] is not None:
                assert np.array_equal(bits(sim.nets[ti].synapses(slot)[2]), bits(nets[ti].synapses(slot)[0]))
    return total


def _driven_sim(seed=7, full=True):
    S.RandomGenerator.srand(seed)
    sim = S.Simulation_API(full_spike_log=full)
    rk4 = {"int_meth": {"name": "RK4", "step": 1e-4, "bdf_order": 1}}
    inp = sim.AddNeuronLayer(4, "InputSpikeNeuronModel", output_activity=True)
    poi = sim.AddNeuronLayer(30, "PoissonGeneratorDeviceVector", {"frequency": 150.0}, output_activity=True)
    adex = sim.AddNeuronLayer(40, "AdExTimeDrivenModel_GPU", rk4, output_activity=True)
    lif = sim.AddNeuronLayer(12, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    adex2 = sim.AddNeuronLayer(10, "AdExTimeDrivenModel_GPU", rk4, output_activity=True)
    r = sim.AddLearningRule("STDP", {"max_LTP": 0.01, "tau_LTP": 0.02, "max_LTD": 0.012, "tau_LTD": 0.03})
    rng = np.random.default_rng(seed)
    tg = rng.integers(0, 40, 60)
    sim.AddSynapticLayer(rng.integers(0, 4, 60) + inp[0], np.array(adex)[tg], {"delay": 0.001, "weight": 30.0})
    src = rng.integers(0, 30, 300)
    sim.AddSynapticLayer(np.array(poi)[src], np.array(adex + adex2)[rng.integers(0, 50, 300)],
                         {"delay": (rng.integers(1, 6, 300) * np.float32(1e-4)).astype(np.float32).tolist(),
                          "weight": rng.uniform(10, 30, 300).astype(np.float32).tolist()})
    sim.AddSynapticLayer(np.array(adex)[rng.integers(0, 40, 200)], np.array(adex)[rng.integers(0, 40, 200)],
                         {"delay": 0.0007, "type": 1, "weight": 5.0})
    sim.AddSynapticLayer(np.array(adex)[rng.integers(0, 40, 120)], np.array(lif)[rng.integers(0, 12, 120)],
                         {"delay": 0.0005, "weight": 0.5, "wchange": r, "max_weight": 2.0})
    sim.AddSynapticLayer(np.array(lif)[rng.integers(0, 12, 60)], np.array(adex2)[rng.integers(0, 10, 60)],
                         {"delay": 0.0002, "type": 1, "weight": 3.0})
    return sim, inp, poi


def _input_spikes(seed, n, lo, hi, first):
    """n spikes of four input neurons between lo and hi seconds, at most one
    per neuron and update (the device's trains hold one spike per source and
    update), off the update grid by 0.3 of a step."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, 4, n)
    slots = rng.choice(np.arange(int(lo / 1e-4), int(hi / 1e-4)), n, replace=False)
    t = slots * 1e-4 + 3e-5
    order = np.argsort(t)
    return t[order], cells[order] + first


def test_simulation_api_network_bit_exact(oracle):
    sim, inp, poi = _driven_sim()
    sim.Initialize()
    t, c = _input_spikes(1, 80, 0.0005, 0.15, inp[0])
    sim.AddExternalSpikeActivity(t, c)
    sim.RunSimulation(0.2)
    assert sim.current_step == int(np.searchsorted(S.update_times(float(np.float32(1e-4)), 2100), 0.2,
                                                   side="right")) - 1
    assert _compare(oracle, sim, sim.current_step) > 50
    times, cells = sim.GetSpikeActivity()
    assert len(times) > 0 and times == sorted(times)
    # every output neuron's spikes: the input neurons' times and the
    # population spikes at their update times
    grid = S.update_times(float(np.float32(1e-4)), sim.current_step)
    want = [(float(ti), int(ci)) for ti, ci in zip(t, c) if ti <= grid[-1]]
    for ti in sim.nets:
        st, ix = sim.nets[ti].spikes()
        T = sim.plan.types[ti]
        want += [(float(grid[s]), int(T.globals[i])) for s, i in zip(st, ix) if T.output[i]]
    # the Poisson layer's generators, replayed on the host from their seeds
    # (RandomGenerator LCG; fire when frequency * step > frand())
    x = sim.plan.poisson_states[1].astype(np.uint64)
    p = float(np.float32(150.0)) * float(np.float32(0.1) / np.float32(1000.0))   # the device's own 0.1 ms step
    n_poi = 0
    for k in range(1, sim.current_step + 1):
        x = (x * np.uint64(1103515245) + np.uint64(12345)) & np.uint64(0x7FFFFFFF)
        u = x.astype(np.float32) / np.float32(2147483647.0)
        for i in np.nonzero(p > u.astype(np.float64))[0]:
            want.append((float(grid[k]), int(poi[0] + i)))
            n_poi += 1
    assert n_poi > 100
    assert sorted(want) == sorted(zip(times, cells))
    assert sim.GetSpikeActivity() == ([], [])      # the buffer was emptied
    w = sim.GetWeights()
    assert len(w) == sim.num_synapses_created
    sim.close()


def test_external_spikes_between_runs(oracle):
    """Spikes added between runs (the closed-loop use of
    AddExternalSpikeActivity) join the event queue when the next run starts,
    as in the reference; the oracle queues them at the same point."""
    sim, inp, _ = _driven_sim(seed=11)
    sim.Initialize()
    t1, c1 = _input_spikes(5, 30, 0.001, 0.05, inp[0])
    t2, c2 = _input_spikes(6, 30, 0.06, 0.1, inp[0])
    sim.AddExternalSpikeActivity(t1, c1)
    sim.RunSimulation(0.05)
    with pytest.raises(EDLUTException, match="already simulated"):
        sim.AddExternalSpikeActivity([0.01], [inp[0]])
        sim.RunSimulation(0.06)
    sim._input_times = sim._input_times[:30]
    sim._input_cells = sim._input_cells[:30]
    sim._input_load = sim._input_load[:30]
    sim._trains_dirty = False
    sim.AddExternalSpikeActivity(t2, c2)
    sim.RunSimulation(0.1)
    assert _compare(oracle, sim, sim.current_step) > 10
    sim.close()


def test_from_files_reference_inputs(oracle, tmp_path):
    """The reference's example input file (three input neurons, AMPA, GABA
    and NMDA synapses) driving GPU LIF neurons at a 0.1 ms step, built from
    network and weights files."""
    (tmp_path / "lif.cfg").write_text("// LIF\n0.0\n-80.0\n-65.0\n-50.0\n2.0\n1.0\n10.0\n1.0\n1.0\n14.0\n"
                                      "Euler\n0.0001\n")
    net = tmp_path / "net.cfg"
    net.write_text("2\n5\n3 InputSpikeNeuronModel Null 0 0\n2 LIFTimeDrivenModel_GPU lif 1 0\n0\n6\n"
                   "0 1 3 2 1 0.001 0 0 1.0 -1\n1 1 3 2 1 0.002 0 1 1.0 -1\n2 1 3 2 1 0.001 0.0005 2 1.0 -1\n")
    w = tmp_path / "w.cfg"
    w.write_text("2 0.9\n2 0.5\n0 0.7\n")
    S.RandomGenerator.srand(1)
    sim = S.Simulation_API.from_files(str(net), str(w), os.path.join(GOLD, "Input_spikes.cfg"),
                                      base_dir=str(tmp_path), full_spike_log=True)
    sim.RunSimulation(1.25)
    n = _compare(oracle, sim, sim.current_step)
    assert n > 0
    assert sim.GetWeights() == [np.float32(0.9)] * 2 + [0.5] * 2 + [np.float32(0.7)] * 2
    sim.close()


def test_reference_example_with_currents(oracle, tmp_path):
    """The reference's example network (Kernel/utils/edlut_input_files:
    three input neurons on AMPA, GABA and NMDA synapses of one neuron, an
    input-current neuron on a current synapse of another) with GPU LIF
    neurons at a 0.1 ms step, driven by the reference's Input_spikes.cfg and
    Input_currents.cfg.  Currents change only between updates: the run is
    split where they reach the population."""
    (tmp_path / "lif.cfg").write_text("0.0\n-80.0\n-65.0\n-50.0\n2.0\n1.0\n10.0\n1.0\n1.0\n14.0\nEuler\n0.0001\n")
    net = tmp_path / "net.cfg"
    net.write_text("3\n6\n3 InputSpikeNeuronModel Null 0 0\n1 InputCurrentNeuronModel Null 0 0\n"
                   "2 LIFTimeDrivenModel_GPU lif 1 0\n0\n4\n0 1 4 1 1 0.001 0 0 1.0 -1\n1 1 4 1 1 0.001 0 1 1.0 -1\n"
                   "2 1 4 1 1 0.001 0 2 1.0 -1\n3 1 5 1 1 0.001 0 3 1.0 -1\n")
    w = tmp_path / "w.cfg"
    w.write_text("3 1.5\n1 1.0\n")
    S.RandomGenerator.srand(1)
    sim = S.Simulation_API.from_files(str(net), str(w), os.path.join(GOLD, "Input_spikes.cfg"), base_dir=str(tmp_path),
                                      input_current_file=os.path.join(GOLD, "Input_currents.cfg"),
                                      full_spike_log=True)
    sim.RunSimulation(0.3)
    sim.RunSimulation(1.25)
    # 20 current changes reach neuron 5 at the first update after t + 1 ms
    grid = S.update_times(float(np.float32(1e-4)), sim.current_step)
    t, _, v = S.NF.parse_input_currents(os.path.join(GOLD, "Input_currents.cfg"), 6)
    D = float(S.round_propagation_delay(np.float32(0.001)))
    want = sorted({int(np.searchsorted(grid, x + D, side="right")) for x in t})
    assert [k for k, _, _ in sim.current_log] == want
    assert all(tot[1] in (np.float32(300.0), np.float32(0.0)) and tot[0] == 0.0 for _, _, tot in sim.current_log)
    n = _compare(oracle, sim, sim.current_step)
    st, ix = sim.nets[2].spikes()
    assert (ix == 1).sum() > 20, "the 300 pA steps should drive neuron 5"
    assert n > 0
    sim.close()


def _poisson_replay(states, freq, steps):
    """(update, generator) of every spike of a Poisson layer, replayed on the
    host from its seeds (RandomGenerator LCG; fire when frequency * 0.1 ms >
    frand(), PoissonGeneratorDeviceVector.cpp:137)."""
    x = np.asarray(states).astype(np.uint64)
    p = float(np.float32(freq)) * float(np.float32(0.1) / np.float32(1000.0))
    out = []
    for k in range(1, steps + 1):
        x = (x * np.uint64(1103515245) + np.uint64(12345)) & np.uint64(0x7FFFFFFF)
        u = x.astype(np.float32) / np.float32(2147483647.0)
        out += [(k, int(i)) for i in np.nonzero(p > u.astype(np.float64))[0]]
    return out


def test_spike_activity_drains_logs():
    """GetSpikeActivity empties the device logs on each read, as the output
    drivers' GetBufferedSpikes (Simulation_API.cpp:173-204): the pieces read
    after each run join to one read of a fully logged simulation, logs far
    smaller than the whole simulation's spikes do not overflow, a population
    without output neurons keeps no log, and a log that is never read
    overflows with an error."""
    print("case drain logs", flush=True)
    full, _, _ = _driven_sim(seed=5)
    full.Initialize()
    for q in range(1, 13):
        full.RunSimulation(0.01 * q)
    want = full.GetSpikeActivity()
    # the largest number of spikes one population logs in one run
    ends = [first + n - 1 for first, n, _ in full.run_log]
    piece = 0
    for ti in full.nets:
        st = full.nets[ti].spikes()[0]
        piece = max([piece] + np.diff(np.searchsorted(st, [0] + ends, side="right")).tolist())
    full.close()
    sim, _, _ = _driven_sim(seed=5, full=False)
    sim.Initialize()
    outs = [ti for ti in sim.nets if sim.plan.types[ti].output.any()]
    quiet = [ti for ti in sim.nets if not sim.plan.types[ti].output.any()]
    assert outs and quiet
    cap = 2 * piece
    for ti in outs:
        sim.nets[ti].set_log_capacity(cap)
    got_t, got_c = [], []
    for q in range(1, 13):
        sim.RunSimulation(0.01 * q)
        t, c = sim.GetSpikeActivity()
        got_t += t
        got_c += c
    assert len(want[0]) > 3 * cap
    assert sorted(zip(got_t, got_c)) == sorted(zip(*want))
    for ti in quiet:
        assert len(sim.nets[ti].spikes()[0]) == 0 and sim.nets[ti].spike_count() > 0
    with pytest.raises(EDLUTException, match="capacity"):
        sim.RunSimulation(0.36)
    sim.close()


def test_poisson_output_layer_without_synapses():
    """A Poisson layer with output neurons and no synapse still reports its
    generators' spikes (a copy without synapses logs them)."""
    print("case poisson output without synapses", flush=True)
    S.RandomGenerator.srand(3)
    sim = S.Simulation_API()
    lone = sim.AddNeuronLayer(25, "PoissonGeneratorDeviceVector", {"frequency": 120.0}, output_activity=True)
    drive = sim.AddNeuronLayer(10, "PoissonGeneratorDeviceVector", {"frequency": 80.0})
    lif = sim.AddNeuronLayer(5, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    sim.AddSynapticLayer(np.array(drive)[np.arange(20) % 10], np.array(lif)[np.arange(20) % 5],
                         {"delay": 0.0003, "weight": 1.0})
    sim.Initialize()
    sim.RunSimulation(0.05)
    t, c = sim.GetSpikeActivity()
    ty = sim.plan.type_of[lone[0]]
    grid = S.update_times(float(np.float32(1e-4)), sim.current_step)
    want = [(float(grid[k]), lone[0] + i) for k, i in _poisson_replay(sim.plan.poisson_states[ty], 120.0,
                                                                       sim.current_step)]
    assert len(want) > 50
    assert sorted(zip(t, c)) == sorted(want)
    sim.close()

