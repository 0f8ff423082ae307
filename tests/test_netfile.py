"""Readers of EDLUT's input files and the simulation front-end's host mapping (CPU).

The fixtures under tests/golden/edlut_input_files are the reference's own
example inputs (Kernel/utils/edlut_input_files: network_file.cfg,
weight_file.cfg, configuration/_LIFTimeDriven.cfg, Input_spikes.cfg).  The
reference holds no expected outputs for them; the expectations below are read
off the files and the format rules of Network::ParseNet
(Kernel/src/spike/Network.cpp:422-706) and FileInputSpikeDriver::LoadInputs
(Kernel/src/communication/FileInputSpikeDriver.cpp:45-85): the weights file's
1.5 exceeds the synapses' max_weight 1.0 and is clamped, delays stay float.
"""
import os

import numpy as np
import pytest

from edlut_amd import netfile as NF
from edlut_amd import simulation as S
from edlut_amd._capi import EDLUTException

GOLD = os.path.join(os.path.dirname(__file__), "golden", "edlut_input_files")


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reference_example_network():
    layers, rules, syn = NF.parse_net(os.path.join(GOLD, "network_file.cfg"),
                                      os.path.join(GOLD, "weight_file.cfg"), base_dir=GOLD)
    assert [(L.num_neurons, L.neuron_model.model_name, L.output_activity, L.log_activity) for L in layers] == [
        (3, "InputSpikeNeuronModel", False, False), (1, "InputCurrentNeuronModel", False, False),
        (2, "LIFTimeDrivenModel", True, False)]
    p = layers[2].neuron_model.param_map
    expect = {"e_exc": 0.0, "e_inh": -80.0, "e_leak": -65.0, "v_thr": -50.0, "c_m": 2.0, "tau_exc": 1.0,
              "tau_inh": 10.0, "tau_ref": 1.0, "g_leak": 1.0, "tau_nmda": 14.0}
    for k, v in expect.items():
        assert p[k] == v, k
    assert p["int_meth"] == {"name": "Bifixed_RK2", "step": float(np.float32(0.001)), "n_steps": 3}
    assert rules == []
    assert syn.source.tolist() == [0, 1, 2, 3]
    assert syn.target.tolist() == [4, 4, 4, 5]
    assert syn.type.tolist() == [0, 1, 2, 3]
    assert syn.delay.dtype == np.float32 and (syn.delay == np.float32(0.001)).all()
    assert (syn.max_weight == 1.0).all()
    assert syn.weight.tolist() == [1.0, 1.0, 1.0, 1.0]      # 1.5 clamped to max_weight
    assert (syn.wchange == -1).all() and (syn.trigger_wchange == -1).all()


def test_reference_example_input_spikes():
    t, c = NF.parse_input_spikes(os.path.join(GOLD, "Input_spikes.cfg"), 6)
    assert len(t) == 300
    for line, (t0, iv, cell) in enumerate([(0.001, 0.010, 0), (0.002, 0.011, 1), (0.001, 0.012, 2)]):
        tt, cc = t[line * 100:(line + 1) * 100], c[line * 100:(line + 1) * 100]
        assert (cc == cell).all()
        # float(time + j * interval), the float arithmetic of LoadInputs
        j = np.arange(100, dtype=np.float32)
        want = (np.float32(t0) + j * np.float32(iv)).astype(np.float32).astype(np.float64)
        assert np.array_equal(tt, want)


def test_reader_primitives(tmp_path):
    f = NF._CFile(_write(tmp_path, "a.cfg", "// c1\n  // c2\n017 0x1f -012 12 t5 7 t x\n 3.5e-1 /c\n"))
    f.skip_comments()
    assert f.line == 3
    assert f.scan_i() == 15          # %i: leading 0 is octal
    assert f.scan_i() == 31          # 0x hex
    assert f.scan_i() == -10
    assert f.scan_d() == 12
    assert f.scan_t_d() == 5
    assert f.scan_t_d() is None      # no 't': nothing but white space consumed
    assert f.scan_d() == 7
    assert f.scan_t_d() is None      # 't' matched, %d failed: 't' stays consumed
    assert f.scan_s() == "x"
    assert f.scan_f() == np.float32(0.35)
    assert f.is_end_line()           # a comment follows


def _net_text(synapse_lines, n_syn, rules="0\n"):
    return ("// two layers\n2\n4\n2 InputSpikeNeuronModel Null 0 0\n2 LIFTimeDrivenModel_GPU lif 1 0\n"
            + rules + f"{n_syn}\n" + synapse_lines)


def _lif_cfg(tmp_path):
    _write(tmp_path, "lif.cfg", "0.0\n-80.0\n-65.0\n-50.0\n2.0\n5.0\n10.0\n1.0\n0.2\n20.0\nRK4\n0.0001\n")


def test_synapse_lines_and_rules(tmp_path):
    _lif_cfg(tmp_path)
    net = _write(tmp_path, "n.cfg", _net_text(
        "0 2 2 2 1 0.001 0.0005 0 5.0 t0 1 // trigger then rule\n2 1 3 1 2 0.002 0 1 2.0 1\n", 6,
        rules="2\nSTDP 0.01 0.02 0.011 0.03\nSTDP 0.5 0.1 0.25 0.2\n"))
    w = _write(tmp_path, "w.cfg", "4 7.5\n0 1.0\n")
    layers, rules, syn = NF.parse_net(net, w, base_dir=str(tmp_path))
    assert [r.param_map for r in rules] == [
        {"max_LTP": np.float32(0.01), "tau_LTP": np.float32(0.02), "max_LTD": np.float32(0.011),
         "tau_LTD": np.float32(0.03)},
        {"max_LTP": 0.5, "tau_LTP": np.float32(0.1), "max_LTD": 0.25, "tau_LTD": np.float32(0.2)}]
    # line 1: sources 0,1 x targets 2,3; line 2: 2 repetitions of 2 -> 3, then 3 -> 4 (out of range)
    assert syn.source.tolist()[:4] == [0, 0, 1, 1]
    assert syn.target.tolist()[:4] == [2, 3, 2, 3]
    d = np.float32(0.001) + np.float32(0.0005) * np.array([0, 1, 0, 1], np.float32)
    assert np.array_equal(syn.delay[:4], d)
    assert syn.trigger_wchange.tolist()[:4] == [0] * 4 and syn.wchange.tolist()[:4] == [1] * 4
    assert syn.weight.tolist() == [5.0] * 4 + [1.0, 1.0]     # 7.5 clamped to max 5.0; "0" = the rest
    assert syn.source.tolist()[4:] == [2, 3] and syn.target.tolist()[4:] == [3, 4]


@pytest.mark.parametrize("lines,n,msg", [
    ("0 2 2 2 1 0.001 0 0 1.0 -1\n", 3, "more synapses than declared"),
    ("0 1 2 1 1 0.001 0 0 1.0 0 1\n", 1, "two non-trigger"),
    ("0 1 2 1 1 0.001 0 0 1.0 t0 t0\n", 1, "two trigger"),
    ("0 1 2 1 1 0.001 0 0 1.0 3\n", 1, "does not exist"),
    ("0 1 2 1 1 0.001 0 0 1.0 0 t2\n", 1, "second learning rule index does not exist"),
    ("0 1 2 1 1 0.001 0 0\n", 1, "cannot be read"),
])
def test_synapse_line_errors(tmp_path, lines, n, msg):
    _lif_cfg(tmp_path)
    net = _write(tmp_path, "n.cfg", _net_text(lines, n, rules="2\nSTDP 0.01 0.02 0.011 0.03\nSTDP 1 1 1 1\n"))
    w = _write(tmp_path, "w.cfg", "0 1.0\n")
    with pytest.raises(NF.EDLUTFileException, match=msg):
        NF.parse_net(net, w, base_dir=str(tmp_path))


def test_file_errors_name_the_line(tmp_path):
    _write(tmp_path, "lif.cfg", "0.0\n-80.0\n-65.0\n// comment\n-50.0\nx\n")
    net = _write(tmp_path, "n.cfg", _net_text("0 1 2 1 1 0.001 0 0 1.0 -1\n", 1))
    w = _write(tmp_path, "w.cfg", "0 1.0\n")
    with pytest.raises(NF.EDLUTFileException) as e:
        NF.parse_net(net, w, base_dir=str(tmp_path))
    assert e.value.filename.endswith("lif.cfg") and e.value.line == 6
    _lif_cfg(tmp_path)
    w = _write(tmp_path, "w.cfg", "2 1.0\n")          # too many weights
    with pytest.raises(NF.EDLUTFileException, match="more weights"):
        NF.parse_net(net, w, base_dir=str(tmp_path))
    with pytest.raises(NF.EDLUTFileException, match="cannot be opened"):
        NF.parse_net(str(tmp_path / "missing.cfg"), w)


def test_random_weights_follow_the_global_seed(tmp_path, oracle):
    """Weights below zero draw frand() * max_weight from a RandomGenerator
    taking the global seed (Network.cpp:684-690); the LCG is the oracle's,
    itself pinned to the reference-built RandomGenerator."""
    _lif_cfg(tmp_path)
    net = _write(tmp_path, "n.cfg", _net_text("0 2 2 2 1 0.001 0 0 4.0 -1\n", 4))
    w = _write(tmp_path, "w.cfg", "0 -1\n")
    S.RandomGenerator.srand(12345)
    rng = S.RandomGenerator()
    assert S.RandomGenerator.global_seed == int(oracle.lcg_seeds(12345, 2)[1])
    _, _, syn = NF.parse_net(net, w, base_dir=str(tmp_path), rng=rng)
    u = oracle.lcg_frand(12345, 4)
    assert np.array_equal(syn.weight, (u.astype(np.float64) * 4.0).astype(np.float32))


def test_round_propagation_delay_and_ring_delay():
    # RoundPropagationDelay: floor to the 0.1 ms grid with a half-grid guard
    d = S.round_propagation_delay(np.array([0.001, 0.00015, 0.0, 0.00034999], np.float32))
    assert np.allclose(d, [0.001, 0.0001, 0.0001, 0.0003], rtol=0, atol=1e-15)
    # brute force on the event-time grid: an event sent at t_s with delay D is
    # consumed by the first update k with t_k > t_s + D
    for step in (1e-4, 2e-4, 1e-3):
        fdt = float(np.float32(step))
        grid = S.update_times(fdt, 3000)
        for D in S.round_propagation_delay(np.arange(1, 40, dtype=np.float32) * np.float32(1e-4)):
            d = S.ring_delay(float(D), fdt)
            for s in (1, 7, 500, 2000):
                k = int(np.searchsorted(grid, grid[s] + D, side="right"))
                assert k == s + d + 1, (step, D, s)
    # at a 0.1 ms step a delay of d grid points is d updates (float(1e-4) < 1e-4)
    assert [S.ring_delay(float(D), float(np.float32(1e-4))) for D in S.round_propagation_delay(
        np.arange(1, 11, dtype=np.float32) * np.float32(1e-4))] == list(range(1, 11))
    # at a 1 ms step a 1 ms delay reaches the very next update: not representable
    assert S.ring_delay(0.001, float(np.float32(1e-3))) == 0


def _small_sim():
    sim = S.Simulation_API()
    rk4 = {"int_meth": {"name": "RK4", "step": 1e-4, "bdf_order": 1}}
    inp = sim.AddNeuronLayer(4, "InputSpikeNeuronModel")
    poi = sim.AddNeuronLayer(3, "PoissonGeneratorDeviceVector", {"frequency": 40.0})
    adex = sim.AddNeuronLayer(5, "AdExTimeDrivenModel_GPU", rk4, output_activity=True)
    lif = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    adex2 = sim.AddNeuronLayer(2, "AdExTimeDrivenModel_GPU", rk4)          # same type as `adex`
    r = sim.AddLearningRule("STDP", {"max_LTP": 0.01, "tau_LTP": 0.02, "max_LTD": 0.012, "tau_LTD": 0.03})
    sim.AddSynapticLayer(inp[:2] * 2, adex[:4], {"delay": 0.001, "type": 0, "weight": 3.0})
    sim.AddSynapticLayer(poi, adex[:3], {"delay": [0.0002, 0.0003, 0.0004], "type": 0, "weight": 2.0})
    sim.AddSynapticLayer(adex, lif * 2 + lif[:1], {"delay": 0.0005, "type": 0, "weight": 1.0, "wchange": r,
                                                    "max_weight": 4.0})
    sim.AddSynapticLayer(lif, adex2, {"delay": 0.0002, "type": 1, "weight": 2.0})
    sim.AddSynapticLayer(adex[:2], adex2, {"delay": 0.0003, "type": 0, "weight": 1.5})
    return sim


def test_plan_maps_layers_and_synapses():
    S.RandomGenerator.srand(7)
    sim = _small_sim()
    plan = sim.build_plan()
    names = [t.model.model_name for t in plan.types]
    assert names == ["InputSpikeNeuronModel", "PoissonGeneratorDeviceVector", "AdExTimeDrivenModel_GPU",
                     "LIFTimeDrivenModel_GPU"]
    adex = plan.types[2]
    assert adex.globals == [7, 8, 9, 10, 11, 14, 15]            # both AdEx layers, in layer order
    assert adex.output.tolist() == [True] * 5 + [False] * 2
    assert plan.max_delay == 10
    kinds = [(s["kind"], s["src_type"], s["stdp"] is not None) for s in plan.sources[2]]
    assert kinds == [("train", 0, False), ("poisson", 1, False), ("self", 2, False), ("link", 3, False)]
    assert [(s["kind"], s["src_type"], s["stdp"] is not None) for s in plan.sources[3]] == [("link", 2, True)]
    # Poisson generators take the global seeds in type order
    assert plan.poisson_states[1].tolist() == [7, (7 * 765426813 + 54321) & 0x7FFFFFFF,
                                               (((7 * 765426813 + 54321) & 0x7FFFFFFF) * 765426813 + 54321)
                                               & 0x7FFFFFFF]
    # delays in updates: 0.2/0.3/0.4 ms -> 2/3/4, 1 ms -> 10
    poi = plan.sources[2][1]
    assert poi["edges"][2].tolist() == [2, 3, 4]
    assert plan.sources[2][0]["edges"][2].tolist() == [10] * 4


def test_plan_refusals():
    sim = S.Simulation_API()
    a = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-3}})
    sim.AddSynapticLayer(a, a[::-1], {"delay": 0.001})
    with pytest.raises(EDLUTException, match="next update|right after"):
        sim.build_plan()
    sim = S.Simulation_API()
    a = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    sim.AddSynapticLayer(a, a[::-1], {"type": 3})
    with pytest.raises(EDLUTException, match="type 3 takes currents"):
        sim.build_plan()
    sim = S.Simulation_API()
    sim.AddNeuronLayer(2, "LIFTimeDrivenModel", {"int_meth": {"name": "Bifixed_RK2", "step": 1e-3, "n_steps": 3}})
    with pytest.raises(EDLUTException, match="Bifixed"):
        sim.build_plan()
    sim = S.Simulation_API()
    with pytest.raises(EDLUTException, match="not available"):
        sim.AddLearningRule("ExpOptimisedBufferedAdditiveKernel", {})
    # Poisson generators sample every 0.1 ms (PoissonGeneratorDeviceVector.cpp:26,137):
    # a 1 ms integration step is refused, 0.1 ms is not
    for step, refused in ((1e-3, True), (1e-4, False)):
        sim = S.Simulation_API()
        p = sim.AddNeuronLayer(5, "PoissonGeneratorDeviceVector", {"frequency": 10.0})
        a = sim.AddNeuronLayer(5, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": step}})
        sim.AddSynapticLayer(p[:2], a[:2], {"delay": 0.001, "weight": 1.0})
        if refused:
            with pytest.raises(EDLUTException, match="0.1 ms"):
                sim.build_plan()
        else:
            sim.build_plan()


def test_input_spike_placement_brute_force():
    """Each input spike lands on the update s whose events are consumed by
    s + d + 1, checked against the first update after t + D on the grid."""
    sim = _small_sim()
    sim.plan = sim.build_plan()
    fdt = float(np.float32(1e-4))
    grid = S.update_times(fdt, 5000)
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0.0002, 0.3, 200))
    c = rng.integers(0, 2, 200)
    steps = sim.place_input_spikes(t, c)
    for q in range(200):
        D = S.round_propagation_delay(np.float32(0.001))[()]
        k = int(np.searchsorted(grid, t[q] + D, side="right"))
        assert k == steps[q] + 10 + 1
    # input neuron 2 has no synapses: nothing to place
    assert sim.place_input_spikes([0.01], [2]).tolist() == [0]
    with pytest.raises(EDLUTException, match="before the first update"):
        sim.place_input_spikes([0.00001], [0])


def test_reference_example_input_currents():
    """FileInputCurrentDriver format (FileInputCurrentDriver.cpp:45-90) on the
    reference's own example: 300 pA on neuron 3 every 50 ms from 10 ms, back
    to 0 pA every 50 ms from 60 ms."""
    t, c, v = NF.parse_input_currents(os.path.join(GOLD, "Input_currents.cfg"), 6)
    assert len(t) == 20 and (c == 3).all()
    j = np.arange(10, dtype=np.float32)
    assert np.array_equal(t[:10], (np.float32(0.010) + j * np.float32(0.050)).astype(np.float64))
    assert np.array_equal(t[10:], (np.float32(0.060) + j * np.float32(0.050)).astype(np.float64))
    assert (v[:10] == np.float32(300.0)).all() and (v[10:] == 0.0).all()


def test_current_synapses_need_current_sources():
    sim = S.Simulation_API()
    inp = sim.AddNeuronLayer(1, "InputSpikeNeuronModel")
    cur = sim.AddNeuronLayer(1, "InputCurrentNeuronModel")
    lif = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    sim.AddSynapticLayer(inp, lif[:1], {"type": 3})
    with pytest.raises(EDLUTException, match="type 3 takes currents"):
        sim.build_plan()
    sim = S.Simulation_API()
    inp = sim.AddNeuronLayer(1, "InputSpikeNeuronModel")
    cur = sim.AddNeuronLayer(1, "InputCurrentNeuronModel")
    lif = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    sim.AddSynapticLayer(cur, lif[:1], {"type": 0})
    with pytest.raises(EDLUTException, match="type 3 takes currents"):
        sim.build_plan()
    # accepted: currents on type 3, kept on the host in creation order per target
    sim = S.Simulation_API()
    inp = sim.AddNeuronLayer(1, "InputSpikeNeuronModel")
    cur = sim.AddNeuronLayer(2, "InputCurrentNeuronModel")
    lif = sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU", {"int_meth": {"name": "Euler", "step": 1e-4}})
    sim.AddSynapticLayer(inp, lif[:1], {"type": 0})
    sim.AddSynapticLayer([cur[1], cur[0], cur[1]], [lif[1], lif[1], lif[0]], {"type": 3, "delay": 0.0005})
    plan = sim.build_plan()
    c = plan.current[2]
    assert c["syn"].tolist() == [1, 2, 3] and c["src"].tolist() == [2, 1, 2] and c["tgt"].tolist() == [1, 1, 0]
    assert [s["kind"] for s in plan.sources[2]] == ["train"]


def test_layers_with_equal_parameters_share_a_population():
    """Network::CreateNeuronLayers merges layers whose models compare equal
    (NeuronModel::compare): defaults given explicitly or left out are the
    same model; another value or another step is not."""
    sim = S.Simulation_API()
    im = {"int_meth": {"name": "RK4", "step": 1e-4}}
    a = sim.AddNeuronLayer(2, "AdExTimeDrivenModel_GPU", dict(im))
    b = sim.AddNeuronLayer(2, "AdExTimeDrivenModel_GPU", dict(im, a=1.0, b=9.0))
    c = sim.AddNeuronLayer(2, "AdExTimeDrivenModel_GPU", dict(im, a=2.0))
    d = sim.AddNeuronLayer(2, "AdExTimeDrivenModel_GPU", {"int_meth": {"name": "RK4", "step": 1e-4, "bdf_order": 1}})
    plan = sim.build_plan()
    assert plan.type_of[a + b + c + d].tolist() == [0, 0, 0, 0, 1, 1, 0, 0]


def test_simulation_api_call_order_and_queries():
    """Simulation_API's call-order checks and registry queries
    (Simulation_API.cpp:63-443): construction calls before Initialize, run
    calls after it; unknown simulation parameters and learning rules raise."""
    sim = S.Simulation_API()
    with pytest.raises(EDLUTException, match="after Initialize"):
        sim.RunSimulation(0.1)
    with pytest.raises(EDLUTException, match="after Initialize"):
        sim.AddExternalSpikeActivity([0.01], [0])
    with pytest.raises(EDLUTException, match="after Initialize"):
        sim.GetSpikeActivity()
    with pytest.raises(EDLUTException, match="invalid simulation parameter"):
        sim.SetSimulationParameters({"no_such_parameter": 1})
    sim.SetSimulationParameters({"num_threads": 2})
    assert sim.simulation_properties["num_threads"] == 2
    assert sim.AddNeuronLayer(3, "InputSpikeNeuronModel") == [0, 1, 2]
    assert sim.AddNeuronLayer(2, "LIFTimeDrivenModel_GPU") == [3, 4]
    assert sim.AddSynapticLayer([0, 1], [3, 4], {"weight": [0.5, 0.25]}) == [0, 1]
    assert sim.AddSynapticLayer([2], [4], {}) == [2]
    with pytest.raises(EDLUTException, match="unknown synapse parameters"):
        sim.AddSynapticLayer([2], [3], {"colour": 1})
    models = S.Simulation_API.GetAvailableNeuronModels()
    assert "AdExTimeDrivenModel_GPU" in models and "InputCurrentNeuronModel" in models
    assert S.Simulation_API.GetNeuronModelDefParams("PoissonGeneratorDeviceVector")["frequency"] == 10.0
    assert S.Simulation_API.GetNeuronModelDefParams("LIFTimeDrivenModel_GPU")["tau_ref"] == 1.0
    assert S.Simulation_API.GetLearningRuleDefParams("STDP")["tau_LTP"] == 0.100
    assert S.Simulation_API.GetIntegrationMethodDefParams("Bifixed_RK4")["n_steps"] == 2
    with pytest.raises(EDLUTException):
        S.Simulation_API.GetNeuronModelDefParams("NoSuchModel")


def test_compressed_weights():
    """Network::GetCompressedWeights (Network.cpp:1204-1231): runs of equal
    consecutive weights, a leading zero weight included."""
    assert S.compress_weights([]) == ([], [])
    assert S.compress_weights([0.0, 0.0, 1.5, 1.5, 1.5, 0.0, 2.0]) == ([2, 3, 1, 1], [0.0, 1.5, 0.0, 2.0])
    c, v = S.compress_weights(np.float32([0.1, 0.1, 0.2]))
    assert c == [2, 1] and v == [float(np.float32(0.1)), float(np.float32(0.2))]
