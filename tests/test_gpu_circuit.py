"""GPU parity of a multi-population circuit (the layered cerebellar network).

Mossy fibres (Poisson devices with one rate per fibre) drive granule cells
(Egidio model, BDF2); granule cells project onto Golgi cells (LIF) and onto
Purkinje cells (HH under BDF2 with the FP64 Newton step, or the Purkinje-cell
model) with additive STDP on granule -> Purkinje.  The three networks run in
lockstep on the device; the oracle runs the same circuit update by update.
Each population's spike raster and state, and every plastic synapse's
weight, traces and last update time, must be bit-identical.
"""
import numpy as np
import pytest

import edlut_amd as E
from edlut_amd import workloads as W
from edlut_amd.network import cerebellum_from_spec
from tests.helpers import bits, defaults, np_out_csr

pytestmark = pytest.mark.gpu


def _oracle_circuit(oracle, spec):
    D = spec.max_delay
    hh = spec.pc_model == "HHTimeDrivenModel_GPU"
    gr = oracle.Population("EgidioGranuleCell_TimeDriven_GPU", "BDF", [], spec.n_grc, mask=1, bdf_order=2)
    go = oracle.Population("LIFTimeDrivenModel_GPU", "Euler", defaults("LIFTimeDrivenModel_GPU"), spec.n_goc, mask=1)
    pc = oracle.Population(spec.pc_model, "BDF" if hh else "RK4", defaults(spec.pc_model), spec.n_pc, mask=1,
                           fp64=2 if hh else False, bdf_order=2)
    ng, no, npc = oracle.Net(gr, D), oracle.Net(go, D), oracle.Net(pc, D)
    ng.add_poisson(spec.mf_states, 0.0, *np_out_csr(spec.n_mf, spec.n_grc, D, *spec.mf_grc))
    ng.set_poisson_rates(0, spec.mf_rates)
    no.add_link(ng, *np_out_csr(spec.n_grc, spec.n_goc, D, *spec.grc_goc))
    npc.add_link(ng, *np_out_csr(spec.n_grc, spec.n_pc, D, *spec.grc_pc))
    npc.set_stdp(0, *spec.stdp)
    return (gr, go, pc), (ng, no, npc)


@pytest.mark.parametrize("pc_model,graph", [("HHTimeDrivenModel_GPU", True),
                                            ("TimeDrivenPurkinjeCell_GPU", True),
                                            ("HHTimeDrivenModel_GPU", False)])
def test_cerebellum_circuit_bit_exact(oracle, pc_model, graph):
    steps = 500
    print(f"case cerebellum pc={pc_model} graph={graph} steps={steps}", flush=True)
    # stronger granule -> Golgi / Purkinje weights than the full-size network:
    # 600 granule cells must make the small target populations fire
    w_pc = 2.0 if pc_model == "HHTimeDrivenModel_GPU" else 0.5
    spec = W.cerebellum(n_grc=600, n_goc=24, n_pc=5, n_mf=120, k_pc=300, pc_model=pc_model, steps=steps,
                        w_goc=0.3, w_pc=w_pc)
    models, nets, circuit = cerebellum_from_spec(spec)
    for n in nets:
        n.set_graph(graph)
    for chunk in (37, 300, 163):   # chunks of the graph path and direct steps
        circuit.run(chunk)
    opops, onets = _oracle_circuit(oracle, spec)
    oracle.circuit_run(list(onets), 1, steps)
    names = ("granule", "Golgi", "Purkinje")
    total = {}
    for name, m, n, op, on in zip(names, models, nets, opops, onets):
        st, ix = n.spikes()
        ost, oix = on.take_log()
        assert np.array_equal(st, ost) and np.array_equal(ix, oix), f"{name} raster"
        total[name] = len(ost)
        for v in range(op.nv):
            assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(op.state(v))), f"{name} variable {v}"
    assert total["granule"] > 50 and total["Golgi"] > 50 and total["Purkinje"] > 0, total
    w_o, pre_o, post_o, last_o = onets[2].synapses(0)
    _, _, w_g = nets[2].synapses(0)
    pre_g, post_g, last_g = nets[2].stdp_state(0)
    assert np.array_equal(bits(w_g), bits(w_o))
    assert np.array_equal(bits(pre_g), bits(pre_o)) and np.array_equal(bits(post_g), bits(post_o))
    assert np.array_equal(last_g.view(np.uint64), last_o.view(np.uint64))
    assert np.any(pre_o != 0), "granule spikes should reach the plastic synapses"
    circuit.close()


def test_cerebellum_full_in_degree(oracle):
    """Config 3's connectivity at a reduced population: 2,500 granule cells,
    Purkinje in-degree 2,000 (config 3's), Golgi in-degree 100, mossy-fibre
    in-degree 4, with the stronger granule weights of the small circuit (so
    the few granule spikes of a short run reach the targets); every raster,
    state and plastic synapse bit-identical to the oracle circuit."""
    steps = 800
    print(f"case cerebellum full in-degree steps={steps}", flush=True)
    spec = W.cerebellum(n_grc=2500, n_goc=50, n_pc=10, n_mf=500, k_pc=2000, steps=steps, w_goc=0.3, w_pc=2.0)
    models, nets, circuit = cerebellum_from_spec(spec)
    circuit.run(steps)
    opops, onets = _oracle_circuit(oracle, spec)
    oracle.circuit_run(list(onets), 1, steps)
    total = {}
    for name, m, n, op, on in zip(("granule", "Golgi", "Purkinje"), models, nets, opops, onets):
        st, ix = n.spikes()
        ost, oix = on.take_log()
        assert np.array_equal(st, ost) and np.array_equal(ix, oix), f"{name} raster"
        total[name] = len(ost)
        for v in range(op.nv):
            assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(op.state(v))), f"{name} variable {v}"
    print("spikes:", total)
    assert total["granule"] > 50 and total["Golgi"] > 0 and total["Purkinje"] > 0, total
    w_o, pre_o, post_o, last_o = onets[2].synapses(0)
    _, _, w_g = nets[2].synapses(0)
    assert np.array_equal(bits(w_g), bits(w_o))
    circuit.close()


def _lif_net(max_delay, n=10):
    dt = float(np.float32(1e-4))
    m = E.CreateNeuronModel("LIFTimeDrivenModel_GPU", {"int_meth": E.integration_method("Euler", dt)})
    m.CheckSynapseType(0)
    m.InitializeStates(n, 0, 0)
    return m, E.Network(m, max_delay)


def test_circuit_rules():
    """A network with projections runs only inside a circuit that holds its
    sources; members do not run on their own while the circuit exists;
    projections need equal max_delay."""
    ma, na = _lif_net(3)
    mb, nb = _lif_net(3)
    mc, nc = _lif_net(4)
    rp, tt, w = E.build_out_csr(10, 10, 3, np.arange(10), np.arange(10), np.ones(10, np.int32),
                                np.zeros(10, np.uint8), np.ones(10, np.float32))
    with pytest.raises(E.EDLUTException):
        nc.add_projection(na, rp, tt, w)
    nb.add_projection(na, rp, tt, w)
    with pytest.raises(E.EDLUTException):
        nb.run(5)
    with pytest.raises(E.EDLUTException):
        E.Circuit([nb])
    c = E.Circuit([na, nb])
    with pytest.raises(E.EDLUTException):
        na.run(5)
    c.run(20)
    c.close()
    # the event queue belonged to the circuit: the network does not continue
    # that run on its own, but runs again from a reset
    with pytest.raises(E.EDLUTException, match="restarts at update 1"):
        na.run(5, first_step=21)
    na.reset()
    na.run(5)
