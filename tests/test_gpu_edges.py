"""Edge cases of the HIP path against the oracle.

* populations of one neuron and of sizes that are not multiples of a tile,
  for every model;
* networks with max_delay 1, with a Poisson source that has no synapses,
  and with synapse types enabled only by the connections (no explicit
  CheckSynapseType), which the granule's run-time synapse mask must follow;
* the external drive (counter-based Poisson stream) on the granule BDF, HH
  BDF and Purkinje kernels;
* a spike log smaller than the run's spikes is reported as an error, and a
  zero-step run changes nothing.
"""
import numpy as np
import pytest

import edlut_amd as E
from edlut_amd import workloads
from edlut_amd.network import binomial_thresholds
from tests.helpers import bits, defaults, np_out_csr

pytestmark = pytest.mark.gpu

MODELS = [
    ("LIFTimeDrivenModel_GPU", "Euler", 1, 1e-4, 0.6),
    ("AdExTimeDrivenModel_GPU", "RK4", 1, 1e-4, 12.0),
    ("HHTimeDrivenModel_GPU", "BDF", 2, 1e-4, 3.0),
    ("EgidioGranuleCell_TimeDriven_GPU", "BDF", 2, 1e-4, 1.5),
    ("TimeDrivenPurkinjeCell_GPU", "RK2", 1, 1e-4, 1.0),
]


@pytest.mark.parametrize("n", [1, 513, 1537])
@pytest.mark.parametrize("model,integ,order,dt,wmax", MODELS)
def test_population_sizes(oracle, model, integ, order, dt, wmax, n):
    print(f"case sizes {model} {integ} n={n}", flush=True)
    rng = np.random.default_rng(n)
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, dt, order)})
    m.CheckSynapseType(0)
    m.CheckSynapseType(1)
    m.InitializeStates(n, 0, 0)
    o = oracle.Population(model, integ, defaults(model), n, dt_s=np.float32(dt), mask=3, bdf_order=order)
    total = 0
    for k in range(1, 201):
        ne = int(rng.integers(0, 4 * n + 2))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        typ = rng.choice(np.array([0, 0, 0, 1], np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        m.ProcessInputSpikes(tgt, typ, w)
        o.add_inputs(tgt, typ, w)
        m.UpdateState(step=k)
        ref = o.update(k)
        assert np.array_equal(m.GetInternalSpikeIndexs(), ref), f"update {k}"
        total += len(ref)
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"variable {v}"


def _edges(rng, n_src, n_tgt, k, max_delay, wmax, types=(0, 0, 0, 1)):
    src = rng.integers(0, n_src, n_tgt * k)
    tgt = np.repeat(np.arange(n_tgt), k)
    d = rng.integers(1, max_delay + 1, n_tgt * k)
    typ = rng.choice(np.array(types, np.uint8), n_tgt * k)
    w = rng.uniform(0.0, wmax, n_tgt * k).astype(np.float32)
    return src, tgt, d, typ, w


def _check_net(net, model, onet, opop, n_steps):
    net.run(n_steps)
    st, ix = net.spikes()
    ost, oix = onet.run(1, n_steps)
    assert np.array_equal(st, ost) and np.array_equal(ix, oix), "raster"
    for v in range(opop.nv):
        assert np.array_equal(bits(model.GetVectorNeuronState(v)), bits(opop.state(v))), f"variable {v}"
    return len(ost)


@pytest.mark.parametrize("model,integ,order,dt,wmax", [MODELS[0], MODELS[3], MODELS[4]])
def test_network_edges(oracle, model, integ, order, dt, wmax):
    """max_delay 1, a Poisson source without synapses, a Poisson source
    with synapses, types enabled by the connections only."""
    print(f"case network edges {model}", flush=True)
    rng = np.random.default_rng(7)
    n, D, steps = 700, 1, 400
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, dt, order)})
    m.InitializeStates(n, 0, 0)   # no CheckSynapseType: the connections enable AMPA and GABA
    net = E.Network(m, D)
    o = oracle.Population(model, integ, defaults(model), n, dt_s=np.float32(dt), mask=3, bdf_order=order)
    onet = oracle.Net(o, D)
    rec = np_out_csr(n, n, D, *_edges(rng, n, n, 8, D, wmax * 0.3))
    net.add_recurrent(*rec)
    onet.add_self(*rec)
    empty = (np.zeros(50 * D + 1, np.int64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    st0 = workloads.lcg_seeds(11, 50)
    net.add_poisson(st0, 40.0, *empty)
    onet.add_poisson(st0, 40.0, *empty)
    pcsr = np_out_csr(300, n, D, *_edges(rng, 300, n, 8, D, wmax))
    st1 = workloads.lcg_seeds(12, 300)
    net.add_poisson(st1, 120.0, *pcsr)
    onet.add_poisson(st1, 120.0, *pcsr)
    total = _check_net(net, m, onet, o, steps)
    assert total > 20


@pytest.mark.parametrize("model,integ,order,w_ext", [("EgidioGranuleCell_TimeDriven_GPU", "BDF", 2, 1.5),
                                                     ("HHTimeDrivenModel_GPU", "BDF", 1, 3.0),
                                                     ("TimeDrivenPurkinjeCell_GPU", "RK4", 1, 1.0)])
def test_external_drive_kernels(oracle, model, integ, order, w_ext):
    print(f"case external drive {model}", flush=True)
    n, D, steps = 900, 4, 300
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, 1e-4, order)})
    m.CheckSynapseType(0)
    m.InitializeStates(n, 0, 0)
    net = E.Network(m, D)
    o = oracle.Population(model, integ, defaults(model), n, mask=1, bdf_order=order)
    onet = oracle.Net(o, D)
    thr = binomial_thresholds(100, 40.0 * 1e-4)
    net.set_external(thr, w_ext, 99)
    onet.set_external(thr, w_ext, 99)
    total = _check_net(net, m, onet, o, steps)
    assert total > 0


def test_log_overflow_and_zero_steps(oracle):
    n = 400
    m = E.CreateNeuronModel("LIFTimeDrivenModel_GPU", {"int_meth": E.integration_method("Euler", 1e-4)})
    m.CheckSynapseType(0)
    m.CheckSynapseType(3, source_emits_spikes=False)
    m.InitializeStates(n, 0, 0)
    m.ProcessInputCurrent(np.full(n, 5.0, np.float32))
    net = E.Network(m, 2)
    v0 = m.GetVectorNeuronState(0).copy()
    net.run(0)
    assert np.array_equal(m.GetVectorNeuronState(0), v0)
    net.set_log_capacity(10)
    with pytest.raises(E.EDLUTException):
        net.run(300)
