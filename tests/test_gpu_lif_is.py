"""GPU parity of the SI-unit LIF (LIFTimeDrivenModel_IS) against the oracle
under Euler, RK2, RK4 and BDF2: spikes, refractory clock in seconds and the
pA -> nA external current, bit for bit."""
import zlib

import numpy as np
import pytest

import edlut_amd as E
from tests.helpers import bits, defaults

pytestmark = pytest.mark.gpu
LI = "LIFTimeDrivenModel_IS_GPU"


@pytest.mark.parametrize("integ,order", [("Euler", 1), ("RK2", 1), ("RK4", 1), ("BDF", 2)])
def test_lif_is_bit_exact(oracle, integ, order):
    n, steps = 1025, 600
    print(f"case LIF_IS {integ} order={order}", flush=True)
    rng = np.random.default_rng(zlib.crc32(f"lifis/{integ}/{order}".encode()))
    m = E.CreateNeuronModel(LI, {"int_meth": E.integration_method(integ, 1e-4, order)})
    for t in (0, 1, 2):
        m.CheckSynapseType(t)
    m.CheckSynapseType(3, source_emits_spikes=False)
    m.InitializeStates(n, 0, 0)
    o = oracle.Population(LI, integ, defaults(LI), n, mask=15, bdf_order=order)
    cur = rng.uniform(0.0, 8.0, n).astype(np.float32)   # pA
    m.ProcessInputCurrent(cur)
    o.set_ext_current(cur)
    total = 0
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, n))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        typ = rng.choice(np.array([0, 0, 1, 2], np.uint8), ne)
        w = rng.uniform(0.0, 0.6, ne).astype(np.float32)   # nS
        m.ProcessInputSpikes(tgt, typ, w)
        o.add_inputs(tgt, typ, w)
        m.UpdateState(step=k)
        ref = o.update(k)
        assert np.array_equal(m.GetInternalSpikeIndexs(), ref), f"update {k}"
        total += len(ref)
    assert total > 200
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v))), f"variable {v}"
