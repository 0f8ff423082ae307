"""GPU parity of the HH model and the BDF integrator against the oracle.

HH runs under Euler, RK2, RK4 and BDF orders 1-6; AdEx and LIF under BDF2.
Every update's spike list and the final state must be bit-identical to the
oracle (float Newton).  The FP64 Newton mode of HH is checked bit-exactly
against the oracle's FP64 variant with the state rounded to float after
each step, which is that mode's definition.
"""
import zlib

import numpy as np
import pytest

import edlut_amd as E
from tests.helpers import bits, defaults

pytestmark = pytest.mark.gpu


def _run(oracle, model, integ, order, n, steps, fp64_newton=False, wmax=None, cur_scale=None):
    rng = np.random.default_rng(zlib.crc32(f"{model}/{integ}/{order}/{fp64_newton}".encode()))
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, 1e-4, order)})
    for t in (0, 1, 2):
        assert m.CheckSynapseType(t)
    m.InitializeStates(n, 0, 0)
    if fp64_newton:
        m.SetNewtonFP64(True)
    o = oracle.Population(model, integ, defaults(model), n, mask=7 | (8 if cur_scale else 0),
                          fp64=2 if fp64_newton else False, bdf_order=order)
    if cur_scale:
        cur = rng.uniform(0.0, 1.0, n).astype(np.float32) * cur_scale
        m.ProcessInputCurrent(cur)
        o.set_ext_current(cur)
    total = 0
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, 3 * n))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        tgt[: ne // 8] = 3
        typ = rng.choice(np.array([0, 0, 1, 2], np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        m.ProcessInputSpikes(tgt, typ, w)
        o.add_inputs(tgt, typ, w)
        m.UpdateState(step=k)
        ref = o.update(k)
        assert np.array_equal(m.GetInternalSpikeIndexs(), ref), f"update {k}"
        total += len(ref)
    for v in range(o.nv):
        assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(o.state(v).astype(np.float32))), f"var {v}"
    return total


@pytest.mark.parametrize("integ,order", [("Euler", 1), ("RK2", 1), ("RK4", 1), ("BDF", 1), ("BDF", 2),
                                         ("BDF", 3), ("BDF", 4), ("BDF", 5), ("BDF", 6)])
def test_hh_bit_exact(oracle, integ, order):
    print(f"case HH {integ} order={order}", flush=True)
    total = _run(oracle, "HHTimeDrivenModel_GPU", integ, order, 513, 400, wmax=3.0, cur_scale=400.0)
    assert total > 0


@pytest.mark.parametrize("model,wmax", [("AdExTimeDrivenModel_GPU", 12.0), ("LIFTimeDrivenModel_GPU", 0.6)])
def test_bdf2_adex_lif_bit_exact(oracle, model, wmax):
    print(f"case {model} BDF2", flush=True)
    total = _run(oracle, model, "BDF", 2, 1025, 300, wmax=wmax)
    assert total > 0


@pytest.mark.parametrize("order", [2, 6])
def test_hh_fp64_newton_bit_exact(oracle, order):
    print(f"case HH BDF{order} fp64 Newton", flush=True)
    total = _run(oracle, "HHTimeDrivenModel_GPU", "BDF", order, 257, 400, fp64_newton=True, wmax=3.0,
                 cur_scale=400.0)
    assert total > 0


def test_fp64_newton_only_for_hh_bdf():
    m = E.AdExTimeDrivenModel_GPU({"int_meth": E.integration_method("BDF", 1e-4, 2)})
    m.InitializeStates(10, 0, 0)
    with pytest.raises(E.EDLUTException):
        m.SetNewtonFP64(True) if hasattr(m, "SetNewtonFP64") else E._capi.check(
            E.lib().edlut_population_set_newton_fp64(m.handle, 1))
