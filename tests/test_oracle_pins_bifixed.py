"""Pins of the oracle's Bifixed integrators (Bifixed_{Euler,RK2,RK4}_GPU2.cuh).

The reference's GPU Bifixed methods take n_steps sub-steps of the base method
in every update (Bifixed_RK4_GPU2.cuh:100-150): sub-step seconds dt/n in
float (BifixedStep_GPU_C_INTERFACE.cuh:157-158), decay factors of the
sub-step (Bifixed_RK4_GPU2.cuh:237-241), the spike clock and the spike
condition after every sub-step.  So a Bifixed population at (dt, n) must
equal the base method at dt/n updated n times per update, when the inputs
arrive only before every n-th of those updates -- the base method being the
oracle's fixed-step restatement, pinned against the reference's CPU models
elsewhere.  A neuron is reported once per update (the reference raises a
per-neuron flag), so the raster of an update is the sorted union of the n
sub-updates' spikes.  Further pins: n = 1 is the base method itself, and the
sub-steps shrink the error against the closed-form LIF decay.
"""
import zlib

import numpy as np
import pytest

from tests.helpers import bits, defaults

# model, input weight scale, external current (or None), float dt
CASES = [
    ("LIFTimeDrivenModel_GPU", 0.6, 4.0, 1e-4),
    ("AdExTimeDrivenModel_GPU", 12.0, None, 1e-4),
    ("HHTimeDrivenModel_GPU", 3.0, None, 1e-4),
    ("EgidioGranuleCell_TimeDriven_GPU", 1.5, None, 2e-5),
    ("TimeDrivenPurkinjeCell_GPU", 1.0, 300.0, 1e-4),
    ("IzhikevichTimeDrivenModel_GPU", 2.0, 300.0, 1e-4),
    ("LIFTimeDrivenModel_IS_GPU", 0.6, 6.0, 1e-4),
]


def _pair(oracle, model, base, n_steps, n, dt, cur):
    mask = 7 | (8 if cur is not None else 0)
    b = oracle.Population(model, "Bifixed_" + base, defaults(model), n, dt_s=np.float32(dt), mask=mask,
                          n_steps=n_steps)
    sub = np.float32(np.float32(dt) / np.float32(n_steps))
    f = oracle.Population(model, base, defaults(model), n, dt_s=sub, mask=mask)
    if cur is not None:
        c = np.full(n, cur, np.float32)
        b.set_ext_current(c)
        f.set_ext_current(c)
    return b, f


@pytest.mark.parametrize("base", ["Euler", "RK2", "RK4"])
@pytest.mark.parametrize("model,wmax,cur,dt", CASES)
def test_bifixed_is_base_method_at_substep(oracle, model, wmax, cur, dt, base):
    if model == "EgidioGranuleCell_TimeDriven_GPU" and base != "RK4":
        pytest.skip("explicit Euler/RK2 are unstable for the granule's Na gate at these steps")
    n, n_steps, steps = 200, 3, 300
    rng = np.random.default_rng(zlib.crc32(f"bif/{model}/{base}".encode()))
    b, f = _pair(oracle, model, base, n_steps, n, dt, cur)
    total = 0
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, 2 * n))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        typ = rng.choice(np.array([0, 0, 0, 1, 2], np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        b.add_inputs(tgt, typ, w)
        f.add_inputs(tgt, typ, w)
        got = b.update(k)
        union = set()
        for j in range(n_steps):
            union.update(f.update((k - 1) * n_steps + j + 1).tolist())
        assert np.array_equal(got, np.array(sorted(union), np.int32)), f"update {k}"
        total += len(got)
    for v in range(b.nv):
        assert np.array_equal(bits(b.state(v)), bits(f.state(v))), f"variable {v}"
    assert total > 0


@pytest.mark.parametrize("base", ["Euler", "RK2", "RK4"])
def test_one_substep_is_the_base_method(oracle, base):
    model, n = "AdExTimeDrivenModel_GPU", 300
    rng = np.random.default_rng(5)
    b = oracle.Population(model, "Bifixed_" + base, defaults(model), n, mask=3, n_steps=1)
    f = oracle.Population(model, base, defaults(model), n, mask=3)
    for k in range(1, 201):
        ne = int(rng.integers(0, n))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        w = rng.uniform(0.0, 12.0, ne).astype(np.float32)
        typ = np.zeros(ne, np.uint8)
        b.add_inputs(tgt, typ, w)
        f.add_inputs(tgt, typ, w)
        assert np.array_equal(b.update(k), f.update(k))
    assert np.array_equal(bits(b.state(0)), bits(f.state(0)))


def test_substeps_converge_to_closed_form(oracle):
    """LIF with a constant current and no synapses: V relaxes to e_leak + I/g
    as exp(-t g/c).  Euler's error at a fixed time falls ~1/n with n sub-steps."""
    model = "LIFTimeDrivenModel_GPU"
    p = defaults(model)
    e_leak, c_m, g_leak = float(p[2]), float(p[4]), float(p[8])
    cur, steps, dt = 1.5, 50, 1e-4   # 5 ms; V stays below threshold
    v_inf = e_leak + cur / g_leak
    t_ms = steps * float(np.float32(dt)) * 1000.0
    exact = v_inf + (e_leak - v_inf) * np.exp(-t_ms * g_leak / c_m)
    errs = []
    for n_steps in (1, 4, 16):
        b = oracle.Population(model, "Bifixed_Euler", p, 1, dt_s=np.float32(dt), mask=8, n_steps=n_steps)
        b.set_ext_current(np.array([cur], np.float32))
        for k in range(1, steps + 1):
            assert len(b.update(k)) == 0
        errs.append(abs(b.state(0)[0] - exact))
    assert errs[0] > 2.5 * errs[1] > 2.5 * 2.5 * errs[2], errs
    assert errs[0] / errs[1] == pytest.approx(4.0, rel=0.15)
