"""Pin of the oracle's SI-unit LIF: it is the millisecond LIF in other units.

With V in volts, t in seconds, capacitance in farads and conductance in
siemens, the same neuron (c_m 2 pF, g_leak 0.2 nS, E in mV / 1000, tau in
ms / 1000, the same nS weights and pA currents) must follow the same
trajectory: in the FP64 oracles V_IS * 1000 equals V_ms to rounding, and the
spike trains (refractory clock included) are the same.
"""
import numpy as np

from tests.helpers import LIF_DEFAULTS, LIF_IS_DEFAULTS


def test_si_lif_is_the_ms_lif(oracle):
    n, steps = 64, 3000
    rng = np.random.default_rng(4)
    ms = oracle.Population("LIFTimeDrivenModel_GPU", "RK4", LIF_DEFAULTS, n, mask=1 | 2 | 8, fp64=True)
    si = oracle.Population("LIFTimeDrivenModel_IS_GPU", "RK4", LIF_IS_DEFAULTS, n, mask=1 | 2 | 8, fp64=True)
    cur = rng.uniform(0.0, 6.0, n).astype(np.float32)
    ms.set_ext_current(cur)                       # ms LIF: pA directly
    si.set_ext_current(cur)                       # SI LIF: pA, taken as nA x 1000
    s_ms, s_si, worst = 0, 0, 0.0
    for k in range(1, steps + 1):
        ne = int(rng.integers(0, 40))
        tgt = rng.integers(0, n, ne).astype(np.int32)
        typ = rng.choice(np.array([0, 0, 1], np.uint8), ne)
        w = rng.uniform(0.0, 0.4, ne).astype(np.float32)
        ms.add_inputs(tgt, typ, w)
        si.add_inputs(tgt, typ, w)
        a = ms.update(k)
        b = si.update(k)
        assert np.array_equal(a, b), f"update {k}"
        s_ms += len(a)
        v_ms = ms.state(0)
        worst = max(worst, float(np.max(np.abs(si.state(0) * 1000.0 - v_ms) / np.abs(v_ms))))
    assert s_ms > 100
    # the float parameters (e.g. -65e-3f vs -65f / 1000) differ by ~3e-8
    assert worst < 1e-6
