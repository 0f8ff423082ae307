"""Pins of the oracle (the CPU restatement the GPU is checked against).

* Look-up table and random generator: bit-exact against the reference's own
  sources, compiled where they lie into oracle/_ref (ExponentialTable.cpp,
  RandomGenerator.cpp).
* Integrators: closed forms on linear ODEs (truncated Taylor polynomials of
  exp(-x)) and the textbook RK4 on a non-autonomous ODE.
* Model equations: FP64 oracle against an independent high-accuracy solution
  of the published AdEx / conductance LIF equations (scipy), with the
  expected convergence order per integrator.
* LIF refractory clock: the double-precision clock of the reference,
  rebuilt step by step in Python.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.helpers import ADEX_DEFAULTS, LIF_DEFAULTS, np_out_csr

REF_DUMP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_dump")


# --------------------------------------------------------------------------
# against the reference's own compiled sources
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def refdump():
    if not os.path.exists(REF_DUMP):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    d = tempfile.mkdtemp()
    probe = np.concatenate([
        np.linspace(-25, 25, 200001, dtype=np.float32),
        np.array([-20.0, 20.0, np.nextafter(np.float32(-20), np.float32(-21)),
                  np.nextafter(np.float32(20), np.float32(21)), 6.0, 0.0, -0.0], np.float32),
        np.random.default_rng(0).uniform(-21, 21, 100000).astype(np.float32),
    ])
    probe.tofile(os.path.join(d, "exp_probe_in.f32"))
    subprocess.run([REF_DUMP, d], check=True, capture_output=True)
    return d, probe


def test_exp_table_bit_exact_vs_reference(oracle, refdump):
    d, _ = refdump
    ref = np.fromfile(os.path.join(d, "exp_table.f32"), np.float32)
    assert ref.shape == (16384,)
    assert np.array_equal(oracle.exp_table().view(np.uint32), ref.view(np.uint32))


def test_exp_table_check_discriminates(oracle, refdump):
    """A table built with the float exp would fail the check above: it
    differs from the reference's table in some entries."""
    d, _ = refdump
    ref = np.fromfile(os.path.join(d, "exp_table.f32"), np.float32)
    grid = np.float32(-20) + (np.float32(40) * np.arange(16384, dtype=np.float32)) / np.float32(16383)
    alt = np.exp(grid.astype(np.float32))  # single-precision exp
    assert np.any(alt.view(np.uint32) != ref.view(np.uint32))


def test_exp_lookup_bit_exact_vs_reference(oracle, refdump):
    d, probe = refdump
    ref = np.fromfile(os.path.join(d, "exp_probe.f32"), np.float32)
    got = oracle.exp_lut(probe)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_random_generator_bit_exact_vs_reference(oracle, refdump):
    d, _ = refdump
    seeds = np.fromfile(os.path.join(d, "lcg_seeds.u32"), np.uint32)
    assert np.array_equal(oracle.lcg_seeds(12345, 64), seeds)
    fr = np.fromfile(os.path.join(d, "lcg_frand.f32"), np.float32)
    assert np.array_equal(oracle.lcg_frand(int(seeds[0]), 4096).view(np.uint32), fr.view(np.uint32))


def test_lcg_closed_form(oracle):
    """x_{n+1} = (1103515245 x_n + 12345) mod 2^31; frand = float(x)/2^31."""
    x = 987654321
    vals = oracle.lcg_frand(x, 50)
    for v in vals:
        x = (1103515245 * x + 12345) % 2**31
        assert v == np.float32(np.float32(x) / np.float32(2**31))


def test_nmda_table_matches_float_evaluation(oracle):
    """g_nmda_inf = 1/(1 + expf(-0.062 V) * (1.2/3.57)) evaluated in float with a
    correctly rounded exp: identical up to the last bit of expf."""
    t = oracle.nmda_table(0.0, -80.0)
    f = np.float32
    v = f(-80.0) + (f(80.0) * np.arange(256, dtype=np.float32)) / f(255)
    e = np.exp((f(-0.062) * v).astype(np.float64)).astype(np.float32)  # correctly rounded expf
    g = f(1.0) / (f(1.0) + e * (f(1.2) / f(3.57)))
    diff = np.abs(t.view(np.int32).astype(np.int64) - g.view(np.int32).astype(np.int64))
    assert diff.max() <= 1 and np.mean(diff == 0) > 0.95


# --------------------------------------------------------------------------
# integrators: closed forms
# --------------------------------------------------------------------------
def _lif_linear(oracle, integ, fp64, v0, current):
    """LIF without conductances, refractoriness or threshold:
    dV/dt = (I + g_leak (E_leak - V)) / C."""
    p = LIF_DEFAULTS.copy()
    p[3] = 1e6  # v_thr out of reach
    pop = oracle.Population("LIFTimeDrivenModel_GPU", integ, p, len(v0), mask=8, fp64=fp64)
    pop.set_state(0, v0)
    pop.set_ext_current(current)
    pop.update(1)
    return pop.state(0), p


@pytest.mark.parametrize("integ,poly", [
    ("Euler", lambda x: 1 - x),
    ("RK2", lambda x: 1 - x + x**2 / 2),
    ("RK4", lambda x: 1 - x + x**2 / 2 - x**3 / 6 + x**4 / 24),
])
def test_integrators_linear_closed_form(oracle, integ, poly):
    v0 = np.linspace(-90, -40, 11)
    cur = np.linspace(-5, 5, 11).astype(np.float32)
    for fp64, tol in [(True, 1e-12), (False, 2e-6)]:
        v1, p = _lif_linear(oracle, integ, fp64, v0, cur)
        g, c, el = float(p[8]), float(p[4]), float(p[2])
        dt = float(np.float32(np.float32(1e-4) * np.float32(1000)))
        x = g / c * dt
        vinf = el + cur.astype(np.float64) / g
        expect = vinf + (v0 - vinf) * poly(x)
        assert np.allclose(v1, expect, rtol=tol, atol=0), (integ, fp64)


def test_integrators_wrong_coefficients_would_fail(oracle):
    """The closed-form check separates the three methods."""
    v0 = np.array([-70.0])
    cur = np.array([3.0], np.float32)
    r = {i: _lif_linear(oracle, i, True, v0, cur)[0][0] for i in ("Euler", "RK2", "RK4")}
    assert abs(r["Euler"] - r["RK2"]) > 1e-9 and abs(r["RK2"] - r["RK4"]) > 1e-12


def test_rk4_matches_textbook_rk4_with_decaying_conductance(oracle):
    """RK4 with exact half-step decays of g equals the classical RK4 on
    dV/dt = g0 exp(-t/tau) (E - V) / C (+ leak), one step.  The reference
    weights the stages with dt*0.166666666667f (RK4.h:178), a float constant
    3e-8 above 1/6; the FP64 oracle keeps it, so the textbook step uses it too."""
    p = LIF_DEFAULTS.copy()
    p[3] = 1e6
    n = 5
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "RK4", p, n, mask=1, fp64=True)
    v0 = np.linspace(-80, -50, n)
    g0 = np.linspace(0.1, 5.0, n)
    pop.set_state(0, v0)
    pop.set_state(1, g0)
    pop.update(1)
    dt = float(np.float32(np.float32(1e-4) * np.float32(1000)))
    e_exc, e_leak, c, tau, gl = 0.0, float(p[2]), float(p[4]), float(p[5]), float(p[8])

    def f(t, v, g0):
        return (g0 * np.exp(-t / tau) * (e_exc - v) + gl * (e_leak - v)) / c

    for i in range(n):
        v, g = v0[i], g0[i]
        k1 = f(0, v, g)
        k2 = f(dt / 2, v + dt / 2 * k1, g)
        k3 = f(dt / 2, v + dt / 2 * k2, g)
        k4 = f(dt, v + dt * k3, g)
        expect = v + (k1 + 2 * (k2 + k3) + k4) * (dt * float(np.float32(0.166666666667)))
        assert abs(pop.state(0)[i] - expect) < 1e-11
        assert abs(pop.state(1)[i] - g * np.exp(-dt / tau)) < 1e-12


# --------------------------------------------------------------------------
# model equations: FP64 oracle vs an independent solution
# --------------------------------------------------------------------------
def _adex_rhs(t, y, p, g0e, g0i):
    a, b, dT, vt, tw, ee, ei, er, el, gl, cm, te, ti, tn = [float(x) for x in p]
    v, w = y
    ge = g0e * np.exp(-t / te)
    gi = g0i * np.exp(-t / ti)
    dv = (gl * (el - v) + gl * dT * np.exp((v - vt) / dT) - w + ge * (ee - v) + gi * (ei - v)) / cm
    dw = (a * (v - el) - w) / tw
    return [dv, dw]


@pytest.mark.parametrize("integ,order,dts", [("Euler", 1, (0.1, 0.05)), ("RK2", 2, (0.1, 0.05)),
                                             ("RK4", 4, (0.8, 0.4))])
def test_adex_fp64_converges_to_published_equations(oracle, integ, order, dts):
    """Brette & Gerstner AdEx with exponentially decaying conductances; the
    FP64 oracle's error against scipy's DOP853 solution falls with the
    integrator's order when dt halves.  (RK4 is probed at larger steps: at
    0.1 ms its error sits below the ~1e-7 mV floor set by the reference's
    float 1/6 stage weight.)"""
    from scipy.integrate import solve_ivp

    T = 3.2  # ms, below threshold throughout
    v0, w0, g0e, g0i = -60.0, 5.0, 8.0, 2.0
    ref = solve_ivp(_adex_rhs, (0, T), [v0, w0], args=(ADEX_DEFAULTS, g0e, g0i), method="DOP853",
                    rtol=1e-13, atol=1e-13).y[:, -1]
    errs = []
    for dt_ms in dts:
        dt_s = np.float32(dt_ms * 1e-3)
        pop = oracle.Population("AdExTimeDrivenModel_GPU", integ, ADEX_DEFAULTS, 1, dt_s=dt_s, mask=3,
                                fp64=True)
        pop.set_state(0, [v0]), pop.set_state(1, [w0]), pop.set_state(2, [g0e]), pop.set_state(3, [g0i])
        dt_model = float(np.float32(dt_s * np.float32(1000)))
        steps = int(round(T / dt_model))
        for k in range(1, steps + 1):
            pop.update(k)
        # compare at the same final time (dt is float-rounded)
        t_end = steps * dt_model
        ref_t = solve_ivp(_adex_rhs, (0, t_end), [v0, w0], args=(ADEX_DEFAULTS, g0e, g0i),
                          method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
        errs.append(abs(pop.state(0)[0] - ref_t[0]))
    ratio = errs[0] / errs[1]
    assert errs[0] < {1: 0.3, 2: 1e-2, 4: 1e-3}[order], errs
    assert 2 ** order * 0.7 < ratio < 2 ** order * 1.4, (errs, ratio)
    assert ref is not None


def test_float_oracle_close_to_fp64_oracle(oracle):
    """Per-model float error: the float restatement (look-up-table exp) stays
    within 1e-3 relative of the FP64 oracle over 50 ms of driven AdEx RK4."""
    n = 64
    rng = np.random.default_rng(3)
    f32 = oracle.Population("AdExTimeDrivenModel_GPU", "RK4", ADEX_DEFAULTS, n, mask=3)
    f64 = oracle.Population("AdExTimeDrivenModel_GPU", "RK4", ADEX_DEFAULTS, n, mask=3, fp64=True)
    for k in range(1, 501):
        tgt = rng.integers(0, n, 8)
        typ = rng.integers(0, 2, 8).astype(np.uint8)
        w = rng.uniform(0, 1.0, 8).astype(np.float32)
        f32.add_inputs(tgt, typ, w)
        f64.add_inputs(tgt, typ, w)
        s32 = f32.update(k)
        s64 = f64.update(k)
        assert len(s32) == 0 and len(s64) == 0
    v32, v64 = f32.state(0), f64.state(0)
    assert np.max(np.abs(v32 - v64) / np.abs(v64)) < 1e-3


# --------------------------------------------------------------------------
# LIF refractory clock
# --------------------------------------------------------------------------
def test_lif_refractory_against_double_clock(oracle):
    """After a spike the reference clock restarts at 0 s and gains
    double(float(dt)) per update; the derivative is zero while
    clock*1000 + dt <= tau_ref, scaled by (clock*1000 + dt - tau_ref)/dt on the
    update that crosses tau_ref, full afterwards (LIFTimeDrivenModel.cpp:189-201)."""
    p = LIF_DEFAULTS.copy()
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "Euler", p, 1, mask=8, fp64=True)
    cur = np.array([50.0], np.float32)  # strong drive: spikes quickly
    pop.set_ext_current(cur)
    k = 0
    while True:
        k += 1
        if len(pop.update(k)):
            break
    el, gl, c = float(p[2]), float(p[8]), float(p[4])
    dt = float(np.float32(np.float32(1e-4) * np.float32(1000)))
    clock = 0.0
    v = el
    for j in range(30):
        k += 1
        lst = clock * 1000.0
        full = (50.0 + gl * (el - v)) / c
        if lst > float(p[7]):
            dv = full
        elif lst + dt > float(p[7]):
            dv = (lst + dt - float(p[7])) / dt * full
        else:
            dv = 0.0
        v = v + dt * dv
        spikes = pop.update(k)
        if len(spikes):
            break
        assert abs(pop.state(0)[0] - v) < 1e-9, j
        clock += float(np.float32(1e-4))


def test_threaded_oracle_equals_serial(oracle):
    """The oracle's population updates run on several OpenMP threads (the
    multi-core CPU baseline); the spikes and states are those of the serial
    loop, bit for bit."""
    from edlut_amd import workloads

    spec = workloads.adex_recurrent(n=6000, in_degree=50, steps=300, w_ext=1.0)
    out = []
    for threads in (1, 4):
        oracle.set_threads(threads)
        try:
            pop, net = oracle.from_spec(spec, ADEX_DEFAULTS, np_out_csr)
            st, ix = net.run(1, 300)
            out.append((st, ix, [pop.state(v) for v in range(4)]))
        finally:
            oracle.set_threads(1)
    assert len(out[0][0]) > 100
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for a, b in zip(out[0][2], out[1][2]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
