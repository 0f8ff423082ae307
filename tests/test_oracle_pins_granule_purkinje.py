"""Pins of the oracle's granule-cell and Purkinje-cell restatements.

What each check fixes independently of the restatement:

* log_cr (the granule's Nernst logarithm) equals the correctly rounded float
  logarithm (numpy's double log rounded to float) on every sampled input,
  subnormals and special values included; glibc's logf, which the reference
  calls, agrees with it except for rare one-ulp differences (measured here).
* The channel tables equal the published rate functions evaluated in double
  at the same float voltage grid, within the float evaluation error, and the
  steady states have the physiological shape (activation gates rise with V,
  inactivation gates fall, all within [0, 1]).
* The right-hand sides, read off one FP64 Euler step of the oracle:
  - every gate obeys dx/dt = (x_inf - x)/tau with x_inf and 1/tau from that
    gate's table columns, independent of every other gate;
  - the membrane equation is the chord-conductance equation: with the gates
    held, dV/dt is affine in V and vanishes at sum(g_k E_k)/sum(g_k), with
    each channel's conductance gmax * (gating product) and its reversal
    potential (the Nernst potential RT/(zF) ln(co/ci) for calcium);
  - calcium relaxes to its resting concentration at 1.5/ms when no calcium
    current flows, and at V equal to its Nernst potential.
* The Purkinje spike: the membrane potential follows the stereotyped
  waveform (linear rise from v_thr to 31 mV over tau_ref/2, linear fall to
  e_leak over the next tau_ref/2, e_leak on the first step after).
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import PURKINJE_DEFAULTS

GR = "EgidioGranuleCell_TimeDriven_GPU"
PK = "TimeDrivenPurkinjeCell_GPU"


# ---------------------------------------------------------------- logarithm
def _ulp_diff(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_log_cr_is_correctly_rounded(oracle):
    rng = np.random.default_rng(11)
    bits = rng.integers(1, 0x7F800000, 400000, dtype=np.int64).astype(np.uint32)
    x = bits.view(np.float32)
    x = np.concatenate([x, np.array([1.0, 2.0, 0.5, 1.4142135, 1.4142137, 0.70710677, 1e-45, 1.1754942e-38,
                                     1.1754944e-38, 3.4028235e38, 2.0 / 1e-4, 2.0 / 3e-4], np.float32)])
    got = oracle.log_cr(x)
    want = np.log(x.astype(np.float64)).astype(np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])
    sp = oracle.log_cr(np.array([0.0, np.inf, np.nan], np.float32))
    assert sp[0] == -np.inf and sp[1] == np.inf and np.isnan(sp[2])
    # a dropped low part of ln 2 or a series cut short changes results
    x2 = np.float32(2.0) ** np.arange(-126, 127, dtype=np.float32)
    assert np.array_equal(oracle.log_cr(x2), np.log(x2.astype(np.float64)).astype(np.float32))


def test_log_cr_against_glibc_logf(oracle):
    """The reference's logf (glibc) differs from the correctly rounded value
    by at most one ulp, rarely, over the granule's Nernst arguments 2/Ca."""
    libm = ctypes.CDLL("libm.so.6")
    libm.logf.restype = ctypes.c_float
    libm.logf.argtypes = [ctypes.c_float]
    rng = np.random.default_rng(3)
    ca = np.exp(rng.uniform(np.log(1e-5), np.log(1e-2), 60000)).astype(np.float32)
    x = (np.float32(2.0) / ca).astype(np.float32)
    glibc = np.array([libm.logf(float(v)) for v in x], np.float32)
    mine = oracle.log_cr(x)
    d = _ulp_diff(glibc, mine)
    assert d.max() <= 1
    frac = float(np.mean(d == 0))
    print(f"glibc logf equals the correctly rounded logf on {frac:.6f} of {len(x)} Nernst arguments")
    assert frac > 0.999


# ---------------------------------------------------------------- tables
def _linoid(x, y):
    return np.where(np.abs(x / y) < 1e-6, y * (1 - x / y / 2), x / np.expm1(x / y))


def _cancel(*xy):
    """Relative-error allowance of the float x/(exp(x/y)-1): exp(x/y)-1 loses
    about ulp(1)/|x/y| of its relative accuracy near x = 0."""
    return sum(6e-8 / np.maximum(np.abs(x / y), 1e-6) for x, y in xy)


def _granule_rates(V):
    """alpha/beta of every granule gate in double (EgidioGranuleCell_TimeDriven.h:402-552)."""
    q = lambda t0: 3.0 ** ((30.0 - t0) / 10.0)
    q20, q22, q30, q63 = q(20.0), q(22.0), q(30.0), q(6.3)
    e = np.exp
    r = {}
    r["xNa_f"] = (q20 * -0.3 * _linoid(V + 19.0, -10.0), q20 * 12.0 * e(-(V + 44.0) / 18.182))
    r["yNa_f"] = (q20 * 0.105 * e(-(V + 44.0) / 3.333), q20 * 1.5 / (1.0 + e(-(V + 11.0) / 5.0)))
    r["xNa_r"] = (q20 * (0.00008 - 0.00493 * _linoid(V - 4.48754, -6.81881)),
                  q20 * (0.04752 + 0.01558 * _linoid(V + 43.97494, 0.10818)))
    r["yNa_r"] = (q20 * 0.31836 * e(-(V + 80.0) / 62.52621), q20 * 0.01014 * e((V + 83.3332) / 16.05379))
    r["xK_V"] = (q63 * -0.01 * _linoid(V + 25.0, -10.0), q63 * 0.125 * e(-0.0125 * (V + 35.0)))
    r["xK_IR"] = (q20 * 0.13289 * e(-(V + 83.94) / 24.3902), q20 * 0.16994 * e((V + 83.94) / 35.714))
    r["xCa"] = (q20 * 0.04944 * e((V + 29.06) / 15.87301587302), q20 * 0.08298 * e(-(V + 18.66) / 25.641))
    r["yCa"] = (q20 * 0.0013 * e(-(V + 48.0) / 18.183), q20 * 0.0013 * e((V + 48.0) / 83.33))
    ss = {
        "xNa_p": (1.0 / (1.0 + e(-(V + 42.0) / 5.0)),
                  (q30 * -0.091 * _linoid(V + 42.0, -5.0) + q30 * 0.062 * _linoid(V + 42.0, 5.0)) * 0.2),
        "xK_A": (1.0 / (1.0 + e((V + 46.7) / -19.8)),
                 q20 * 4.88826 / (1 + e(-(V + 9.17203) / 23.32708)) + q20 * 0.99285 / e((V + 18.27914) / 19.47175)),
        "yK_A": (1.0 / (1.0 + e((V + 78.8) / 8.4)),
                 q20 * 0.11042 / (1.0 + e((V + 111.33209) / 12.8433)) + q20 * 0.10353 / (1.0 + e(-(V + 49.9537) / 8.90123))),
        "xK_sl": (1.0 / (1.0 + e(-(V + 30.0) / 6.0)),
                  q22 * 0.0033 * e((V + 30.0) / 40.0) + q22 * 0.0033 * (-(V + 30.0) / 20.0)),
    }
    cancel = {"xNa_f": _cancel((V + 19.0, -10.0)), "xNa_r": _cancel((V - 4.48754, -6.81881), (V + 43.97494, 0.10818)),
              "xK_V": _cancel((V + 25.0, -10.0)), "xNa_p": _cancel((V + 42.0, -5.0), (V + 42.0, 5.0))}
    return r, ss, cancel


# table columns of each gate: (x_inf, 1/tau)
GR_COLS = {"xNa_f": 0, "yNa_f": 2, "xNa_r": 4, "yNa_r": 6, "xNa_p": 8, "xK_V": 10, "xK_A": 12, "yK_A": 14,
           "xK_IR": 16, "xCa": 20, "yCa": 22, "xK_sl": 24}
# state index of each gate
GR_VAR = {"xNa_f": 1, "yNa_f": 2, "xNa_r": 3, "yNa_r": 4, "xNa_p": 5, "xK_V": 6, "xK_A": 7, "yK_A": 8,
          "xK_IR": 9, "xK_Ca": 10, "xCa": 11, "yCa": 12, "xK_sl": 13}


def _grid(lo, hi):
    i = np.arange(1024)
    return (np.float32(lo) + (np.float32(hi - lo) * i.astype(np.float32)) / np.float32(1023)).astype(np.float32)


def test_granule_table_matches_rate_functions(oracle):
    t = oracle.granule_table().astype(np.float64)
    V = _grid(-100.0, 50.0).astype(np.float64)
    with np.errstate(over="ignore"):
        rates, ss, cancel = _granule_rates(V)
    for g, (a, b) in rates.items():
        c = GR_COLS[g]
        s = np.abs(a) + np.abs(b)
        rel = 2e-6 + cancel.get(g, 0.0)
        assert np.all(np.abs(t[:, c + 1] - (a + b)) <= rel * s + 1e-30), g
        assert np.all(np.abs(t[:, c] - a / (a + b)) <= rel * s / np.abs(a + b) + 1e-7), g
    for g, (inf, itau) in ss.items():
        c = GR_COLS[g]
        rel = 2e-6 + cancel.get(g, 0.0)
        assert np.all(np.abs(t[:, c] - inf) <= 2e-6 * inf + 1e-30), g
        assert np.all(np.abs(t[:, c + 1] - itau) <= rel * np.abs(itau) + 1e-6), g
    aux = 0.0015 * np.exp(-V / 11.765)
    assert np.allclose(t[:, 18], aux, rtol=2e-6, atol=0)
    assert np.allclose(t[:, 19], 1.0 / (0.00015 * np.exp(-V / 11.765)), rtol=2e-6, atol=0)
    # physiology: steady states in [0, 1]; activation rises and inactivation
    # falls with V (inward rectifier: falls); a swapped column fails this
    for g in ("xNa_f", "xNa_p", "xK_V", "xK_A", "xCa", "xK_sl"):
        x = t[:, GR_COLS[g]]
        assert x.min() >= 0 and x.max() <= 1 and np.all(np.diff(x) >= -1e-7) and x[-1] - x[0] > 0.5, g
    for g in ("yNa_f", "yK_A", "yCa", "xK_IR"):
        x = t[:, GR_COLS[g]]
        assert x.min() >= 0 and x.max() <= 1 and np.all(np.diff(x) <= 1e-7) and x[0] - x[-1] > 0.5, g


def test_purkinje_table_matches_rate_functions(oracle):
    t = oracle.purkinje_table().astype(np.float64)
    V = _grid(-100.0, 35.0).astype(np.float64)
    a_ca = 1.6 / (1 + np.exp(-0.072 * (V - 5.0)))
    b_ca = 0.02 * (V + 8.9) / np.expm1((V + 8.9) * 0.2)
    a_m = 0.3 / (1 + np.exp((-V - 2.0) * 0.2))
    b_m = 0.001 * np.exp((-V - 60.0) * 0.055555555555555)
    assert np.allclose(t[:, 0], a_ca, rtol=2e-6, atol=0)
    assert np.allclose(t[:, 1], a_ca + b_ca, rtol=1e-5, atol=1e-7)
    assert np.allclose(t[:, 2], a_m, rtol=2e-6, atol=0)
    assert np.allclose(t[:, 3], a_m + b_m, rtol=2e-6, atol=0)
    c_inf, m_inf = t[:, 0] / t[:, 1], t[:, 2] / t[:, 3]
    assert np.all(np.diff(c_inf) > -1e-7) and np.all(np.diff(m_inf) > -1e-7)
    assert c_inf[0] < 0.01 and c_inf[-1] > 0.9 and m_inf[0] < 0.01 and m_inf[-1] > 0.9


# ---------------------------------------------------------------- right-hand sides
DT_S = np.float32(1e-9)
DT_MS = np.float64(np.float32(DT_S * np.float32(1000.0)))


def _deriv(oracle, model, params, states, nv, mask=1 | 2 | 8):
    """f(S) of the FP64 oracle for each row of `states` [n, nv], from one
    Euler step of length DT: f = (S1 - S0)/DT."""
    n = states.shape[0]
    p = oracle.Population(model, "Euler", params, n, dt_s=DT_S, mask=mask, fp64=True)
    for v in range(nv):
        p.set_state(v, states[:, v])
    p.update(1)
    s1 = np.stack([p.state(v) for v in range(nv)], axis=1)
    return (s1 - states) / DT_MS


def _granule_state(rng, n, V):
    s = np.zeros((n, 19))
    s[:, 0] = V
    s[:, 1:14] = rng.uniform(0.05, 0.95, (n, 13))
    s[:, 14] = rng.uniform(5e-5, 5e-4, n)
    return s


def test_granule_gates_relax_to_their_table_steady_states(oracle):
    rng = np.random.default_rng(7)
    t = oracle.granule_table().astype(np.float64)
    for V in (-75.0, -52.3, -20.4, 10.7):
        base = _granule_state(rng, 1, V)
        pos = int(np.float32((np.float32(V) - np.float32(-100.0)) * np.float32(1023.0 / 150.0)))
        s = np.repeat(base, 3, axis=0)
        for g, c in GR_COLS.items():
            j = GR_VAR[g]
            s = np.repeat(base, 3, axis=0)
            s[0, j], s[1, j], s[2, j] = 0.0, 1.0, 0.37
            f = _deriv(oracle, GR, [], s, 19)[:, j]
            itau = f[0] - f[1]
            x_inf = f[0] / itau
            assert abs(itau - t[pos, c + 1]) <= 1e-7 * abs(t[pos, c + 1]) + 1e-9, (g, V)
            assert abs(x_inf - t[pos, c]) <= 1e-7, (g, V)
            assert abs(f[2] - (x_inf - 0.37) * itau) <= 1e-7 * abs(itau), (g, V)
            # no other gate's derivative moved
            others = [GR_VAR[h] for h in GR_COLS if h != g]
            assert np.allclose(f.reshape(3, 1) * 0 + _deriv(oracle, GR, [], s, 19)[:, others],
                               _deriv(oracle, GR, [], np.repeat(base, 3, axis=0), 19)[:, others], rtol=1e-9, atol=1e-9)
        # the Ca-dependent K gate: Hill-type rates of Ca with the table's V factors
        ca = base[0, 14]
        a = 2.5 / (1.0 + t[pos, 18] / ca)
        b = 1.5 / (1.0 + ca * t[pos, 19])
        s = np.repeat(base, 2, axis=0)
        s[0, 10], s[1, 10] = 0.0, 1.0
        f = _deriv(oracle, GR, [], s, 19)[:, 10]
        assert abs((f[0] - f[1]) - (a + b)) <= 1e-7 * (a + b) and abs(f[0] / (f[0] - f[1]) - a / (a + b)) <= 1e-7


def _nernst_ca(ca):
    # RT/(zF) ln(co/ci) in mV, z = 2, 30 C, co = 2 mM
    return 1000.0 * 8.3134 * (30.0 + 273.15) / (2.0 * 96485.309) * np.log(2.0 / ca)


def test_granule_membrane_is_the_chord_conductance_equation(oracle):
    rng = np.random.default_rng(5)
    gmax = dict(Na_f=0.013, Na_r=0.0005, Na_p=0.0002, K_V=0.003, K_A=0.004, K_IR=0.0009, K_Ca=0.004,
                Ca=0.00046, K_sl=0.00035)
    # (reversal, gating product as a function of the state)
    ch = {
        "Na_f": (87.39, lambda s: s[1] ** 3 * s[2]), "Na_r": (87.39, lambda s: s[3] * s[4]),
        "Na_p": (87.39, lambda s: s[5]), "K_V": (-84.69, lambda s: s[6] ** 4),
        "K_A": (-84.69, lambda s: s[7] ** 3 * s[8]), "K_IR": (-84.69, lambda s: s[9]),
        "K_Ca": (-84.69, lambda s: s[10]), "Ca": (None, lambda s: s[11] ** 2 * s[12]),
        "K_sl": (-84.69, lambda s: s[13]),
    }
    for trial in range(6):
        s0 = _granule_state(rng, 1, -60.0)[0]
        g = [gmax[k] * f(s0) for k, (e, f) in ch.items()] + [5.68e-5, 2.17e-5]
        e_ca = _nernst_ca(s0[14])
        E = [e if e is not None else e_ca for k, (e, f) in ch.items()] + [-58.0, -65.0]
        s = np.repeat(s0[None, :], 2, axis=0)
        s[0, 0], s[1, 0] = -90.0, 30.0
        f = _deriv(oracle, GR, [], s, 19)[:, 0]
        slope = (f[1] - f[0]) / 120.0
        v_zero = -90.0 - f[0] / slope
        G = np.sum(g)
        # constants are float in the model: 2e-7 relative
        assert abs(slope + G / 1e-3) <= 2e-7 * G / 1e-3
        assert abs(v_zero - np.dot(g, E) / G) <= 2e-5
        # external current: 1 pA adds 3.34156941e-4 / Cm per ms
        s[:, 18] = 7.0
        f2 = _deriv(oracle, GR, [], s, 19)[:, 0]
        assert np.allclose(f2 - f, 7.0 * 3.34156941e-4 / 1e-3, rtol=2e-7)


def test_granule_calcium_buffering(oracle):
    rng = np.random.default_rng(9)
    s = _granule_state(rng, 3, -50.0)
    s[0, 11] = 0.0                       # no calcium current
    s[1, 0] = _nernst_ca(s[1, 14])       # at the Nernst potential
    f = _deriv(oracle, GR, [], s, 19)[:, 14]
    ca0 = float(np.float32(1e-4))   # the model's float resting concentration
    assert abs(f[0] + 1.5 * (s[0, 14] - ca0)) <= 1e-13
    # the Nernst potential in double against the model's float constants
    assert abs(f[1] + 1.5 * (s[1, 14] - ca0)) <= 1e-8
    # inward calcium current (V below the Nernst potential) raises Ca faster
    assert f[2] > -1.5 * (s[2, 14] - ca0) + 1e-6


def test_purkinje_membrane_and_gates(oracle):
    rng = np.random.default_rng(2)
    t = oracle.purkinje_table().astype(np.float64)
    for V in (-68.0, -52.5, -40.1):
        s = np.zeros((2, 7))
        s[:, 1] = rng.uniform(0.05, 0.9)
        s[:, 2] = rng.uniform(0.05, 0.9)
        s[0, 0], s[1, 0] = V, V + 1e-3
        f = _deriv(oracle, PK, PURKINJE_DEFAULTS, s, 7)
        ca, M = s[0, 1], s[0, 2]
        g = np.array([0.02, 0.001 * ca * ca, 0.75 * M])
        E = np.array([-70.0, 125.0, -95.0])
        want = np.sum(g * (E - V)) / 0.95
        assert abs(f[0, 0] - want) <= 2e-7 * np.sum(np.abs(g * (E - V))) / 0.95   # float constants
        pos = int(np.float32((np.float32(V) + np.float32(100.0)) * np.float32(1023.0 / 135.0)))
        assert abs(f[0, 1] - (t[pos, 0] - ca * t[pos, 1])) <= 1e-9
        assert abs(f[0, 2] - (t[pos, 2] - M * t[pos, 3])) <= 1e-9


def test_purkinje_spike_waveform(oracle):
    """After each spike V follows the stereotyped waveform: a linear rise from
    v_thr to 31 mV over tau_ref/2, a linear fall to e_leak over the next
    tau_ref/2, e_leak at the first update past tau_ref."""
    p = PURKINJE_DEFAULTS
    v_thr, e_leak, tau_ref = float(p[2]), float(p[3]), float(p[7])
    dt_ms = 0.05
    n = 4
    o = oracle.Population(PK, "RK4", p, n, dt_s=np.float32(dt_ms * 1e-3), mask=8)
    o.set_ext_current(np.array([0.0, 150.0, 300.0, 600.0], np.float32))
    V = []
    fired = []
    for k in range(1, 4001):
        sp = o.update(k)
        V.append(o.state(0).copy())
        fired.append(sp)
    V = np.array(V)
    half = tau_ref / 2
    checked = 0
    for i in (1, 2, 3):
        ks = [k for k in range(len(fired)) if i in fired[k]]
        assert len(ks) >= 3, "the current step should make the cell fire repeatedly"
        for k0 in ks[:-1]:
            for j in range(0, int(tau_ref / dt_ms) + 2):
                tj = j * dt_ms
                if k0 + j >= len(V):
                    break
                if tj <= half - 1e-9:
                    want = v_thr + (31.0 - v_thr) * tj / half
                elif tj < tau_ref - 1e-9:
                    want = 31.0 - (31.0 - e_leak) * (tj - half) / half
                elif tj - tau_ref < dt_ms - 1e-9:
                    want = e_leak
                else:
                    break
                assert abs(V[k0 + j, i] - want) <= 1e-4, (i, k0, j, V[k0 + j, i], want)
                checked += 1
    assert checked > 100
    assert np.all(V[:, 0] < v_thr), "no input, no spike"


@pytest.mark.parametrize("model,integ,order,dt,steps,cur", [
    (GR, "BDF", 2, 1e-4, 2000, 5.0),
    (GR, "RK4", 1, 2e-6, 20000, 5.0),
    (PK, "RK4", 1, 1e-4, 2000, 200.0),
    (PK, "BDF", 2, 1e-4, 2000, 200.0),
])
def test_fp64_error(oracle, model, integ, order, dt, steps, cur):
    """Distance of the float restatement from the FP64 oracle (same
    discretisation, double arithmetic, exact logarithm): spike rasters and
    the largest relative V error below the threshold region."""
    from tests.helpers import defaults
    n = 16
    res = []
    for fp64 in (False, True):
        o = oracle.Population(model, integ, defaults(model), n, dt_s=np.float32(dt), mask=8, fp64=fp64,
                              bdf_order=order)
        o.set_ext_current(np.linspace(0, cur, n).astype(np.float32))
        V, spk = [], []
        for k in range(1, steps + 1):
            spk.append(tuple(o.update(k)))
            V.append(o.state(0).copy())
        res.append((np.array(V), spk))
    (vf, sf), (vd, sd) = res
    first = next((k for k in range(steps) if sf[k] != sd[k]), steps)
    sub = vd[:first] < (-50.0 if model == GR else -40.0)
    err = np.max(np.abs(vf[:first] - vd[:first])[sub] / np.abs(vd[:first][sub])) if sub.any() else 0.0
    print(f"{model} {integ}{order}: rasters equal for {first}/{steps} steps, max rel V error {err:.2e}")
    # the BDF Newton stops at an absolute step of 1e-6 in both, so float and
    # double take different iteration counts; near threshold that grows
    assert err < 5e-3
