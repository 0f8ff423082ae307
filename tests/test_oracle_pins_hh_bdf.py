"""Pins of the oracle's HH model and BDF integrator against independent math.

  * HH channel rates: the Traub-Miles rate functions (alpha/beta of m, h, n
    with the v_thr shift), evaluated in double at the table's grid points;
  * HH initial gates: steady state alpha / (alpha + beta) at e_leak;
  * HH right-hand side: one explicit Euler step of the FP64 oracle equals
    S0 + dt f(S0) with f written out from the published equations;
  * BDF: on a linear ODE (LIF leak plus a constant current) the corrector is
    solved exactly, so the FP64 oracle must follow the textbook BDF-k
    recursion (orders rising 1, 1, 2, ... up to bdf_order);
  * BDF on HH: the FP64 oracle's BDF1 step satisfies the backward-Euler
    equation S1 = S0 + dt f(S1) to the Newton tolerance.
"""
import numpy as np
import pytest

from tests.helpers import HH_DEFAULTS, LIF_DEFAULTS

N_TAB = 262144
V_MIN, V_MAX = np.float32(-90.0), np.float32(50.0)
AUX = np.float32(np.float32(N_TAB - 1) / (V_MAX - V_MIN))
DT_MODEL = float(np.float32(np.float32(1e-4) * np.float32(1000)))


def traub_rates(V, vt):
    """alpha_m, beta_m, alpha_h, beta_h, alpha_n, beta_n (1/ms) of the Traub
    model with threshold shift vt, in double."""
    V = np.asarray(V, np.float64)
    x_am = 13.0 - V + vt
    x_bm = V - vt - 40.0
    x_an = 15.0 - V + vt
    with np.errstate(divide="ignore", invalid="ignore"):
        am = np.where(x_am == 0, 1.28, 0.32 * x_am / np.expm1(x_am / 4.0))
        bm = np.where(x_bm == 0, 1.4, 0.28 * x_bm / np.expm1(x_bm / 5.0))
        an = np.where(x_an == 0, 0.16, 0.032 * x_an / np.expm1(x_an / 5.0))
    ah = 0.128 * np.exp((17.0 - V + vt) / 18.0)
    bh = 4.0 / (1.0 + np.exp((40.0 - V + vt) / 5.0))
    bn = 0.5 * np.exp((10.0 - V + vt) / 40.0)
    return np.stack([am, bm, ah, bh, an, bn], axis=-1), np.stack([x_am, x_bm, x_an], axis=-1)


def grid_v(i):
    i = np.asarray(i, np.float32)
    return V_MIN + (np.float32(V_MAX - V_MIN) * i) / np.float32(N_TAB - 1)


def grid_index(V):
    p = np.floor((np.float32(V) - V_MIN) * AUX).astype(np.int64)
    return np.clip(p, 0, N_TAB - 1)


def hh_rhs(S, p, I_ext):
    """dV/dt, dm/dt, dh/dt, dn/dt of the HH model (Traub rates at the table
    grid point below V), no synaptic conductances."""
    e_exc, e_inh, e_leak, g_leak, c_m, vt, _, _, _, g_na, g_kd, e_na, e_k = [float(x) for x in p]
    V, m, h, n = S
    r, _ = traub_rates(grid_v(grid_index(V)), vt)
    am, bm, ah, bh, an, bn = r
    dV = (I_ext + g_leak * (e_leak - V) + g_na * m**3 * h * (e_na - V) + g_kd * n**4 * (e_k - V)) / c_m
    return np.array([dV, am * (1 - m) - bm * m, ah * (1 - h) - bh * h, an * (1 - n) - bn * n])


def test_hh_rate_table_matches_traub_formulas(oracle):
    vt = float(HH_DEFAULTS[5])
    V = grid_v(np.arange(N_TAB)).astype(np.float64)
    t = oracle.hh_table(vt).astype(np.float64)
    ref, _ = traub_rates(V, vt)
    rel = np.abs(t - ref) / np.abs(ref)
    # float evaluation: the exponent argument carries a rounding error
    # proportional to its size, and exp(a) - 1 cancels for small |a|
    arg = np.stack([(13.0 - V + vt) / 4.0, (V - vt - 40.0) / 5.0, (17.0 - V + vt) / 18.0,
                    (40.0 - V + vt) / 5.0, (15.0 - V + vt) / 5.0, (10.0 - V + vt) / 40.0], 1)
    singular = np.array([1, 1, 0, 0, 1, 0], bool)
    tol = 3e-7 + 2.4e-7 * np.abs(arg) + singular * 2.4e-7 / np.maximum(np.abs(arg), 1e-12)
    assert np.all(rel < tol), np.max(rel / tol)
    # away from the removable singularities the table is accurate to ~1e-6
    far = np.abs(arg) > 0.05
    assert np.max(rel[far]) < 5e-6


def test_hh_rate_check_discriminates(oracle):
    """The check above fails for a table with beta_m and alpha_n swapped or a
    sign error in the h-gate exponent."""
    vt = float(HH_DEFAULTS[5])
    ref, _ = traub_rates(grid_v(np.arange(0, N_TAB, 97)), vt)
    bad = ref.copy()
    bad[:, [1, 4]] = bad[:, [4, 1]]
    assert np.max(np.abs(bad - ref) / np.abs(ref)) > 1e-2
    V = grid_v(np.arange(0, N_TAB, 97)).astype(np.float64)
    bad_ah = 0.128 * np.exp((17.0 + V - vt) / 18.0)
    assert np.max(np.abs(bad_ah - ref[:, 2]) / ref[:, 2]) > 1e-2


def test_hh_initial_gates_are_steady_state(oracle):
    pop = oracle.Population("HHTimeDrivenModel_GPU", "Euler", HH_DEFAULTS, 3, fp64=True, mask=8)
    r, _ = traub_rates(grid_v(grid_index(HH_DEFAULTS[2])), float(HH_DEFAULTS[5]))
    am, bm, ah, bh, an, bn = r
    for var, inf in ((1, am / (am + bm)), (2, ah / (ah + bh)), (3, an / (an + bn))):
        assert np.allclose(pop.state(var), inf, rtol=2e-6, atol=0), var
    assert np.all(pop.state(0) == np.float32(HH_DEFAULTS[2]))


@pytest.mark.parametrize("S0,I", [((-60.0, 0.05, 0.6, 0.32), 20.0), ((-45.0, 0.3, 0.4, 0.5), 0.0),
                                  ((-70.0, 0.01, 0.9, 0.1), 300.0)])
def test_hh_rhs_one_euler_step(oracle, S0, I):
    pop = oracle.Population("HHTimeDrivenModel_GPU", "Euler", HH_DEFAULTS, 1, fp64=True, mask=8)
    for v, x in enumerate(S0):
        pop.set_state(v, [x])
    pop.set_ext_current(np.array([I], np.float32))
    pop.update(1)
    got = np.array([pop.state(v)[0] for v in range(4)])
    inc = DT_MODEL * hh_rhs(np.array(S0), HH_DEFAULTS, I)
    assert np.all(np.abs(got - np.array(S0) - inc) <= 1e-12 + 2e-6 * np.abs(inc)), (got - S0, inc)


BDF_TEXTBOOK = {1: (1.0, [1.0]), 2: (2.0 / 3.0, [4.0 / 3.0, -1.0 / 3.0]),
                3: (6.0 / 11.0, [18.0 / 11.0, -9.0 / 11.0, 2.0 / 11.0]),
                4: (12.0 / 25.0, [48.0 / 25.0, -36.0 / 25.0, 16.0 / 25.0, -3.0 / 25.0]),
                5: (60.0 / 137.0, [300.0 / 137.0, -300.0 / 137.0, 200.0 / 137.0, -75.0 / 137.0, 12.0 / 137.0]),
                6: (60.0 / 147.0, [360.0 / 147.0, -450.0 / 147.0, 400.0 / 147.0, -225.0 / 147.0, 72.0 / 147.0,
                                   -10.0 / 147.0])}


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_bdf_linear_ode_follows_textbook_recursion(oracle, order):
    """dV/dt = (I + g_leak (e_leak - V)) / c_m; step n uses BDF-k with
    k = max(1, min(n - 1, order)) (the variable-order start)."""
    p = LIF_DEFAULTS.copy()
    I = 3.0
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "BDF", p, 2, fp64=True, mask=8, bdf_order=order)
    v0 = np.array([-55.0, -72.0])
    pop.set_state(0, v0)
    pop.set_ext_current(np.array([I, I], np.float32))
    el, c, gl = float(p[2]), float(p[4]), float(p[8])
    lam, b = -gl / c, (I + gl * el) / c
    hist = [v0.copy()]
    for n in range(1, 41):
        pop.update(n)
        k = max(1, min(n - 1, order))
        # the reference stores the coefficients as floats (BDFn.h:621-669)
        c0, cs = BDF_TEXTBOOK[k]
        c0, cs = float(np.float32(c0)), [float(np.float32(x)) for x in cs]
        rhs = c0 * DT_MODEL * b + sum(cj * hist[-1 - j] for j, cj in enumerate(cs))
        hist.append(rhs / (1.0 - c0 * DT_MODEL * lam))
        assert np.allclose(pop.state(0), hist[-1], rtol=0, atol=1e-10), (n, pop.state(0) - hist[-1])


def test_bdf_check_discriminates(oracle):
    """Running the order-2 oracle against the order-1 recursion fails."""
    p = LIF_DEFAULTS.copy()
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "BDF", p, 1, fp64=True, mask=8, bdf_order=2)
    pop.set_state(0, [-55.0])
    pop.set_ext_current(np.array([3.0], np.float32))
    el, c, gl = float(p[2]), float(p[4]), float(p[8])
    lam, b = -gl / c, (3.0 + gl * el) / c
    v = -55.0
    worst = 0.0
    for n in range(1, 21):
        pop.update(n)
        v = (v + DT_MODEL * b) / (1.0 - DT_MODEL * lam)
        worst = max(worst, abs(pop.state(0)[0] - v))
    assert worst > 1e-7


@pytest.mark.parametrize("S0,I", [((-60.0, 0.05, 0.6, 0.32), 20.0), ((-45.0, 0.3, 0.4, 0.5), 0.0)])
def test_hh_bdf1_solves_backward_euler(oracle, S0, I):
    pop = oracle.Population("HHTimeDrivenModel_GPU", "BDF", HH_DEFAULTS, 1, fp64=True, mask=8, bdf_order=1)
    for v, x in enumerate(S0):
        pop.set_state(v, [x])
    pop.set_ext_current(np.array([I], np.float32))
    pop.update(1)
    S1 = np.array([pop.state(v)[0] for v in range(4)])
    resid = S1 - np.array(S0) - DT_MODEL * hh_rhs(S1, HH_DEFAULTS, I)
    assert np.max(np.abs(resid)) < 1e-6, resid
    # and it is not the explicit step
    expl = np.array(S0) + DT_MODEL * hh_rhs(np.array(S0), HH_DEFAULTS, I)
    assert np.max(np.abs(S1 - expl)) > 10 * np.max(np.abs(resid))


def test_hh_fp64_newton_mode_tracks_fp64_oracle(oracle):
    """FP64 Newton with float state between steps stays within float rounding
    of the all-double oracle below the spike upstroke (1e-5 relative; float
    state stalls where the per-step change is below half an ulp).  On the
    upstroke at dt = 0.1 ms the ten Newton iterations (BDFn.h:288-423) do not
    always converge, and the two precisions may then take different branches,
    so traces are compared until the first time V exceeds -45 mV."""
    n = 16
    cur = np.linspace(0.0, 400.0, n).astype(np.float32)
    a = oracle.Population("HHTimeDrivenModel_GPU", "BDF", HH_DEFAULTS, n, fp64=2, mask=8, bdf_order=2)
    b = oracle.Population("HHTimeDrivenModel_GPU", "BDF", HH_DEFAULTS, n, fp64=True, mask=8, bdf_order=2)
    a.set_ext_current(cur)
    b.set_ext_current(cur)
    live = np.ones(n, bool)
    for k in range(1, 601):
        a.update(k)
        b.update(k)
        va, vb = a.state(0), b.state(0)
        live &= (va < -45.0) & (vb < -45.0)
        assert np.all(np.abs(va - vb)[live] <= 1e-5 * np.abs(vb)[live]), k
    assert 0 < live.sum() < n
