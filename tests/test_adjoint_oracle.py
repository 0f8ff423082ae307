"""Pins of the FP64 trace-fitting oracle (oracle/adjoint_oracle.cpp).

The reference has no gradient path, so the FP64 adjoint is pinned by the
mathematics: its gradient of a loss over the V trace equals central finite
differences of its own FP64 forward map (steps small enough to keep the
spike pattern), for every trained quantity (g_leak, e_leak, tau_exc,
tau_inh, synaptic weights), for LIF/Euler and AdEx/RK4, through spikes
(resets) and refractory periods.  A mistake in one term (a dropped
derivative, a wrong stage weight, the reset jump not stopping the adjoint)
makes the comparison fail, as the discrimination test shows.
"""
import numpy as np
import pytest

from edlut_amd import workloads

MODELS = ["LIFTimeDrivenModel_GPU", "AdExTimeDrivenModel_GPU"]


def _problem(oracle, model, n=6, steps=3000, seed=3):
    spec = workloads.adjoint_fit(model, n=n, steps=steps, seed=seed)
    refr = None
    if model == "LIFTimeDrivenModel_GPU":
        from edlut_amd.adjoint import refractory_table

        refr = refractory_table(1e-4, float(spec.params[7]))
    o = oracle.FitOracle(model, spec.params, n, steps, spec.syn_type, spec.events, refr)
    return spec, o


def _loss_grad(o, x, target):
    v, spk = o.forward(*x)
    r = v - target
    return float((r ** 2).mean()), 2.0 * r / r.size, spk


@pytest.mark.parametrize("model", MODELS)
def test_fp64_adjoint_matches_finite_differences(oracle, model):
    spec, o = _problem(oracle, model)
    tgt = spec.target
    target, tspk = o.forward(tgt["g_leak"], tgt["e_leak"], tgt["tau_exc"], tgt["tau_inh"], tgt["weight"])
    st = spec.start
    x = [st["g_leak"].astype(np.float64), st["e_leak"].astype(np.float64), st["tau_exc"].astype(np.float64),
         st["tau_inh"].astype(np.float64), st["weight"].astype(np.float64)]
    L, gv, spk = _loss_grad(o, x, target)
    assert spk.sum() > 5 and tspk.sum() > 5, "the neurons should spike"
    gp, gw = o.backward(*x, gv)
    rng = np.random.default_rng(0)
    checks = [(j, i) for j in range(4) for i in range(spec.n)]
    checks += [(4, (i, s)) for i, s in zip(rng.integers(0, spec.n, 12), rng.integers(0, spec.syn_type.shape[1], 12))]
    n_ok = 0
    for j, idx in checks:
        base = x[j][idx]
        h = 1e-6 * max(abs(base), 1e-3)
        vals = []
        for sgn in (1, -1):
            xx = [a.copy() for a in x]
            xx[j][idx] = base + sgn * h
            l, _, s = _loss_grad(o, xx, target)
            vals.append((l, s))
        if not (np.array_equal(vals[0][1], spk) and np.array_equal(vals[1][1], spk)):
            continue   # the step moved a spike: finite differences do not apply
        fd = (vals[0][0] - vals[1][0]) / (2 * h)
        ad = gp[j][idx] if j < 4 else gw[idx]
        scale = max(abs(fd), abs(ad), 1e-12)
        assert abs(fd - ad) <= 2e-5 * scale + 1e-11, (j, idx, fd, ad)
        n_ok += 1
    assert n_ok >= len(checks) * 3 // 4


def test_fp64_adjoint_check_discriminates(oracle):
    """Dropping the reset's jump (letting the V adjoint pass through a spike)
    changes the e_leak gradient of LIF by far more than the tolerance."""
    model = "LIFTimeDrivenModel_GPU"
    spec, o = _problem(oracle, model)
    st = spec.start
    x = [st[k].astype(np.float64) for k in ("g_leak", "e_leak", "tau_exc", "tau_inh", "weight")]
    v, spk = o.forward(*x)
    gv = np.random.default_rng(1).normal(size=v.shape)
    gp, _ = o.backward(*x, gv)
    # the part of dL/de_leak routed through the resets: sum of the V adjoint at
    # each spike, which the reset sends to e_leak; it is a large share
    i = int(np.argmax(spk.sum(0)))
    assert spk[:, i].sum() > 3
    h = 1e-6
    xx = [a.copy() for a in x]
    xx[1][i] += h
    vp, sp = o.forward(*xx)
    xx[1][i] -= 2 * h
    vm, sm = o.forward(*xx)
    assert np.array_equal(sp, spk) and np.array_equal(sm, spk)
    fd = ((vp - vm) * gv).sum() / (2 * h)
    assert abs(fd - gp[1][i]) <= 1e-5 * abs(fd)
    # the same derivative with V continuing through the resets differs
    contrib = (gv[:, i] * (spk[:, i] == 1)).sum()
    assert abs(contrib) > 1e-3 * abs(fd)


@pytest.mark.parametrize("model,bound", [("LIFTimeDrivenModel_GPU", 1e-5), ("AdExTimeDrivenModel_GPU", None)])
def test_gradient_conditioning_float_vs_fp64_trajectory(oracle, model, bound):
    """How much float rounding of the forward trajectory alone moves the
    gradient: the FP64 adjoint along a float run vs along the FP64 run (same
    spike trains).  LIF: ~1e-6.  AdEx: tens of percent, from the steps on the
    exponential upswing before each spike; this is why the GPU adjoint is
    checked against the FP64 adjoint along its own trajectory."""
    n, steps = 32, 10000
    spec = workloads.adjoint_fit(model, n=n, steps=steps, seed=7)
    refr = None
    if model == "LIFTimeDrivenModel_GPU":
        from edlut_amd.adjoint import refractory_table

        refr = refractory_table(1e-4, float(spec.params[7]))
    o = oracle.FitOracle(model, spec.params, n, steps, spec.syn_type, spec.events, refr)
    x = [spec.start[k] for k in ("g_leak", "e_leak", "tau_exc", "tau_inh", "weight")]
    tgt, _ = o.forward(*[spec.target[k] for k in ("g_leak", "e_leak", "tau_exc", "tau_inh", "weight")], fp64=False)
    v64, s64 = o.forward(*x, fp64=True)
    gp, _ = o.backward(*x, 2 * (v64 - tgt) / v64.size)
    traj, sf, since = o.trajectory(*x)
    vf, _ = o.forward(*x, fp64=False)
    gp2, _ = o.backward_along(*x, 2 * (vf - tgt) / vf.size, traj, sf, since)
    same = np.all(sf == s64, axis=0)
    assert same.mean() > 0.9 and s64.sum() > 20
    err = np.linalg.norm(gp2[0][same] - gp[0][same]) / np.linalg.norm(gp[0][same])
    if bound is not None:
        assert err < bound, err
    else:
        assert err > 1e-3, err   # ill-conditioned: far above float rounding
