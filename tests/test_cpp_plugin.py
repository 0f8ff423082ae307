"""The C++ host plug-ins (edlut_amd/host/edlut_plugin.hpp) over the C ABI.

CPU: parameter handling, errors and identifiers without a device.  GPU: the
plug-in driven like EDLUT's simulation loop (input spikes of each step in
delivery order, UpdateState, spike indices) reproduces the oracle bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import ADEX_DEFAULTS, HH_DEFAULTS, LIF_DEFAULTS, bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "edlut_amd", "host", "plugin_driver")


def _driver():
    if not os.path.exists(DRIVER):
        from edlut_amd import build

        build.build_driver()
    return DRIVER


def test_cpp_plugin_parameter_handling():
    r = subprocess.run([_driver(), "selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def _write_inputs(path, n, steps, ext, rng, wmax):
    events = []
    with open(path, "wb") as f:
        np.array([n, steps], np.int32).tofile(f)
        np.asarray(ext, np.float32).tofile(f)
        for _ in range(steps):
            c = int(rng.integers(0, 3 * n // 4))
            tg = rng.integers(0, n, c).astype(np.int32)
            ty = rng.integers(0, 3, c).astype(np.int32)
            w = rng.uniform(0, wmax, c).astype(np.float32)
            np.array([c], np.int32).tofile(f)
            tg.tofile(f), ty.tofile(f), w.tofile(f)
            events.append((tg, ty, w))
    return events


def _read_outputs(path, n):
    with open(path, "rb") as f:
        ns = int(np.fromfile(f, np.int64, 1)[0])
        sp = np.fromfile(f, np.int32, 2 * ns).reshape(ns, 2)
        nv = int(np.fromfile(f, np.int32, 1)[0])
        st = np.fromfile(f, np.float32, nv * n).reshape(nv, n)
    return sp[:, 0], sp[:, 1], st


@pytest.mark.gpu
@pytest.mark.parametrize("model,integ,order,params,wmax,cur", [
    ("AdExTimeDrivenModel_GPU", "RK4", 1, ADEX_DEFAULTS, 6.0, 150.0),
    ("LIFTimeDrivenModel_GPU", "Euler", 1, LIF_DEFAULTS, 0.3, 1.0),
    ("HHTimeDrivenModel_GPU", "BDF", 2, HH_DEFAULTS, 6.0, 100.0),
    ("LIFTimeDrivenModel_GPU", "Bifixed_RK2", 3, LIF_DEFAULTS, 0.3, 1.0),
])
def test_cpp_plugin_matches_oracle(oracle, tmp_path, model, integ, order, params, wmax, cur):
    n, steps = 700, 400
    rng = np.random.default_rng(17)
    ext = rng.uniform(0, cur, n).astype(np.float32)
    events = _write_inputs(tmp_path / "in.bin", n, steps, ext, rng, wmax)
    print(f"case cpp plugin {model} {integ}", flush=True)
    r = subprocess.run([_driver(), "run", model, integ, str(order), str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    st, ix, state = _read_outputs(tmp_path / "out.bin", n)
    pop = oracle.Population(model, integ, params, n, mask=15, bdf_order=order, n_steps=order)
    pop.set_ext_current(ext)
    ost, oix = [], []
    for k, (tg, ty, w) in enumerate(events, start=1):
        pop.add_inputs(tg, ty.astype(np.uint8), w)
        s = pop.update(k)
        ost += [k] * len(s)
        oix += list(s)
    assert len(oix) > 20, "the case should spike"
    assert np.array_equal(st, ost) and np.array_equal(ix, oix)
    for v in range(pop.nv):
        assert np.array_equal(bits(state[v]), bits(pop.state(v))), f"variable {v}"
