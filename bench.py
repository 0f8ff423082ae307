"""Benchmark: neuron-updates/s of the time-driven AdEx RK4 update on one MI355X.

Workload (one GPU): 1,000,000-neuron AdEx recurrent network, fp32 SoA state
(V, w, g_exc, g_inh), RK4 dt = 0.1 ms, fixed in-degree 200 CSR (80/20
exc/inh), lognormal(-1, 0.5) nS weights, delays 1-10 steps, external drive of
100 Poisson synapses at 5 Hz per neuron.  One "step" is one full network
update on the device: ordered delivery of the spikes due, the fused update
kernel (integrate, decay, threshold/reset, compaction) and spike logging.
Inputs are resident in HBM before the timed region.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_UPDATE = 32      # V, w, g_exc, g_inh: 16 B read + 16 B written


def read_traffic(path: str):
    """HBM bytes per update-kernel launch from the committed rocprof summary."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def calibration_copy(n_bytes: int) -> None:
    """One device-to-device copy of n_bytes (HIP runtime blit kernel), so that a
    counter profile of this run holds a kernel of known traffic to scale the
    FETCH_SIZE / WRITE_SIZE readings of the update kernel by."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n_bytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n_bytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n_bytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n_bytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(spec, steps: int):
    """The oracle (single-thread C++ restatement) on a bounded sample of the
    same workload: `steps` updates of the full network from rest."""
    from edlut_amd.models import AdExTimeDrivenModel_GPU
    from edlut_amd.network import build_out_csr
    from oracle import pyoracle

    params = np.array([v for _, v in AdExTimeDrivenModel_GPU.PARAMS], np.float32)
    pop, net = pyoracle.from_spec(spec, params, build_out_csr)
    t0 = time.perf_counter()
    net.run(1, steps)
    dt = time.perf_counter() - t0
    return spec.n * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--w-ext", type=float, default=0.65)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device copy of this size (for counter profiles)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import edlut_amd as E
    from edlut_amd import workloads

    if args.calibrate_copy_mb:
        calibration_copy(args.calibrate_copy_mb << 20)

    t_setup = time.perf_counter()
    spec = workloads.adex_recurrent(n=args.neurons, steps=args.warmup + args.steps, w_ext=args.w_ext)
    model, net = E.from_spec(spec)
    print(f"setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    net.run(args.warmup)                       # untimed warmup (synchronised)
    net.set_timing(args.steps)
    net.synchronize()
    t0 = time.perf_counter()
    net.run(args.steps, sync=False)
    net.synchronize()
    wall = time.perf_counter() - t0
    ms_launch, launches = net.update_timing()
    st, ix = net.spikes()
    timed = st > args.warmup
    rate_hz = timed.sum() / (spec.n * args.steps * 1e-4)

    updates = spec.n * args.steps
    value = updates / wall
    avg_launch_s = ms_launch / 1e3 / max(launches, 1)
    achieved = BYTES_PER_UPDATE * spec.n / avg_launch_s
    traffic = read_traffic(os.path.join(ROOT, "profiles", "adex_rk4_update_1m_pmc.txt"))

    out = {
        "metric": "neuron-updates/sec (fp32, RK4) at 1 GPU",
        "value": value,
        "unit": "neuron-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded network and Poisson drive)",
        "config": {
            "workload": "AdEx recurrent network, RK4 dt 0.1 ms, device-resident delivery",
            "neurons": spec.n,
            "in_degree": 200,
            "max_delay_steps": spec.max_delay,
            "external_drive": f"100 synapses x 5 Hz, w {args.w_ext} nS",
            "mean_rate_hz": round(float(rate_hz), 3),
            "update_kernel_us": round(avg_launch_s * 1e6, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
        },
    }
    if not args.no_cpu:
        cpu_spec = workloads.adex_recurrent(n=args.neurons, steps=args.cpu_steps, w_ext=args.w_ext)
        v, dt = cpu_baseline(cpu_spec, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": v,
            "unit": "neuron-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle C++ restatement, first {args.cpu_steps} updates of the same "
                      f"{spec.n}-neuron network ({dt:.1f} s)",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
