"""Benchmark: neuron-updates/s of the time-driven AdEx RK4 update on one MI355X.

Workload (one GPU): 1,000,000-neuron AdEx recurrent network, fp32 SoA state
(V, w, g_exc, g_inh), RK4 dt = 0.1 ms, fixed in-degree 200 CSR (80/20
exc/inh), lognormal(-1, 0.5) nS weights, delays 1-10 steps, external drive of
100 Poisson synapses at 5 Hz per neuron.  One "step" is one full network
update on the device: ordered delivery of the spikes due, the fused update
kernel (integrate, decay, threshold/reset, compaction) and spike logging.
Inputs are resident in HBM before the timed region.

A second leg measures the batched parameter sweep: 4,096 independent AdEx
networks (lognormal sizes, median 5,000, clipped to [1,000, 20,000], in-degree
100, per-network g_leak, E_leak, a, b, tau_w within +-30%), RK4, 0.5 s,
advanced together as one grouped population with the autotuner's launch
shape and hipGraph-replayed steps, next to the same networks run one after
another as plain networks (each with its own captured steps), on a
size-stratified sample.

Further legs: the 1,000-neuron LIF network with 125,000 Poisson inputs (the
reference's CPU-sized case), the cerebellar circuit, and the trace-fitting
adjoint.  Each leg carries a `roofline` object for its dominant kernel and,
unless --no-cpu, a `cpu_baseline` from the oracle on one thread and on every
thread of the host share.

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
BYTES_PER_EVENT = 24       # delivery: 8 B synapse read, 4 B fill atomic, 12 B (order, w) entry
# VALU issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane-instructions per second
VALU_LANE_OPS_PEAK = 256 * 4 * 16 * 2.4e9
CLOCK_HZ = 2.4e9


def cpu_threads() -> int:
    """Threads of the multi-core CPU baseline: the host share set for this
    process (OMP_NUM_THREADS where it is set, else every CPU)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def read_profile_value(name: str, key: str):
    """A `key: value` line of a committed rocprof summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
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
    """The oracle (C++ restatement) on a bounded sample of the same workload:
    `steps` updates of the full network from rest on one thread, then the
    next `steps` updates on every thread of the host share (population
    updates and Poisson draws in parallel, delivery in order)."""
    from edlut_amd.models import AdExTimeDrivenModel_GPU
    from edlut_amd.network import build_out_csr
    from oracle import pyoracle

    params = np.array([v for _, v in AdExTimeDrivenModel_GPU.PARAMS], np.float32)
    pop, net = pyoracle.from_spec(spec, params, build_out_csr)
    out = []
    for i, threads in enumerate((1, cpu_threads())):
        pyoracle.set_threads(threads)
        t0 = time.perf_counter()
        net.run(1 + i * steps, steps)
        dt = time.perf_counter() - t0
        out.append((threads, spec.n * steps / dt, dt))
    pyoracle.set_threads(1)
    return out


def cpu_entry(samples, what: str):
    """cpu_baseline object: the all-thread figure, the one-thread one beside it."""
    (t1, v1, d1), (tn, vn, dn) = samples
    return {
        "value": vn,
        "unit": "neuron-updates/s",
        "cores": tn,
        "kind": "port",
        "nproc": os.cpu_count(),
        "sample": f"{what}: {d1:.1f} s on 1 thread, {dn:.1f} s on {tn} threads",
        "single_thread": {"value": v1, "cores": t1},
    }


def bench_sweep(n_networks: int, steps: int, seq_sample: int, seq_steps: int, cpu: bool):
    """Batched sweep vs sequential plain networks (see the module docstring)."""
    import edlut_amd as E
    from edlut_amd import workloads
    from edlut_amd.network import Network

    spec = workloads.adex_sweep(n_networks=n_networks, steps=steps)
    thr, w_ext, ext_seed = spec.ext
    t0 = time.perf_counter()
    m = E.CreateNeuronModel("AdExTimeDrivenModel_GPU")
    m.SetParameters({"int_meth": E.integration_method("RK4", 1e-4)})
    m.InitializeGroups(spec.sizes, spec.params)
    net = Network(m, spec.max_delay)
    net.set_log_capacity(0)                      # spike count only
    net.generate_recurrent(spec.in_degree, spec.exc_fraction, spec.w_mu, spec.w_sigma, spec.wiring_seed)
    net.set_external(thr, w_ext, ext_seed)
    net.synchronize()
    setup = time.perf_counter() - t0
    # launch shape chosen in the ignited regime: 500 settling steps, then
    # each shape continues the run for its trial (reset afterwards)
    items, lanes, blocks, tune_ms = net.autotune_settled(500, 32, blocks=(256, 1024, 4096))
    print(f"sweep setup {setup:.1f}s, autotune -> items {items} lanes {lanes} blocks {blocks} "
          f"({tune_ms:.4f} ms/step)", file=sys.stderr, flush=True)
    net.synchronize()
    t0 = time.perf_counter()
    chunk = 1000
    for k in range(0, steps, chunk):     # a progress line per chunk (one sync each)
        net.run(min(chunk, steps - k), sync=False)
        net.synchronize()
        print(f"sweep: {k + min(chunk, steps - k)} of {steps} steps, {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    n_spk = net.spike_count()
    real = int(spec.sizes.sum())
    batched = real * steps / wall
    rate_hz = n_spk / (real * steps * 1e-4)
    net.close()
    m.close()

    # sequential plain networks: one network at a time, each with its own
    # captured steps (the plain path's default), on a sample spread evenly
    # over the size distribution
    order = np.argsort(spec.sizes, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, seq_sample).round().astype(int)]
    seq_updates, seq_time = 0, 0.0
    cpu_nets = []
    for g in pick:
        pm = E.CreateNeuronModel("AdExTimeDrivenModel_GPU")
        names = [k for k, _ in type(pm).PARAMS]
        pm.SetParameters({k: float(v) for k, v in zip(names, spec.params[g])})
        pm.SetParameters({"int_meth": E.integration_method("RK4", 1e-4)})
        pm.InitializeStates(int(spec.sizes[g]))
        pn = Network(pm, spec.max_delay)
        pn.set_log_capacity(0)
        pn.generate_recurrent(spec.in_degree, spec.exc_fraction, spec.w_mu, spec.w_sigma,
                              spec.wiring_seed + int(g))
        pn.set_external(thr, w_ext, ext_seed + int(g))
        pn.run(16)
        t0 = time.perf_counter()
        pn.run(seq_steps, sync=False)
        pn.synchronize()
        seq_time += time.perf_counter() - t0
        seq_updates += int(spec.sizes[g]) * seq_steps
        print(f"sequential: network {g} ({spec.sizes[g]} neurons) {seq_time:.1f}s", file=sys.stderr, flush=True)
        if cpu and len(cpu_nets) < 2 and spec.sizes[g] >= np.median(spec.sizes):
            cpu_nets.append((int(g), pn.synapses(0)))
        pn.close()
        pm.close()
    sequential = seq_updates / seq_time
    # delivery traffic: each spike walks its synapses (mean out-degree =
    # in-degree) at BYTES_PER_EVENT each, besides the state of every neuron
    events_per_step = n_spk / steps * spec.in_degree
    bytes_per_step = BYTES_PER_UPDATE * real + BYTES_PER_EVENT * events_per_step
    out = {
        "workload": f"{n_networks} AdEx networks, RK4 dt 0.1 ms, {steps} steps, in-degree {spec.in_degree}",
        "neurons": real,
        "value": batched,
        "unit": "neuron-updates/s",
        "ms_per_step": wall * 1e3 / steps,
        "mean_rate_hz": round(float(rate_hz), 3),
        "events_per_step": round(float(events_per_step)),
        "launch": {"items_per_thread": items, "fold_lanes": lanes, "fold_blocks": blocks,
                   "autotune_ms_per_step": tune_ms, "autotune": "500 settling steps, 32-step trials"},
        "roofline": {
            "bound": "hbm",
            "achieved": bytes_per_step / (wall / steps) / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": bytes_per_step / (wall / steps) / HBM_PEAK,
            "traffic": None,
            "basis": f"whole step: {BYTES_PER_UPDATE} B per neuron + {BYTES_PER_EVENT} B per delivered event",
        },
        "sequential_plain": {
            "value": sequential,
            "unit": "neuron-updates/s",
            "sample": f"{seq_sample} networks spread over the size distribution, {seq_steps} steps each, "
                      f"captured-step replay",
        },
        "speedup_vs_sequential": batched / sequential,
    }
    if cpu and cpu_nets:
        from oracle import pyoracle

        c_steps = 1000
        samples = []
        for threads in (1, cpu_threads()):
            pyoracle.set_threads(threads)
            upd, secs = 0, 0.0
            for g, (rp, tt, ww) in cpu_nets:
                pop = pyoracle.Population("AdExTimeDrivenModel_GPU", "RK4", spec.params[g], int(spec.sizes[g]),
                                          mask=3)
                on = pyoracle.Net(pop, spec.max_delay)
                on.add_self(rp, tt, ww)
                on.set_external(thr, w_ext, ext_seed + g)
                t0 = time.perf_counter()
                on.run(1, c_steps)
                secs += time.perf_counter() - t0
                upd += int(spec.sizes[g]) * c_steps
            samples.append((threads, upd / secs, secs))
        pyoracle.set_threads(1)
        out["cpu_baseline"] = cpu_entry(samples, f"oracle C++ restatement, networks "
                                                 f"{[g for g, _ in cpu_nets]} of the sweep, first {c_steps} updates")
    return out


def bench_adjoint(n: int, steps: int, checkpoint: int, cpu: bool, reps: int = 3):
    """Trace-fitting forward and adjoint kernels (256 LIF/Euler + 256 AdEx/RK4
    neurons, 50 Poisson inputs at 20 Hz, 10,000 steps, checkpoints every 100
    steps): kernel time of the backward pass over the forward pass.

    Both kernels run one lane per neuron through a sequential recurrence, so
    they are bound by the latency of each lane's dependent instruction chain:
    `roofline` compares the measured forward time with the issue time of its
    VALU instructions on one SIMD (4 cycles per wave64 instruction), from the
    instruction count of the committed counter profile."""
    import torch

    from edlut_amd import workloads
    from edlut_amd.adjoint import FitProblem

    out = {"workload": f"{n} LIF (Euler) + {n} AdEx (RK4) neurons, 50 inputs at 20 Hz, {steps} steps, "
                       f"checkpoint every {checkpoint}"}
    total_f = total_b = 0.0
    specs = {}
    for model, key in (("LIFTimeDrivenModel_GPU", "lif"), ("AdExTimeDrivenModel_GPU", "adex")):
        spec = workloads.adjoint_fit(model, n=n, steps=steps, seed=5)
        specs[key] = spec
        prob = FitProblem(model, spec.params, n, steps, spec.syn_type, spec.events, checkpoint=checkpoint)
        x = [torch.tensor(spec.start[k], dtype=torch.float32, device="cuda")
             for k in ("g_leak", "e_leak", "tau_exc", "tau_inh", "weight")]
        gv = torch.randn(steps, n, device="cuda")
        prob.forward_raw(*x)
        prob.backward_raw(*x, gv)
        fs, bs = [], []
        for _ in range(reps):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record()
            prob.forward_raw(*x)
            e[1].record()
            prob.backward_raw(*x, gv)
            e[2].record()
            torch.cuda.synchronize()
            fs.append(e[0].elapsed_time(e[1]))
            bs.append(e[1].elapsed_time(e[2]))
        f, b = min(fs), min(bs)
        out[key] = {"forward_ms": f, "backward_ms": b, "ratio": b / f}
        instr = read_profile_value("adjoint_fit_kernels.txt", f"{key}_forward_valu_per_wave")
        if instr:
            issue_ms = instr * 4 / CLOCK_HZ * 1e3
            out[key]["roofline"] = {"bound": "latency", "achieved": issue_ms, "peak": f, "unit": "ms",
                                    "frac": issue_ms / f,
                                    "basis": "VALU issue time of one wave's forward instructions / forward time"}
        total_f += f
        total_b += b
        prob.close()
        print(f"adjoint {key}: forward {f:.3f} ms, backward {b:.3f} ms", file=sys.stderr, flush=True)
    out["cost_ratio"] = total_b / total_f
    if cpu:
        from oracle import pyoracle

        spec = specs["lif"]
        fo = pyoracle.FitOracle("LIFTimeDrivenModel_GPU", spec.params, n, steps, spec.syn_type, spec.events)
        x = [spec.start[k] for k in ("g_leak", "e_leak", "tau_exc", "tau_inh", "weight")]
        gv = np.random.default_rng(0).standard_normal((steps, n))
        t0 = time.perf_counter()
        fo.forward(*x)
        fo.backward(*x, gv)
        secs = time.perf_counter() - t0
        gpu_ms = out["lif"]["forward_ms"] + out["lif"]["backward_ms"]
        out["cpu_baseline"] = {
            "value": n * steps / secs,
            "unit": "neuron-steps/s (forward + adjoint)",
            "cores": 1,
            "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"FP64 oracle forward + adjoint of the LIF problem, {secs:.1f} s "
                      f"(GPU: {gpu_ms:.1f} ms, {n * steps / (gpu_ms / 1e3):.3g} neuron-steps/s)",
        }
    return out


def bench_cerebellum(steps: int, warmup: int, cpu_steps: int):
    """Layered cerebellar circuit at full size (200,000 granule cells under
    BDF2, 2,000 Golgi LIF, 100 HH Purkinje cells under BDF2 with the FP64
    Newton step, 10,000 mossy fibres at U[20,80] Hz, STDP on the 200,000
    granule -> Purkinje synapses): lockstep circuit steps, hipGraph replays.

    The granule kernel is bound by VALU issue: `roofline` relates its VALU
    instructions per launch (committed counter profile) to the measured step
    time, against 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz."""
    from edlut_amd import workloads
    from edlut_amd.network import cerebellum_from_spec

    spec = workloads.cerebellum(steps=warmup + steps)
    t_setup = time.perf_counter()
    models, nets, circuit = cerebellum_from_spec(spec)
    print(f"cerebellum setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    circuit.run(warmup)
    t0 = time.perf_counter()
    done = 0
    while done < steps:   # progress every 2000 steps
        k = min(2000, steps - done)
        circuit.run(k)
        done += k
        print(f"cerebellum {done}/{steps} steps", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    sizes = (spec.n_grc, spec.n_goc, spec.n_pc)
    rates = {}
    for name, n, size in zip(("granule", "golgi", "purkinje"), nets, sizes):
        st, _ = n.spikes()
        rates[name] = round(float((st > warmup).sum() / (size * steps * 1e-4)), 3)
    updates = sum(sizes) * steps
    step_s = wall / steps
    out = {
        "workload": f"{spec.n_grc} granule (BDF2) + {spec.n_goc} Golgi (LIF, Euler) + {spec.n_pc} Purkinje "
                    f"(HH, BDF2, FP64 Newton); {spec.n_mf} mossy fibres U[20,80] Hz; in-degree 4 / 100 / 2000; "
                    f"STDP granule->Purkinje; dt 0.1 ms",
        "value": updates / wall,
        "unit": "neuron-updates/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": step_s * 1e3,
        "rates_hz": rates,
    }
    valu = read_profile_value("granule_bdf2_cerebellum_200k.txt", "valu_instr_per_launch")
    if valu:
        ach = valu * 64 / step_s
        out["roofline"] = {"bound": "valu", "achieved": ach / 1e12, "peak": VALU_LANE_OPS_PEAK / 1e12,
                           "unit": "T lane-instr/s", "frac": ach / VALU_LANE_OPS_PEAK,
                           "basis": "granule BDF2 kernel VALU instructions x 64 lanes per whole circuit step"}
    circuit.close()
    for n in reversed(nets):
        n.close()
    for m in models:
        m.close()
    if cpu_steps > 0:
        import edlut_amd as E
        from oracle import pyoracle

        def defaults(name):
            return np.array([v for _, v in E.REGISTRY[name].PARAMS], np.float32)

        np_out_csr = E.build_out_csr
        D = spec.max_delay
        t_c = time.perf_counter()
        gr = pyoracle.Population("EgidioGranuleCell_TimeDriven_GPU", "BDF", [], spec.n_grc, mask=1, bdf_order=2)
        go = pyoracle.Population("LIFTimeDrivenModel_GPU", "Euler", defaults("LIFTimeDrivenModel_GPU"),
                                 spec.n_goc, mask=1)
        pc = pyoracle.Population(spec.pc_model, "BDF", defaults(spec.pc_model), spec.n_pc, mask=1, fp64=2,
                                 bdf_order=2)
        ng, no, npc = pyoracle.Net(gr, D), pyoracle.Net(go, D), pyoracle.Net(pc, D)
        ng.add_poisson(spec.mf_states, 0.0, *np_out_csr(spec.n_mf, spec.n_grc, D, *spec.mf_grc))
        ng.set_poisson_rates(0, spec.mf_rates)
        no.add_link(ng, *np_out_csr(spec.n_grc, spec.n_goc, D, *spec.grc_goc))
        npc.add_link(ng, *np_out_csr(spec.n_grc, spec.n_pc, D, *spec.grc_pc))
        npc.set_stdp(0, *spec.stdp)
        samples = []
        k = 1
        for threads in (1, cpu_threads()):
            pyoracle.set_threads(threads)
            t0 = time.perf_counter()
            pyoracle.circuit_run([ng, no, npc], k, cpu_steps)
            dt = time.perf_counter() - t0
            k += cpu_steps
            samples.append((threads, sum(sizes) * cpu_steps / dt, dt))
        pyoracle.set_threads(1)
        out["cpu_baseline"] = cpu_entry(samples, f"oracle C++ restatement, updates 1-{cpu_steps} (1 thread) and "
                                                 f"{cpu_steps + 1}-{2 * cpu_steps} of the same circuit (setup "
                                                 f"{time.perf_counter() - t_c - sum(x[2] for x in samples):.1f} s)")
    return out


def bench_lif_poisson(cpu: bool):
    """The reference's CPU-sized case: 1,000 LIF neurons (Euler, dt 0.1 ms),
    100 AMPA + 25 GABA Poisson inputs each at 10 Hz (125,000 generators),
    10,000 updates (1 s).  The GPU network runs the whole second; the oracle
    runs the same second on one thread and on every thread, and its raster
    is compared with the GPU's."""
    import edlut_amd as E
    from edlut_amd import workloads

    steps = 10000
    spec = workloads.lif_poisson(steps=steps)
    model, net = E.from_spec(spec)
    net.run(16)
    net.reset()
    net.synchronize()
    t0 = time.perf_counter()
    net.run(steps, sync=False)
    net.synchronize()
    wall = time.perf_counter() - t0
    st, ix = net.spikes()
    net.close()
    model.close()
    out = {
        "workload": "1000 LIF (Euler), 125,000 Poisson inputs at 10 Hz, delays 1-5 ms, 10000 updates",
        "value": spec.n * steps / wall,
        "unit": "neuron-updates/s",
        "ms_per_step": wall * 1e3 / steps,
        "spikes": int(len(st)),
    }
    if cpu:
        from edlut_amd.network import build_out_csr
        from oracle import pyoracle

        params = np.array([v for _, v in E.REGISTRY[spec.model].PARAMS], np.float32)
        samples = []
        for threads in (1, cpu_threads()):
            pyoracle.set_threads(threads)
            pop, onet = pyoracle.from_spec(spec, params, build_out_csr)
            t0 = time.perf_counter()
            ost, oix = onet.run(1, steps)
            dt = time.perf_counter() - t0
            samples.append((threads, spec.n * steps / dt, dt))
        pyoracle.set_threads(1)
        out["raster_equal_to_oracle"] = bool(np.array_equal(st, ost) and np.array_equal(ix, oix))
        out["cpu_baseline"] = cpu_entry(samples, "oracle C++ restatement, the whole 10000-update run")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--settle", type=int, default=2000,
                    help="untimed steps after the launch-shape search, before the warmup, so that the "
                         "timed steps see the network's settled activity")
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--kernel-steps", type=int, default=2000,
                    help="timed steps whose update kernels stamp their workgroups on the device clock")
    ap.add_argument("--w-ext", type=float, default=0.65)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device copy of this size (for counter profiles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default="",
                    help="items,fold_lanes,fold_blocks of the main leg instead of the autotuner's choice")
    ap.add_argument("--sweep-networks", type=int, default=4096)
    ap.add_argument("--sweep-steps", type=int, default=5000)
    ap.add_argument("--seq-sample", type=int, default=16)
    ap.add_argument("--seq-steps", type=int, default=0,
                    help="steps per sequential network (default: the sweep's steps)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-cerebellum", action="store_true")
    ap.add_argument("--no-lif", action="store_true")
    ap.add_argument("--only", choices=["cerebellum", "sweep", "adjoint", "lif"],
                    help="run one secondary leg alone (for profiling its kernels)")
    ap.add_argument("--cerebellum-steps", type=int, default=20000)
    ap.add_argument("--cerebellum-cpu-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import edlut_amd as E
    from edlut_amd import workloads

    if args.calibrate_copy_mb:
        calibration_copy(args.calibrate_copy_mb << 20)
    cpu = not args.no_cpu
    if args.only == "cerebellum":
        print(json.dumps({"cerebellum": bench_cerebellum(args.cerebellum_steps, 100, 0)}))
        return
    if args.only == "sweep":
        print(json.dumps({"batched": bench_sweep(args.sweep_networks, args.sweep_steps, args.seq_sample,
                                                 args.seq_steps or args.sweep_steps, False)}))
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": bench_adjoint(256, 10000, 100, False)}))
        return
    if args.only == "lif":
        print(json.dumps({"lif_poisson": bench_lif_poisson(False)}))
        return

    t_setup = time.perf_counter()
    spec = workloads.adex_recurrent(n=args.neurons, steps=args.settle + args.warmup + args.steps + args.kernel_steps,
                                    w_ext=args.w_ext)
    model, net = E.from_spec(spec)
    print(f"setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    # launch shape: timed in the settled regime the timed region runs in
    # (the network is reset after), or given by --launch
    if args.launch:
        items, lanes, blocks = (int(x) for x in args.launch.split(","))
        net.set_launch(items, lanes, blocks)
        tune_ms = None
    else:
        items, lanes, blocks, tune_ms = net.autotune_settled(args.settle, 256)
    print(f"launch -> items {items} lanes {lanes} blocks {blocks} (autotune {tune_ms} ms/step)",
          file=sys.stderr, flush=True)
    net.run(args.settle)                       # settle into the asynchronous state
    net.run(args.warmup)                       # untimed warmup (synchronised)
    print("timed region", file=sys.stderr, flush=True)
    first = net.next_step
    net.set_timing(min(args.steps, args.kernel_steps))
    net.synchronize()
    t0 = time.perf_counter()
    net.run(args.steps, sync=False)            # hipGraph replays
    net.synchronize()
    wall = time.perf_counter() - t0
    # update-kernel duration in the timed region (the first kernel_steps
    # steps stamp the device wall clock): the launch duration as a kernel
    # trace sees it (from
    # its first workgroup to the next launch's first workgroup: dispatch and
    # end-of-kernel write-back included), and its workgroups' active span
    ms_launch, launches = net.update_dispatch_timing()
    ms_active, launches_active = net.update_timing()
    net.set_timing(0)
    st, ix = net.spikes()
    timed = (st >= first) & (st < first + args.steps)
    rate_hz = timed.sum() / (spec.n * args.steps * 1e-4)
    settle_rate = ((st >= first - args.warmup - args.settle) & (st < first)).sum() / (
        spec.n * (args.settle + args.warmup) * 1e-4)

    updates = spec.n * args.steps
    value = updates / wall
    avg_launch_s = ms_launch / 1e3 / max(launches, 1)
    achieved = BYTES_PER_UPDATE * spec.n / avg_launch_s
    traffic = read_profile_value("adex_rk4_update_1m_pmc.txt", f"traffic_bytes_per_launch_items{items}")

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
            "settle_steps": args.settle,
            "mean_rate_hz": round(float(rate_hz), 3),
            "settle_rate_hz": round(float(settle_rate), 3),
            "update_kernel_us": round(avg_launch_s * 1e6, 3),
            "update_kernel_timed_launches": launches,
            "update_kernel_timing": "device wall clock, first workgroup start to the next launch's, timed region",
            "update_kernel_active_us": round(ms_active / 1e3 / max(launches_active, 1) * 1e6, 3),
            "launch": {"items_per_thread": items, "fold_lanes": lanes, "fold_blocks": blocks,
                       "autotune": "24 shapes, 256 steps each, after settling" if tune_ms else "given"},
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
    net.close()
    model.close()
    if not args.no_lif:
        out["lif_poisson"] = bench_lif_poisson(cpu)
        print("lif_poisson done", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(256, 10000, 100, cpu)
    if not args.no_cerebellum:
        out["cerebellum"] = bench_cerebellum(args.cerebellum_steps, 100, args.cerebellum_cpu_steps if cpu else 0)
    if not args.no_sweep:
        out["batched"] = bench_sweep(args.sweep_networks, args.sweep_steps, args.seq_sample,
                                     args.seq_steps or args.sweep_steps, cpu)
    if cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        cpu_spec = workloads.adex_recurrent(n=args.neurons, steps=2 * args.cpu_steps, w_ext=args.w_ext)
        out["cpu_baseline"] = cpu_entry(cpu_baseline(cpu_spec, args.cpu_steps),
                                        f"oracle C++ restatement, updates 1-{args.cpu_steps} (1 thread) and "
                                        f"{args.cpu_steps + 1}-{2 * args.cpu_steps} (all threads) of the same "
                                        f"{spec.n}-neuron network from rest")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
