"""Seeded synthetic workloads of the benchmark configurations.

Each builder returns plain numpy arrays (parameters, edge lists, RNG states) so
the same inputs can be fed to the HIP path and to any CPU implementation.

  lif_poisson      1,000 LIF neurons (Euler, dt 0.1 ms), each with 100 AMPA +
                   25 GABA Poisson inputs at 10 Hz; weights U[0,0.5] / U[0,1] nS;
                   delays U{1..5 ms} on the 0.1 ms grid.
  adex_recurrent   AdEx network, RK4 dt 0.1 ms, fixed in-degree CSR (sources
                   without replacement, no self-connections), 80/20 exc/inh
                   sources, weights lognormal(-1, 0.5) nS, delays U{1..10}
                   steps, external drive = 100 Poisson synapses at 5 Hz each.
  adex_sweep       parameter sweep: many independent AdEx networks (RK4,
                   dt 0.1 ms) of lognormal sizes, each with its own g_leak,
                   E_leak, a, b, tau_w drawn uniformly within +-30% of the
                   defaults; in-degree 100 wiring and the external drive are
                   generated per network from its own seed.
  cerebellum       layered cerebellar network: mossy fibres (Poisson, one
                   rate per fibre, U[20,80] Hz) -> granule cells (Egidio model,
                   BDF2; in-degree 4, AMPA) -> Golgi cells (LIF, Euler;
                   in-degree 100, AMPA) and Purkinje cells (HH, BDF2 with the
                   FP64 Newton step, or the TimeDrivenPurkinjeCell model;
                   in-degree 2,000, AMPA, additive STDP); dt 0.1 ms.
  adjoint_fit      trace fitting: independent LIF (Euler) or AdEx (RK4)
                   neurons, 50 input synapses each (40 exc, 10 inh) driven by
                   fixed seeded Poisson trains at 20 Hz; per-neuron g_leak,
                   e_leak, tau_exc, tau_inh and per-synapse weights, with a
                   target parameter set perturbed by up to +-20%.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .network import binomial_thresholds

DT_S = np.float32(1e-4)


@dataclass
class NetworkSpec:
    model: str
    integrator: str
    n: int
    max_delay: int
    params: dict = field(default_factory=dict)
    # recurrent edges (src, tgt, delay steps, type, weight)
    rec: tuple | None = None
    # Poisson devices: list of (lcg_state, rate_hz, edges)
    poisson: list = field(default_factory=list)
    # external drive (thresholds, w_ext, seed)
    ext: tuple | None = None
    steps: int = 0


def lcg_seeds(seed: int, n: int) -> np.ndarray:
    """States of n RandomGenerator objects built after RandomGenerator::srand(seed)
    (Kernel/include/simulation/RandomGenerator.h:48-57, RandomGenerator.cpp:21-24):
    g_0 = seed, g_{i+1} = (765426813*g_i + 54321) mod 2^31."""
    out = np.empty(n, np.uint32)
    g = seed & 0xFFFFFFFF
    a, c, m = 765426813, 54321, 0x7FFFFFFF
    for i in range(n):
        out[i] = g
        g = (a * g + c) & m
    return out


def lif_poisson(n: int = 1000, n_exc: int = 100, n_inh: int = 25, rate: float = 10.0,
                steps: int = 10000, seed: int = 1) -> NetworkSpec:
    rng = np.random.default_rng(seed)
    per = n_exc + n_inh
    n_src = n * per
    src = np.arange(n_src, dtype=np.int32)
    tgt = np.repeat(np.arange(n, dtype=np.int32), per)
    is_inh = np.tile(np.r_[np.zeros(n_exc, bool), np.ones(n_inh, bool)], n)
    typ = is_inh.astype(np.uint8)
    w = np.where(is_inh, rng.uniform(0.0, 1.0, n_src), rng.uniform(0.0, 0.5, n_src)).astype(np.float32)
    delay = rng.integers(10, 51, n_src).astype(np.int32)  # 1..5 ms in 0.1 ms steps
    states = lcg_seeds(seed * 7919 + 17, n_src)
    spec = NetworkSpec("LIFTimeDrivenModel_GPU", "Euler", n, 50, steps=steps)
    spec.poisson.append((states, rate, (src, tgt, delay, typ, w)))
    return spec


def adex_recurrent(n: int = 1_000_000, in_degree: int = 200, exc_frac: float = 0.8,
                   max_delay: int = 10, n_ext: int = 100, ext_rate: float = 5.0,
                   w_ext: float = 1.0, steps: int = 10000, seed: int = 2) -> NetworkSpec:
    rng = np.random.default_rng(seed)
    k = min(in_degree, n - 1)
    src = rng.integers(0, n - 1, size=(n, k), dtype=np.int32)
    # redraw duplicates until every row holds distinct sources
    while True:
        src.sort(axis=1)
        dup = np.zeros_like(src, dtype=bool)
        dup[:, 1:] = src[:, 1:] == src[:, :-1]
        nd = int(dup.sum())
        if nd == 0:
            break
        src[dup] = rng.integers(0, n - 1, size=nd, dtype=np.int32)
    # no self-connections: sources drawn from n-1 values, shifted past the target
    tgt = np.repeat(np.arange(n, dtype=np.int32), k).reshape(n, k)
    src = src + (src >= tgt)
    src = src.reshape(-1)
    tgt = tgt.reshape(-1)
    n_exc = int(round(exc_frac * n))
    typ = (src >= n_exc).astype(np.uint8)
    w = rng.lognormal(-1.0, 0.5, size=src.shape[0]).astype(np.float32)
    delay = rng.integers(1, max_delay + 1, size=src.shape[0]).astype(np.int32)
    thr = binomial_thresholds(n_ext, float(ext_rate) * float(DT_S))
    spec = NetworkSpec("AdExTimeDrivenModel_GPU", "RK4", n, max_delay, steps=steps)
    spec.rec = (src, tgt, delay, typ, w)
    spec.ext = (thr, np.float32(w_ext), seed * 1000003 + 11)
    return spec


@dataclass
class CerebellumSpec:
    n_mf: int
    n_grc: int
    n_goc: int
    n_pc: int
    pc_model: str
    max_delay: int
    mf_states: np.ndarray
    mf_rates: np.ndarray
    # edge lists (src, tgt, delay steps, type, weight)
    mf_grc: tuple
    grc_goc: tuple
    grc_pc: tuple
    # additive STDP on granule -> Purkinje: (max_LTP, tau_LTP, max_LTD, tau_LTD, max_weight)
    stdp: tuple
    steps: int


def _fixed_in_degree(rng, n_src, n_tgt, k, d_lo, d_hi, w_lo, w_hi):
    src = rng.integers(0, n_src, n_tgt * k).astype(np.int32)
    tgt = np.repeat(np.arange(n_tgt, dtype=np.int32), k)
    delay = rng.integers(d_lo, d_hi + 1, n_tgt * k).astype(np.int32)
    typ = np.zeros(n_tgt * k, np.uint8)
    w = rng.uniform(w_lo, w_hi, n_tgt * k).astype(np.float32)
    return src, tgt, delay, typ, w


def cerebellum(n_grc: int = 200_000, n_goc: int = 2_000, n_pc: int = 100, n_mf: int = 10_000,
               k_grc: int = 4, k_goc: int = 100, k_pc: int = 2_000, pc_model: str = "HHTimeDrivenModel_GPU",
               max_delay: int = 10, steps: int = 20_000, seed: int = 3, w_goc: float = 0.02,
               w_pc: float | None = None) -> CerebellumSpec:
    """Layered cerebellar network.  Weights (nS): mossy fibre -> granule
    U[0.5, 1.5], granule -> Golgi U[0, w_goc], granule -> Purkinje U[0, w_pc]
    (default 0.3 for HH, 0.05 for the Purkinje-cell model), clamped by STDP
    to [0, 2 w_pc]; delays 1-5 steps into granule cells, 1-max_delay elsewhere."""
    rng = np.random.default_rng(seed)
    pc_w = w_pc if w_pc is not None else (0.3 if pc_model == "HHTimeDrivenModel_GPU" else 0.05)
    return CerebellumSpec(
        n_mf=n_mf, n_grc=n_grc, n_goc=n_goc, n_pc=n_pc, pc_model=pc_model, max_delay=max_delay,
        mf_states=lcg_seeds(seed * 7919 + 5, n_mf),
        mf_rates=rng.uniform(20.0, 80.0, n_mf).astype(np.float32),
        mf_grc=_fixed_in_degree(rng, n_mf, n_grc, k_grc, 1, min(5, max_delay), 0.5, 1.5),
        grc_goc=_fixed_in_degree(rng, n_grc, n_goc, k_goc, 1, max_delay, 0.0, w_goc),
        grc_pc=_fixed_in_degree(rng, n_grc, n_pc, k_pc, 1, max_delay, 0.0, pc_w),
        stdp=(0.005 * pc_w, 0.020, 0.0055 * pc_w, 0.020, 2.0 * pc_w),
        steps=steps)


@dataclass
class SweepSpec:
    sizes: np.ndarray            # int32 [networks]
    params: np.ndarray           # float32 [networks, 14] (AdEx parameter order)
    in_degree: int = 100
    exc_fraction: float = 0.8
    w_mu: float = -1.0
    w_sigma: float = 0.5
    max_delay: int = 10
    ext: tuple = ()              # (thresholds, w_ext, seed)
    wiring_seed: int = 0
    steps: int = 5000


SWEPT = ("g_leak", "e_leak", "a", "b", "tau_w")


def adex_sweep(n_networks: int = 4096, median: float = 5000.0, sigma: float = 0.5,
               lo: int = 1000, hi: int = 20000, spread: float = 0.3, w_ext: float = 0.7,
               steps: int = 5000, seed: int = 4) -> SweepSpec:
    from .models import AdExTimeDrivenModel_GPU

    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, n_networks)), lo, hi).astype(np.int32)
    names = [k for k, _ in AdExTimeDrivenModel_GPU.PARAMS]
    base = np.array([v for _, v in AdExTimeDrivenModel_GPU.PARAMS], np.float32)
    params = np.tile(base, (n_networks, 1))
    for k in SWEPT:
        j = names.index(k)
        params[:, j] = (base[j] * rng.uniform(1.0 - spread, 1.0 + spread, n_networks)).astype(np.float32)
    thr = binomial_thresholds(100, 5.0 * float(DT_S))
    return SweepSpec(sizes=sizes, params=params, ext=(thr, np.float32(w_ext), seed * 1000003 + 29),
                     wiring_seed=seed * 7777 + 3, steps=steps)


@dataclass
class FitSpec:
    model: str
    params: np.ndarray          # model parameter vector (GetDefaultParameters order)
    n: int
    steps: int
    syn_type: np.ndarray        # [n, n_syn] uint8
    events: tuple               # (ev_ptr [n+1], ev_step, ev_syn) int32
    start: dict                 # g_leak, e_leak, tau_exc, tau_inh [n], weight [n, n_syn]
    target: dict                # the perturbed set the target trace comes from


FIT_SCALE = {"LIFTimeDrivenModel_GPU": (0.03, 0.06), "AdExTimeDrivenModel_GPU": (2.0, 4.0)}


def adjoint_fit(model: str, n: int = 256, n_syn: int = 50, n_exc: int = 40, rate: float = 20.0,
                steps: int = 10000, perturb: float = 0.2, seed: int = 5) -> FitSpec:
    from .models import REGISTRY

    cls = REGISTRY[model]
    names = [k for k, _ in cls.PARAMS]
    params = np.array([v for _, v in cls.PARAMS], np.float32)
    rng = np.random.default_rng(seed)
    syn_type = np.zeros((n, n_syn), np.uint8)
    syn_type[:, n_exc:] = 1
    # fixed Poisson trains: each synapse fires before update k with probability rate*dt
    p = rate * float(DT_S)
    ev_neuron, ev_step, ev_syn = [], [], []
    chunk = 1000
    for k0 in range(0, steps, chunk):
        m = min(chunk, steps - k0)
        hit = rng.random((m, n, n_syn)) < p
        kk, ii, ss = np.nonzero(hit)
        ev_step.append(kk + k0 + 1)
        ev_neuron.append(ii)
        ev_syn.append(ss)
    ev_neuron = np.concatenate(ev_neuron)
    ev_step = np.concatenate(ev_step)
    ev_syn = np.concatenate(ev_syn)
    order = np.lexsort((ev_syn, ev_step, ev_neuron))
    ev_neuron, ev_step, ev_syn = ev_neuron[order], ev_step[order], ev_syn[order]
    ev_ptr = np.zeros(n + 1, np.int64)
    np.add.at(ev_ptr, ev_neuron + 1, 1)
    ev_ptr = np.cumsum(ev_ptr)
    we, wi = FIT_SCALE[model]
    weight = np.where(syn_type == 0, rng.uniform(0, we, (n, n_syn)), rng.uniform(0, wi, (n, n_syn)))
    start = {
        "g_leak": np.full(n, params[names.index("g_leak")], np.float32),
        "e_leak": np.full(n, params[names.index("e_leak")], np.float32),
        "tau_exc": np.full(n, params[names.index("tau_exc")], np.float32),
        "tau_inh": np.full(n, params[names.index("tau_inh")], np.float32),
        "weight": weight.astype(np.float32),
    }
    target = {k: (v * rng.uniform(1 - perturb, 1 + perturb, v.shape)).astype(np.float32) for k, v in start.items()}
    return FitSpec(model, params, n, steps, syn_type,
                   (ev_ptr.astype(np.int32), ev_step.astype(np.int32), ev_syn.astype(np.int32)), start, target)
