"""The oracle's delivery order is the reference engine's.

The oracle restates the reference's event loop (one binary heap of update
events, spike groups and input spikes; EventQueue.cpp:121-198).  Pins:

  * its heap equals the reference's own EventQueue.cpp (compiled from the
    reference's sources into oracle/_ref) on random operation sequences full
    of equal keys;
  * driven with the same spike rasters, it applies the same increments in the
    same order, update by update, as tests/eventqueue_model.py, an
    independent Python restatement of the heap, the update events (one stream
    per model at its own step), TimeDrivenInternalSpike's groups of 1,024
    sources and RunSimulationSlot's stop events.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.eventqueue_model import POISSON_STEP, Heap, SourceModel, simulate
from tests.helpers import LIF_DEFAULTS, np_out_csr


def lcg_raster(states, rate, n_steps, step_s=POISSON_STEP):
    """Spiking generators of a PoissonGeneratorDeviceVector per update (1-based):
    frequency * step (double) > frand()."""
    x = states.astype(np.uint64)
    p = float(np.float32(rate)) * float(step_s)
    raster = [[]]
    for _ in range(n_steps):
        x = (x * np.uint64(1103515245) + np.uint64(12345)) & np.uint64(0x7FFFFFFF)
        u = x.astype(np.float32) / np.float32(2147483648.0)
        raster.append(np.nonzero(p > u.astype(np.float64))[0].tolist())
    return raster


def random_ops(rng, n_ops, n_times, p_insert=0.6):
    op = (rng.random(n_ops) < p_insert).astype(np.int8)
    t = rng.integers(0, n_times, n_ops).astype(np.float64) * 0.25
    prio = rng.choice(np.array([1, 8, 10], np.int32), n_ops)
    return op, t, prio


def py_heap_replay(op, t, prio):
    h = Heap()
    out, nxt = [], 0
    for o, tt, pp in zip(op, t, prio):
        if o:
            h.insert(float(tt), int(pp), nxt)
            nxt += 1
        elif h.size() > 0:
            out.append(h.remove()[2])
    return np.array(out, np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_heap_equals_reference_eventqueue(oracle, seed):
    """The oracle's heap against the reference's EventQueue.cpp, built from
    its own sources (oracle/_ref/ref_eventqueue): identical removal order."""
    if not os.path.exists(oracle.REF_EVENTQUEUE):
        pytest.skip("oracle/_ref/ref_eventqueue not built (needs the reference checkout)")
    rng = np.random.default_rng(seed)
    op, t, prio = random_ops(rng, 20000, 6 + 20 * seed)
    with tempfile.TemporaryDirectory() as d:
        rec = np.zeros(len(op), dtype=[("op", "i1"), ("t", "<f8"), ("p", "<i4")])
        rec["op"], rec["t"], rec["p"] = op, t, prio
        with open(os.path.join(d, "ops.bin"), "wb") as f:
            f.write(np.int64(len(op)).tobytes())
            f.write(rec.tobytes())
        subprocess.run([oracle.REF_EVENTQUEUE, os.path.join(d, "ops.bin"), os.path.join(d, "out.bin")], check=True)
        ref = np.fromfile(os.path.join(d, "out.bin"), np.int32)
    got = oracle.heap_replay(op, t, prio)
    assert len(ref) > 5000
    assert np.array_equal(got, ref)
    # the Python model of the tests agrees, and ties really are permuted
    assert np.array_equal(py_heap_replay(op, t, prio), ref)
    order = np.argsort(t[op == 1], kind="stable")
    assert not np.array_equal(ref[: len(order)], order[: len(ref)])


def test_heap_reorders_equal_keys():
    """Five equal-key inserts leave in the order the heap's layout gives (the
    reference's EventQueue returns 0 4 3 2 1 for this sequence)."""
    op = np.array([1] * 5 + [0] * 5, np.int8)
    t = np.ones(10)
    prio = np.full(10, 8, np.int32)
    assert py_heap_replay(op, t, prio).tolist() == [0, 4, 3, 2, 1]


@pytest.fixture(autouse=True)
def heap_order(oracle):
    oracle.set_order(1)
    yield
    oracle.set_order(0)


def oracle_trace(oracle, n_tgt, sources, steps, runs=None, self_edges=None, target_first=True):
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "Euler", LIF_DEFAULTS, n_tgt, mask=7)
    net = oracle.Net(pop, 16)
    if self_edges is not None:
        net.add_self(*np_out_csr(n_tgt, n_tgt, 16, *self_edges))
    for states, rate, (s, t, d, y, w) in sources:
        net.add_poisson(states, rate, *np_out_csr(len(states), n_tgt, 16, s, t, d, y, w))
    if not target_first:
        net.set_model_rank(5000)
    net.set_trace(True)
    st_all, ix_all = [], []
    k = 1
    for r in runs or [steps]:
        st, ix = net.run(k, r)
        st_all.append(st)
        ix_all.append(ix)
        k += r
    st, tg, ty, w = net.trace()
    out = {}
    for k in range(1, steps + 1):
        m = st == k
        out[k] = list(zip(tg[m].tolist(), ty[m].tolist(), w[m].astype(float).tolist()))
    return out, np.concatenate(st_all), np.concatenate(ix_all)


def make_source(rng, n_src, n_tgt, fanout, max_d, seed):
    src = np.repeat(np.arange(n_src), fanout)
    tgt = np.concatenate([rng.choice(n_tgt, fanout, replace=False) for _ in range(n_src)])
    d = rng.integers(1, max_d + 1, len(src))
    y = rng.integers(0, 3, len(src)).astype(np.uint8)
    w = rng.random(len(src)).astype(np.float32)
    states = np.array([(seed * 7919 + 13 * i) & 0x7FFFFFFF for i in range(n_src)], np.uint32)
    return states, (src, tgt, d, y, w)


def compare(a, b, steps):
    for k in range(1, steps + 1):
        assert [(t, y, np.float32(w)) for t, y, w in a[k]] == [(t, y, np.float32(w)) for t, y, w in b[k]], k


@pytest.mark.parametrize("target_first", [True, False])
def test_single_source_model_order(oracle, target_first):
    """One Poisson layer: its groups arrive after the update they follow, in
    descending delay (its step, 0.1f/1000f s, exceeds the population's
    float(1e-4) s, so later emissions arrive later)."""
    rng = np.random.default_rng(1)
    steps = 300
    states, edges = make_source(rng, 200, 30, 3, 6, 5)
    model = SourceModel(lcg_raster(states, 200.0, steps), *edges)
    ref = simulate([model], steps, target_first=target_first, steps=[POISSON_STEP])
    got, _, _ = oracle_trace(oracle, 30, [(states, 200.0, edges)], steps, target_first=target_first)
    assert sum(len(v) for v in ref.values()) > 1000
    compare(ref, got, steps)
    # the canonical (ascending delay) order differs
    oracle.set_order(0)
    canon, _, _ = oracle_trace(oracle, 30, [(states, 200.0, edges)], steps)
    oracle.set_order(1)
    assert any(canon[k] != got[k] for k in range(1, steps + 1))


def test_large_groups_split_at_1024_sources(oracle):
    """More than 1023 sources of one model firing at one update with one delay:
    the groups of 1,024 share their arrival time and leave the heap in the
    order its layout gives (not always ascending); the oracle follows it."""
    rng = np.random.default_rng(2)
    steps = 40
    n_src = 4000
    src = np.arange(n_src)
    tgt = rng.integers(0, 20, n_src)
    d = np.full(n_src, 2)
    y = np.zeros(n_src, np.uint8)
    w = rng.random(n_src).astype(np.float32)
    states = np.array([(77 + 31 * i) & 0x7FFFFFFF for i in range(n_src)], np.uint32)
    model = SourceModel(lcg_raster(states, 6000.0, steps), src, tgt, d, y, w)
    assert max(len(r) for r in model.raster) > 2048
    ref = simulate([model], steps, steps=[POISSON_STEP])
    got, _, _ = oracle_trace(oracle, 20, [(states, 6000.0, (src, tgt, d, y, w))], steps)
    compare(ref, got, steps)
    # some update applies the groups out of ascending source order: the
    # canonical order differs there, by whole groups only
    oracle.set_order(0)
    canon, _, _ = oracle_trace(oracle, 20, [(states, 6000.0, (src, tgt, d, y, w))], steps)
    oracle.set_order(1)
    permuted = [k for k in range(1, steps + 1) if canon[k] != got[k]]
    assert permuted
    for k in permuted:
        assert sorted(canon[k]) == sorted(got[k])


def test_two_poisson_models_tie(oracle):
    """Two Poisson layers of different types (two models) with equal steps:
    their groups of one update and delay share a time; the heap orders them."""
    rng = np.random.default_rng(3)
    steps = 200
    s1, e1 = make_source(rng, 150, 25, 4, 5, 7)
    s2, e2 = make_source(rng, 120, 25, 4, 5, 9)
    m1 = SourceModel(lcg_raster(s1, 300.0, steps), *e1)
    m2 = SourceModel(lcg_raster(s2, 250.0, steps), *e2)
    ref = simulate([m1, m2], steps, steps=[POISSON_STEP, POISSON_STEP], runs=[37, 100, 63])
    got, _, _ = oracle_trace(oracle, 25, [(s1, 300.0, e1), (s2, 250.0, e2)], steps, runs=[37, 100, 63])
    assert sum(len(v) for v in ref.values()) > 2000
    compare(ref, got, steps)


def test_recurrent_and_poisson_burst(oracle):
    """A population with recurrent synapses (its own groups, split at 1,024
    sources in bursts) and a Poisson layer; runs split by stop events."""
    rng = np.random.default_rng(4)
    steps, n = 120, 3000
    src = np.repeat(np.arange(n), 30)
    tgt = rng.integers(0, n, len(src))
    d = rng.integers(1, 4, len(src))
    y = np.zeros(len(src), np.uint8)
    w = rng.uniform(10.0, 20.0, len(src)).astype(np.float32)
    states, pe = make_source(rng, 400, n, 20, 4, 3)
    pe = (pe[0], pe[1], pe[2], np.zeros(len(pe[0]), np.uint8), (pe[4] * 20.0).astype(np.float32))
    runs = [50, 7, 63]
    got, st, ix = oracle_trace(oracle, n, [(states, 2000.0, pe)], steps, runs=runs, self_edges=(src, tgt, d, y, w))
    raster = [[] for _ in range(steps + 1)]
    for k, i in zip(st.tolist(), ix.tolist()):
        raster[k].append(i)
    assert max(len(r) for r in raster) > 1024, "the recurrent source should burst"
    self_model = SourceModel(raster, src, tgt, d, y, w)
    pm = SourceModel(lcg_raster(states, 2000.0, steps), *pe)
    ref = simulate([pm], steps, steps=[POISSON_STEP], runs=runs, self_model=self_model)
    compare(ref, got, steps)


def test_spike_train_source_delivers_like_the_poisson_device(oracle):
    """A spike-train source whose input spikes sit at update times: a lone
    spike at update s through a delay of d updates arrives at update s+d+1."""
    pop = oracle.Population("LIFTimeDrivenModel_GPU", "Euler", LIF_DEFAULTS, 2, mask=7)
    net = oracle.Net(pop, 16)
    ptr = np.zeros(8, np.int64)
    ptr[5:] = 1                                   # source 0 fires at update 5
    net.add_train(1, ptr, np.array([0], np.int32),
                  *np_out_csr(1, 2, 16, [0], [1], [3], np.array([1], np.uint8), np.array([0.25], np.float32)))
    net.set_trace(True)
    net.run(1, 20)
    st, tg, ty, w = net.trace()
    assert st.tolist() == [5 + 3 + 1] and tg.tolist() == [1] and ty.tolist() == [1] and w.tolist() == [0.25]
