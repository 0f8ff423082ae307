"""Brute-force model of the reference's event-driven delivery, for ordering tests.

Restates, in plain Python, the pieces of EDLUT's CPU engine that decide in
which order input spikes reach a time-driven population:

  * EventQueue: a 1-based binary heap keyed by (time, priority), with the
    sift-up / sift-down rules of Kernel/src/simulation/EventQueue.cpp:121-198
    (equal times: higher EventPriority first, Kernel/include/simulation/Event.h:37);
  * TimeEventUpdateNeurons: one event per model, next one at t + its step
    (Kernel/src/simulation/TimeEventUpdateNeurons.cpp:45-79), first at one
    step (Kernel/src/simulation/Simulation.cpp:228-268); a population's step
    is double(float dt), a PoissonGeneratorDeviceVector's 0.1f / 1000.0f in
    float (PoissonGeneratorDeviceVector.cpp:26,137);
  * RunSimulationSlot: a StopSimulationEvent at the stop time, events until
    it leaves (Kernel/src/simulation/Simulation.cpp:322-373);
  * TimeDrivenInternalSpike: spiking sources in ascending index, one
    PropagatedSpikeGroup per model delay (delays sorted ascending,
    NeuronModelPropagationDelayStructure.cpp:42-73), a group is queued as soon
    as it holds MaxSize = 1024 sources and the rest at the end
    (TimeDrivenInternalSpike.cpp:66-104,163-178, PropagatedSpikeGroup.cpp:37-45);
  * PropagatedSpikeGroup: per source, its synapses of that delay in the order
    of Network::FindOutConnections (source, delay, target; Network.cpp:45-68);
  * delays: RoundPropagationDelay (Network.cpp:1311-1317) of the float delay.

The model records, for every update of the target population, the sequence
of (target, type, weight) increments applied before it.
"""
from __future__ import annotations

import math

import numpy as np

# Event.h:37
ENDSIMULATIONEVENT, STOPSIMULATIONEVENT, PROPAGATEDSPIKE, TIMEEVENT = 0, 1, 8, 10
MAX_GROUP = 1024
POISSON_STEP = float(np.float32(0.1) / np.float32(1000.0))


class Heap:
    """EventQueue restricted to one queue."""

    def __init__(self):
        self.e = [None]  # slot 0 unused; len(self.e) == NumberOfElements

    def size(self):
        return len(self.e) - 1

    @staticmethod
    def _gt(a, b):  # a should sit below b
        return a[0] > b[0] or (a[0] == b[0] and a[1] < b[1])

    def insert(self, time, prio, payload):
        self.e.append((time, prio, payload))
        c = self.size()
        while c > 1 and self._gt(self.e[c // 2], self.e[c]):
            self.e[c], self.e[c // 2] = self.e[c // 2], self.e[c]
            c //= 2

    def remove(self):
        n = len(self.e)
        if n > 2:
            first = self.e[1]
            self.e[1] = self.e[self.size()]
            self.e.pop()
            p = 1
            c = 2
            while c < self.size():
                a, b = self.e[c], self.e[c + 1]
                if a[0] > b[0] or (a[0] == b[0] and a[1] < b[1]):
                    c += 1
                x, y = self.e[c], self.e[p]
                if x[0] < y[0] or (x[0] == y[0] and x[1] > y[1]):
                    self.e[p], self.e[c] = self.e[c], self.e[p]
                else:
                    break
                p = c
                c = p * 2
            if c == self.size():
                x, y = self.e[c], self.e[p]
                if x[0] < y[0] or (x[0] == y[0] and x[1] > y[1]):
                    self.e[p], self.e[c] = self.e[c], self.e[p]
            return first
        if n == 2:
            first = self.e[1]
            self.e.pop()
            return first
        return None


def round_delay(delay_s: float) -> float:
    """Network::RoundPropagationDelay with minpropagationdelay = 0.0001."""
    m = 0.0001
    r = math.floor(delay_s * (1.0 / 0.0001) + m * 0.5) * m
    return m if r <= 0 else r


class SourceModel:
    """A time-driven spike source: raster[s] = ascending list of sources
    firing at its s-th update; edges grouped per (source, delay)."""

    def __init__(self, raster, src, tgt, delay_steps, typ, w):
        self.raster = raster
        order = np.lexsort((tgt, delay_steps, src))
        self.syn = {}
        d_s = [round_delay(float(np.float32(d * 1e-4))) for d in delay_steps]
        self.delays = sorted(set(d_s))
        for q in order:
            self.syn.setdefault((int(src[q]), d_s[q]), []).append((int(tgt[q]), int(typ[q]), float(w[q])))
        self.src_delays = {}
        for (s, d) in self.syn:
            self.src_delays.setdefault(s, set()).add(d)


def simulate(models, n_updates: int, dt_s=np.float32(1e-4), target_first: bool = True, steps=None,
             runs=None, self_model=None):
    """Run the event loop; `models` are SourceModel objects, each updated by
    its own update events at steps[i] seconds (default: the target's step),
    like the target population.  `runs` splits the updates into successive
    RunSimulationSlot calls, each stopping at the time of its last update.
    `self_model` is the target's own recurrent source (its raster is the
    target's spikes), emitted by the target's update events.
    Returns {update k: [(tgt, type, w), ...]} for the target's updates."""
    step = float(np.float32(dt_s))
    steps = [step] * len(models) if steps is None else [float(x) for x in steps]
    runs = [n_updates] if runs is None else list(runs)
    h = Heap()
    order = ["target"] + list(range(len(models))) if target_first else list(range(len(models))) + ["target"]
    for m in order:
        h.insert(step if m == "target" else steps[m], TIMEEVENT, ("time", m, 1))
    out = {}
    pending = []
    done = 0
    for run in runs:
        last = done + run
        t_stop = 0.0
        for _ in range(last):
            t_stop += step
        h.insert(t_stop, STOPSIMULATIONEVENT, ("stop",))
        while True:
            ev = h.remove()
            t, prio, payload = ev
            if payload[0] == "stop":
                break
            if payload[0] == "time":
                _, m, s = payload
                if m == "target":
                    out[s] = pending
                    pending = []
                    if self_model is None:
                        h.insert(t + step, TIMEEVENT, ("time", m, s + 1))
                        continue
                mod = models[m] if m != "target" else self_model
                fired = mod.raster[s] if s < len(mod.raster) else []
                groups = {d: [] for d in mod.delays}
                for src in fired:
                    for d in sorted(mod.src_delays.get(src, ())):
                        groups[d].append(src)
                        if len(groups[d]) >= MAX_GROUP:
                            h.insert(t + d, PROPAGATEDSPIKE, ("group", m, list(groups[d]), d))
                            groups[d] = []
                for d in mod.delays:
                    if groups[d]:
                        h.insert(t + d, PROPAGATEDSPIKE, ("group", m, groups[d], d))
                h.insert(t + (step if m == "target" else steps[m]), TIMEEVENT, ("time", m, s + 1))
            else:
                _, m, srcs, d = payload
                mod = models[m] if m != "target" else self_model
                for src in srcs:
                    pending.extend(mod.syn[(src, d)])
        done = last
    return out
