"""Differentiable trace fitting on the GPU: a torch.autograd.Function over the
HIP forward simulation and its discrete adjoint (edlut_fit_* in
include/edlut_amd.h).

    prob = FitProblem("AdExTimeDrivenModel_GPU", params, n, steps, syn_type, events)
    v = prob.simulate(g_leak, e_leak, tau_exc, tau_inh, weight)   # [steps, n]
    loss = ((v - target) ** 2).mean(); loss.backward()

The tensors live on the GPU; torch provides device memory and the stream,
the kernels are the library's.  Gradients are those of the discrete forward
map with the spike pattern of the forward run held fixed.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C
from ._capi import EDLUTException

_MODELS = {"LIFTimeDrivenModel_GPU": C.EDLUT_LIF, "AdExTimeDrivenModel_GPU": C.EDLUT_ADEX}


class _Tensors(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "g_leak", "e_leak", "tau_exc", "tau_inh", "weight", "syn_type", "ev_ptr", "ev_step", "ev_syn",
        "v_trace", "spikes", "traj", "traj_since", "grad_v", "grad_par", "grad_w")]


def _key(*ts):
    """Identity of a parameter set: storage and in-place version of each tensor."""
    return tuple((t.data_ptr(), t._version) for t in ts)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class FitProblem:
    """Fixed part of a fitting problem: model, fixed parameters, input trains.

    params: the model's full parameter vector (GetDefaultParameters order);
    syn_type: [n, n_syn] 0 exc / 1 inh; events: (ev_ptr [n+1], ev_step, ev_syn)
    with each neuron's events sorted by step (ev_step = the update they precede).
    """

    def __init__(self, model: str, params, n: int, steps: int, syn_type, events, dt_s=1e-4,
                 checkpoint: int = 100, device: int = 0):
        if model not in _MODELS:
            raise EDLUTException(C.EDLUT_E_ARG, f"trace fitting takes LIF or AdEx, not {model!r}")
        self.model, self.n, self.steps = model, int(n), int(steps)
        self.dev = torch.device("cuda", device)
        st = np.ascontiguousarray(syn_type, np.uint8)
        self.n_syn = st.shape[1]
        p = np.ascontiguousarray(params, np.float32)
        h = ctypes.c_void_p()
        C.check(C.lib().edlut_fit_create(_MODELS[model], C.ptr(p), len(p), float(np.float32(dt_s)), self.n,
                                         self.steps, self.n_syn, int(checkpoint), int(device), C.byref(h)))
        self._h = h
        self._last = None
        self.syn_type = torch.as_tensor(st, device=self.dev)
        ev_ptr, ev_step, ev_syn = events
        self.ev_ptr = torch.as_tensor(np.ascontiguousarray(ev_ptr, np.int32), device=self.dev)
        self.ev_step = torch.as_tensor(np.ascontiguousarray(ev_step, np.int32), device=self.dev)
        self.ev_syn = torch.as_tensor(np.ascontiguousarray(ev_syn, np.int32), device=self.dev)

    def close(self):
        if getattr(self, "_h", None):
            C.lib().edlut_fit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _tensors(self, gl, el, te, ti, w, v=None, gv=None, gp=None, gw=None, spk=None, traj=None,
                 since=None):
        for x in (gl, el, te, ti, w):
            if x.dtype != torch.float32 or not x.is_contiguous() or x.device != self.dev:
                raise EDLUTException(C.EDLUT_E_ARG, "parameters must be contiguous float32 on the GPU")
        if gl.shape != (self.n,) or w.shape != (self.n, self.n_syn):
            raise EDLUTException(C.EDLUT_E_ARG, "parameter shapes do not match the problem")
        t = _Tensors()
        t.g_leak, t.e_leak, t.tau_exc, t.tau_inh = (_ptr(x) for x in (gl, el, te, ti))
        t.weight = _ptr(w)
        t.syn_type = _ptr(self.syn_type)
        t.ev_ptr, t.ev_step, t.ev_syn = _ptr(self.ev_ptr), _ptr(self.ev_step), _ptr(self.ev_syn)
        t.v_trace, t.grad_v, t.grad_par, t.grad_w = _ptr(v), _ptr(gv), _ptr(gp), _ptr(gw)
        t.spikes = _ptr(spk)
        t.traj, t.traj_since = _ptr(traj), _ptr(since)
        return t

    def forward_raw(self, gl, el, te, ti, w, spikes: bool = False, trajectory: bool = False):
        """V trace [steps, n]; with spikes, also the spike flags [steps, n]; with
        trajectory, also the states before each step's integration
        [steps, 4, n] and the refractory counts [steps, n]."""
        v = torch.empty((self.steps, self.n), dtype=torch.float32, device=self.dev)
        spk = torch.empty((self.steps, self.n), dtype=torch.uint8, device=self.dev) if spikes else None
        traj = since = None
        if trajectory:
            traj = torch.empty((self.steps, 4, self.n), dtype=torch.float32, device=self.dev)
            since = torch.empty((self.steps, self.n), dtype=torch.int32, device=self.dev)
        t = self._tensors(gl, el, te, ti, w, v=v, spk=spk, traj=traj, since=since)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        C.check(C.lib().edlut_fit_forward(self._h, ctypes.byref(t), stream))
        self._last = _key(gl, el, te, ti, w)
        out = [v]
        if spikes:
            out.append(spk)
        if trajectory:
            out += [traj, since]
        return tuple(out) if len(out) > 1 else v

    def backward_raw(self, gl, el, te, ti, w, grad_v):
        grad_v = grad_v.contiguous().to(torch.float32)
        gp = torch.empty((4, self.n), dtype=torch.float32, device=self.dev)
        gw = torch.empty((self.n, self.n_syn), dtype=torch.float32, device=self.dev)
        t = self._tensors(gl, el, te, ti, w, gv=grad_v, gp=gp, gw=gw)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        C.check(C.lib().edlut_fit_backward(self._h, ctypes.byref(t), stream))
        return gp, gw

    def simulate(self, g_leak, e_leak, tau_exc, tau_inh, weight):
        """V trace [steps, n], differentiable w.r.t. the five parameter tensors."""
        return NeuronTrace.apply(self, g_leak, e_leak, tau_exc, tau_inh, weight)


class NeuronTrace(torch.autograd.Function):
    """V trace of a FitProblem; backward runs the adjoint kernel, which replays
    the forward pass from its checkpoints (the forward pass must be the last
    one run with these parameters)."""

    @staticmethod
    def forward(ctx, prob, g_leak, e_leak, tau_exc, tau_inh, weight):
        args = [x.detach().contiguous() for x in (g_leak, e_leak, tau_exc, tau_inh, weight)]
        v = prob.forward_raw(*args)
        ctx.prob = prob
        ctx.save_for_backward(*args)
        return v

    @staticmethod
    def backward(ctx, grad_v):
        prob = ctx.prob
        gl, el, te, ti, w = ctx.saved_tensors
        # the adjoint replays segments from the checkpoints of the forward
        # run; rerun it if another forward pass has replaced them since
        if prob._last != _key(gl, el, te, ti, w):
            prob.forward_raw(gl, el, te, ti, w)
        gp, gw = prob.backward_raw(gl, el, te, ti, w, grad_v)
        return None, gp[0], gp[1], gp[2], gp[3], gw


def refractory_table(dt_s: float, tau_ref: float) -> np.ndarray:
    n = ctypes.c_int(0)
    C.check(C.lib().edlut_refractory_table(float(dt_s), float(tau_ref), None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), np.float32)
    C.check(C.lib().edlut_refractory_table(float(dt_s), float(tau_ref), C.ptr(out), len(out), C.byref(n)))
    return out[: n.value]
