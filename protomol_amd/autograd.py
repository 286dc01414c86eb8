"""torch.autograd binding of a GPU force group.

The potential energy of a group of nonbonded cutoff and PME forces is a
differentiable function of the positions, the scaled charges and, with
per-atom LJ parameters, sigma and epsilon:

* first order: -forces (pm_evaluate), dE/dq, dE/dsigma, dE/depsilon
  (pm_param_gradients);
* second order in the positions: backward of the position gradient is the
  Hessian-vector product (pm_hessian_vec), so torch.autograd.grad with
  create_graph=True, torch.autograd.functional.hvp and gradgradcheck work.
  Mixed second derivatives (positions x parameters) are not provided.

Tensors may live on any device; values cross the C ABI as host float64
arrays (the ABI's pointer convention, as in the reference's Vector3DBlock).
"""
from __future__ import annotations

import numpy as np
import torch

from . import gpu


def _np(t):
    # a copy: numpy() of a CPU float64 tensor shares its storage, and a caller
    # may change the tensor in place after the call
    return np.array(t.detach().to("cpu", torch.float64).numpy(), dtype=np.float64, copy=True, order="C")


class ForceGroup:
    """A gpu.Context plus the force handles whose summed energy is the
    function; remembers the charges and LJ parameters last sent so that a call
    with unchanged values sends nothing."""

    def __init__(self, ctx: gpu.Context, handles, charges=None, sigma=None, epsilon=None):
        self.ctx = ctx
        self.handles = list(handles)
        self._q = None if charges is None else np.array(charges, np.float64)
        self._lj = None if sigma is None else (np.array(sigma, np.float64), np.array(epsilon, np.float64))

    def sync(self, q, sigma, eps):
        if q is not None:
            qn = _np(q)
            if self._q is None or not np.array_equal(qn, self._q):
                self.ctx.set_charges(qn)
                self._q = qn
        if sigma is not None:
            sn, en = _np(sigma), _np(eps)
            if self._lj is None or not (np.array_equal(sn, self._lj[0]) and np.array_equal(en, self._lj[1])):
                self.ctx.set_lj_per_atom(sn, en)
                self._lj = (sn, en)

    def energy(self, pos, q=None, sigma=None, epsilon=None):
        """E(pos[, q, sigma, epsilon]) as a float64 scalar tensor on pos.device."""
        if (sigma is None) != (epsilon is None):
            raise ValueError("sigma and epsilon go together")
        return _Energy.apply(pos, q, sigma, epsilon, self)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, q, sigma, eps, group):
        group.sync(q, sigma, eps)
        f, e = group.ctx.evaluate(group.handles, _np(pos))
        ctx.group = group
        ctx.save_for_backward(pos, q, sigma, eps)
        ctx.minus_f = torch.from_numpy(-f).to(pos.device)
        return torch.tensor(float(e[:4].sum()), dtype=torch.float64, device=pos.device)

    @staticmethod
    def backward(ctx, g):
        pos, q, sigma, eps = ctx.saved_tensors
        group = ctx.group
        gpos = gq = gs = ge = None
        if ctx.needs_input_grad[0]:
            gpos = g * _Gradient.apply(pos, q, sigma, eps, group, ctx.minus_f)
        if any(ctx.needs_input_grad[1:4]):
            group.sync(q, sigma, eps)
            want_lj = sigma is not None and (ctx.needs_input_grad[2] or ctx.needs_input_grad[3])
            dq, ds, de = group.ctx.param_gradients(group.handles, _np(pos), lj=want_lj)
            dev = pos.device
            if ctx.needs_input_grad[1]:
                gq = g * torch.from_numpy(dq).to(dev)
            if want_lj and ctx.needs_input_grad[2]:
                gs = g * torch.from_numpy(ds).to(dev)
            if want_lj and ctx.needs_input_grad[3]:
                ge = g * torch.from_numpy(de).to(dev)
        return gpos, gq, gs, ge, None


class _Gradient(torch.autograd.Function):
    """dE/dpos (= -forces, computed by the forward evaluation); its backward
    is the Hessian-vector product."""

    @staticmethod
    def forward(ctx, pos, q, sigma, eps, group, minus_f):
        ctx.group = group
        ctx.save_for_backward(pos, q, sigma, eps)
        return minus_f.clone()

    @staticmethod
    def backward(ctx, gg):
        pos, q, sigma, eps = ctx.saved_tensors
        return _HessianVec.apply(pos, q, sigma, eps, ctx.group, gg), None, None, None, None, None


class _HessianVec(torch.autograd.Function):
    """H u as a function of u (linear, H symmetric: its backward is H again),
    so that torch's double-backward constructions (functional.hvp) see it."""

    @staticmethod
    def forward(ctx, pos, q, sigma, eps, group, u):
        ctx.group = group
        ctx.save_for_backward(pos, q, sigma, eps)
        group.sync(q, sigma, eps)
        hv = group.ctx.hessian_vec(group.handles, _np(pos), _np(u))
        return torch.from_numpy(hv).to(u.device, u.dtype)

    @staticmethod
    def backward(ctx, g):
        pos, q, sigma, eps = ctx.saved_tensors
        return None, None, None, None, None, _HessianVec.apply(pos, q, sigma, eps, ctx.group, g)
