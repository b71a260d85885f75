"""torch.autograd Functions of the propagators and of the fused BCLP loss.

Every backward is a hand-written adjoint kernel of libvamx_hip.so:

* ``ParallelProject``      forward: sinogram of a volume (D1 forward projector);
                           backward: exact transpose (vamx_par_forward_adjoint).
* ``ParallelBackproject``  forward: dose of a sinogram (D1 back projector, circle clip);
                           backward: exact transpose (vamx_par_backward_adjoint).
* ``RayTrace`` / ``RayDeposit``  the ray tracer's integrate / deposit, each the
                           other's adjoint; trajectories are recomputed, not stored.
* ``ResponseMap``          the material response m = map(dose) of
                           vamtoolbox/response.py:191-291 (one kernel,
                           vamx_response_map, that also keeps map'(dose));
                           backward multiplies by the kept derivative.
* ``BCLPLoss``             L(g) of vamtoolbox/optimizer/BCLP.py:332-351 with the
                           dose = B g, the response, the band and the Lp weighting
                           fused into one kernel (vamx_bclp_eval) that also emits the
                           operand v w |e-eps|^(p-1) sign(e) map'(dose); backward is
                           q L^((q-p)/q) B^T operand -- the exact gradient.
* ``RayTraceBCLPLoss``     the same loss on the ray tracer (vamx_rt_bclp_eval: the
                           deposit fused with the response, band and Lp weighting);
                           backward is q L^((q-p)/q) W (alpha_abs operand), the
                           integrate pass -- the exact gradient.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


class ParallelProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, projector):
        ctx.projector = projector
        return projector.plan.forward(vol.detach().float().contiguous())

    @staticmethod
    def backward(ctx, grad):
        return ctx.projector.plan.forward_adjoint(grad.float().contiguous()), None


class ParallelBackproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, projector):
        ctx.projector = projector
        return projector.plan.backward(sino.detach().float().contiguous())

    @staticmethod
    def backward(ctx, grad):
        return ctx.projector.plan.backward_adjoint(grad.float().contiguous()), None


class RayTrace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, propagator):
        ctx.propagator = propagator
        return propagator.forward(f.detach().float().contiguous())

    @staticmethod
    def backward(ctx, grad):
        return ctx.propagator.backward(grad.float().contiguous()).reshape(ctx.propagator.vol_shape), None


class RayDeposit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, propagator):
        ctx.propagator = propagator
        return propagator.backward(g.detach().float().contiguous())

    @staticmethod
    def backward(ctx, grad):
        return ctx.propagator.forward(grad.float().contiguous()), None


class ResponseMap(torch.autograd.Function):
    """m = response_model.map(x) on the device; d m / d x = response_model.dmapdf(x)."""

    @staticmethod
    def forward(ctx, x, response_model):
        lib = _lib.load()
        xf = x.detach().float().contiguous()
        mapped = torch.empty_like(xf)
        dmap = torch.empty_like(xf)
        resp, tabs = response_model.vamx_struct(xf.device)
        with torch.cuda.device(xf.device):
            _lib.check(lib.vamx_response_map(ctypes.byref(resp), xf.data_ptr(), mapped.data_ptr(), dmap.data_ptr(),
                                             xf.numel(), _lib.stream_ptr(xf.device)), "vamx_response_map")
        del tabs
        ctx.save_for_backward(dmap)
        return mapped

    @staticmethod
    def backward(ctx, grad):
        (dmap,) = ctx.saved_tensors
        return grad.float() * dmap, None


def _bclp_params(eps, weight, p, q, dev):
    eps_t = None if np.isscalar(eps) else torch.as_tensor(eps, device=dev, dtype=torch.float32).contiguous()
    w_t = None if np.isscalar(weight) else torch.as_tensor(weight, device=dev, dtype=torch.float32).contiguous()
    prm = _lib.BclpParams(p=float(p), q=float(q), eps=float(eps) if eps_t is None else 0.0,
                          weight=float(weight) if w_t is None else 0.0, learning_rate=0.0,
                          glb=-np.inf, gub=np.inf, dvol=1.0, grad_op=0)
    return prm, eps_t, w_t


def _loss_scale(L, p, q):
    return torch.where(L != 0, q * L.clamp_min(1e-300) ** ((q - p) / q), torch.ones_like(L))


class RayTraceBCLPLoss(torch.autograd.Function):
    """L(g) for the ray tracer: the deposit fused with the response, band and
    Lp weighting (vamx_rt_bclp_eval); backward = q L^((q-p)/q) W (alpha_abs operand)."""

    @staticmethod
    def forward(ctx, g, propagator, response_model, target, eps=0.1, weight=1.0, p=2.0, q=1.0):
        dev = g.device
        lib = _lib.load()
        g32 = g.detach().float().reshape(-1).contiguous()
        t32 = target.float().reshape(propagator.vol_shape).contiguous()
        dose = torch.empty(propagator.vol_shape, device=dev, dtype=torch.float32)
        operand = torch.empty(propagator.vol_shape, device=dev, dtype=torch.float32)
        loss = torch.empty(1, device=dev, dtype=torch.float64)
        nbytes = ctypes.c_int64()
        _lib.check(lib.vamx_rt_bclp_workspace_bytes(propagator.handle, ctypes.byref(nbytes)),
                   "vamx_rt_bclp_workspace_bytes")
        ws = torch.empty(int(nbytes.value), device=dev, dtype=torch.uint8)
        resp, tabs = response_model.vamx_struct(dev)
        prm, eps_t, w_t = _bclp_params(eps, weight, p, q, dev)
        with torch.cuda.device(dev):
            _lib.check(lib.vamx_rt_bclp_eval(propagator.handle, ctypes.byref(resp), ctypes.byref(prm),
                                             g32.data_ptr(), t32.data_ptr(),
                                             0 if eps_t is None else eps_t.data_ptr(),
                                             0 if w_t is None else w_t.data_ptr(), dose.data_ptr(),
                                             operand.data_ptr(), loss.data_ptr(), ws.data_ptr(),
                                             _lib.stream_ptr(dev)), "vamx_rt_bclp_eval")
        del tabs
        ctx.propagator = propagator
        ctx.p, ctx.q = float(p), float(q)
        ctx.g_shape = g.shape
        ctx.save_for_backward(operand, loss)
        ctx.dose = dose
        return loss[0]

    @staticmethod
    def backward(ctx, grad_out):
        operand, loss = ctx.saved_tensors
        scale = _loss_scale(loss[0], ctx.p, ctx.q)
        gray = ctx.propagator.forward(operand)
        grad = (grad_out.double() * scale).float() * gray
        return grad.reshape(ctx.g_shape), None, None, None, None, None, None, None


class BCLPLoss(torch.autograd.Function):
    """L(g) for the parallel-beam projector (fused forward, exact adjoint backward)."""

    @staticmethod
    def forward(ctx, g, projector, response_model, target, eps=0.1, weight=1.0, p=2.0, q=1.0):
        plan = projector.plan
        dev = g.device
        lib = _lib.load()
        g32 = g.detach().float().reshape(plan.sino_shape).contiguous()
        t32 = target.float().reshape(plan.vol_shape).contiguous()
        dose = torch.empty(plan.vol_shape, device=dev, dtype=torch.float32)
        operand = torch.empty(plan.vol_shape, device=dev, dtype=torch.float32)
        loss = torch.empty(1, device=dev, dtype=torch.float64)
        nbytes = ctypes.c_int64()
        _lib.check(lib.vamx_bclp_workspace_bytes(plan.handle, ctypes.byref(nbytes)), "vamx_bclp_workspace_bytes")
        ws = torch.empty(int(nbytes.value), device=dev, dtype=torch.uint8)
        resp, tabs = response_model.vamx_struct(dev)
        eps_t = None if np.isscalar(eps) else torch.as_tensor(eps, device=dev, dtype=torch.float32).contiguous()
        w_t = None if np.isscalar(weight) else torch.as_tensor(weight, device=dev, dtype=torch.float32).contiguous()
        prm = _lib.BclpParams(p=float(p), q=float(q), eps=float(eps) if eps_t is None else 0.0,
                              weight=float(weight) if w_t is None else 0.0, learning_rate=0.0,
                              glb=-np.inf, gub=np.inf, dvol=1.0, grad_op=1)
        _lib.check(lib.vamx_bclp_eval(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g32.data_ptr(),
                                      t32.data_ptr(), 0 if eps_t is None else eps_t.data_ptr(),
                                      0 if w_t is None else w_t.data_ptr(), dose.data_ptr(), operand.data_ptr(),
                                      loss.data_ptr(), ws.data_ptr(), _lib.stream_ptr(dev)), "vamx_bclp_eval")
        del tabs
        ctx.projector = projector
        ctx.p, ctx.q = float(p), float(q)
        ctx.g_shape = g.shape
        ctx.save_for_backward(operand, loss)
        ctx.dose = dose
        return loss[0]

    @staticmethod
    def backward(ctx, grad_out):
        operand, loss = ctx.saved_tensors
        L = loss[0]
        p, q = ctx.p, ctx.q
        scale = torch.where(L != 0, q * L.clamp_min(1e-300) ** ((q - p) / q), torch.ones_like(L))
        gsino = ctx.projector.plan.backward_adjoint(operand)
        grad = (grad_out.double() * scale).float() * gsino
        return grad.reshape(ctx.g_shape), None, None, None, None, None, None, None
