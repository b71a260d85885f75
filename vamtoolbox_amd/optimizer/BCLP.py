"""Band-constraint Lp-norm (BCLP) sinogram optimisation on the GPU.

Mirrors ``vamtoolbox/optimizer/BCLP.py``: ``LogPerf`` (14-135), ``BCLPNorm``
(138-590) and ``minimizeBCLP`` (593-641), with the same options, iteration
order (one back projection and one forward projection per iteration, the
loss logged at index k for the k-th iterate, a final evaluation after the
loop) and outputs ("packaged", "full", "components").

Three execution modes for the gradient-descent loop, all on GPU-resident
tensors:

* fused (default): one ``vamx_bclp_step`` (parallel beam) or
  ``vamx_rt_bclp_step`` (ray tracer) per iteration -- the back projection
  (deposit) fused with response, band and Lp weighting and the loss
  partials, a fixed-order loss reduction, and the forward projection
  (integrate) fused with the sinogram update -- optionally replayed from a
  captured HIP graph; the loss log stays on the device.  With ``bit_depth``
  the step's update is followed by imposeSinogramConstraints (eager, no
  graph); with a zero-dose body by zeroing the iterate's rays through it.
* unfused (``fused=False``): projector calls plus torch element-wise
  operations with the reference's host-side branches, the plain GPU loop.
* ``gradient="exact"``: the gradient is taken through the exact transpose of
  the back projector (the true gradient of the loss) instead of the forward
  projector the reference uses in its place (BCLP.py:372).
"""

from __future__ import annotations

import ctypes
import logging
import time
import warnings

import numpy as np
import torch

from .. import _lib, geometry
from ..projectorconstructor import projectorconstructor
from ..util import data


class LogPerf:
    """BCLP.py:14-92 (performance and norm logs, n_iter + 1 entries)."""

    def __init__(self, options):
        self.options = options
        self.curr_iter = 0
        n = options.n_iter + 1
        self.loss = np.full(n, np.nan)
        for k in ("l0", "l1", "l2", "lf", "l0_eps0", "l1_eps0", "l2_eps0", "lf_eps0", "ver"):
            setattr(self, k, np.full(n, np.nan))
        self.iter_times = np.zeros(n)
        self.t0 = np.nan
        self.is_converged = False

    def save(self, filename):
        np.savetxt(filename, np.column_stack([range(self.options.n_iter + 1), self.iter_times, self.loss, self.l0,
                                              self.l1, self.l2, self.lf, self.l0_eps0, self.l1_eps0, self.l2_eps0,
                                              self.lf_eps0]),
                   delimiter=",",
                   header="n,compute time,loss,L0 norm,L1 norm,L2 norm,Linf norm,L0 norm eps=0,L1 norm eps=0,"
                          "L2 norm eps=0,Linf norm eps=0,")

    def show(self, normalation_flags=(True, True, True, True)):
        """The loss and the other norms over the iterations (BCLP.py:94-139),
        three matplotlib panels; a norm is divided by its largest value when
        its flag is set."""
        import matplotlib.pyplot as plt
        fig, axs = plt.subplots(1, 3)
        fig.set_size_inches(12, 8)
        axs[0].plot(self.loss)
        axs[0].set(title="Weighted $L_{p}$ norm cost function", xlabel="Iteration", ylabel="Cost")
        names = [r"$L_{0}$", r"$L_{1}$", r"$L_{2}$", r"$L_{\infty}$"]
        for ax, keys, title in ((axs[1], ("l0", "l1", "l2", "lf"), r"Other $L_{p}$ norms with original tolerance"),
                                (axs[2], ("l0_eps0", "l1_eps0", "l2_eps0", "lf_eps0"),
                                 r"Other $L_{p}$ norms with zero tolerance")):
            for k, name, norm in zip(keys, names, normalation_flags):
                v = np.asarray(getattr(self, k), dtype=np.float64)
                top = np.nanmax(np.abs(v)) if np.any(np.isfinite(v)) else 0.0
                ax.plot(v / top if norm and top > 0 else v, label=name)
            ax.set(title=title, xlabel="Iteration", ylabel="Normalized norm value")
            ax.legend()
        plt.show()
        return fig, axs

    def startTiming(self):
        self.t0 = time.perf_counter()

    def recordIterTime(self):
        self.iter_times[self.curr_iter] = time.perf_counter() - self.t0


def _opt(options, name, default):
    return getattr(options, name, default) if getattr(options, name, None) is not None else default


class BCLPNorm:
    def __init__(self, target_geo, proj_geo, options):
        self.logger = logging.getLogger(__name__)
        self.target_geo = target_geo
        self.target_geo.array = np.atleast_3d(self.target_geo.array)
        self.proj_geo = proj_geo
        self.logs = LogPerf(options)
        self.logs.startTiming()
        self.response_model = options.response_model
        self.eps = options.eps
        self.weight = options.weight
        self.p = float(options.p)
        self.q = float(options.q)
        self.learning_rate = float(options.learning_rate)
        self.g0 = options.g0
        self.glb = options.blb
        self.gub = options.bub
        self.bit_depth = options.bit_depth
        self.dvol = 1
        self.exit_param = options.exit_param
        self.verbose = options.verbose
        self.gradient = _opt(options, "gradient", "reference")
        if self.gradient not in ("reference", "exact"):
            raise ValueError("gradient must be 'reference' or 'exact'")
        self.P = projectorconstructor(self.target_geo, self.proj_geo)
        self.device = self.P.device
        self.parallel = hasattr(self.P, "plan")
        if self.gradient == "exact" and not self.parallel:
            # the ray tracer's forward is already the exact transpose of its backward
            self.gradient = "reference"
        # fused iteration on the GPU: vamx_bclp_step (parallel beam) or
        # vamx_rt_bclp_step (ray tracer); a zero-dose constraint zeroes the
        # iterate's rays through the zero-dose body after each step (see
        # _zero_dose and _gradient_descent_fused)
        zero_dose = getattr(self.P, "imposeZeroDose", None) is not None and \
            getattr(self.proj_geo, "zero_dose_sino", None) is not None
        self.fused = bool(_opt(options, "fused", True))
        self.zero_dose = zero_dose
        self.hip_graph = bool(_opt(options, "hip_graph", True))
        self.log_metrics = bool(_opt(options, "log_metrics", not self.fused))
        dev = self.device
        self.T = torch.as_tensor(np.asarray(self.target_geo.array, dtype=np.float32), device=dev)
        if self.response_model.checkResponseTarget(self.target_geo.array) is False:
            warnings.warn("Target is either out of range of response function over non-negative real dose input, "
                          "or contains inf/nan.")
        self.eps_t = None if np.isscalar(self.eps) else torch.as_tensor(
            np.broadcast_to(self.eps, self.T.shape).astype(np.float32), device=dev).contiguous()
        self.w_t = None if np.isscalar(self.weight) else torch.as_tensor(
            np.broadcast_to(self.weight, self.T.shape).astype(np.float32), device=dev).contiguous()
        # initial sinogram (BCLP.py:244-271)
        if self.g0 is None:
            if callable(getattr(self.P, "inverseBackward", None)):
                self.g0 = self.P.inverseBackward(self.response_model.map_inv(self.T.double()).float(),
                                                 filter_option=options.filter)
            else:
                g0 = self.P.forward(self.response_model.map_inv(self.T.double()).float())
                g0 = data.filterSinogram(g0, options.filter)
                g0 = g0 * (np.pi / (2 * proj_geo.n_angles))
                self.g0 = g0
        g0 = self.g0 if isinstance(self.g0, torch.Tensor) else torch.as_tensor(np.asarray(self.g0), device=dev)
        self.g0_shape = tuple(g0.shape)
        self.g0 = self.imposeSinogramConstraints(g0.to(dev, torch.float32))
        self.dose = None
        self.mapped_dose = None
        self.loss = None

    # -- reference building blocks (BCLP.py:277-437) --------------------------
    def imposeSinogramConstraints(self, g):
        if self.bit_depth is not None:
            g = data.discretize(g, self.bit_depth, [0.0, torch.amax(g)])
        lo = -np.inf if self.glb is None else float(self.glb)
        hi = np.inf if self.gub is None else float(self.gub)
        return torch.clamp(g, min=lo, max=hi)

    def _cyl(self, g):
        return g.reshape(self.g0_shape)

    def _zero_dose(self, g):
        """g with the zero-dose rays set to 0: the reference's back projection
        zeroes them in g itself (Projector3DParallelCUDA.py:57-58), so its
        update and its result continue from the zeroed sinogram."""
        f = getattr(self.P, "imposeZeroDose", None)
        return g if f is None else f(self._cyl(g)).reshape(g.shape)

    def _eval_state(self, g):
        """dose, mapped dose, errors, v, loss and operand of sinogram g (updateVariables)."""
        dose = self.P.backward(self._cyl(g)).reshape(self.T.shape)
        dd = dose.double()
        mapped = self.response_model.map(dd)
        eT = mapped - self.T.double()
        eps = self.eps if self.eps_t is None else self.eps_t.double()
        w = self.weight if self.w_t is None else self.w_t.double()
        eb = torch.abs(eT) - eps
        v = eb > 0
        loss = (torch.sum(v * w * torch.abs(eb) ** self.p) * self.dvol) ** (self.q / self.p)
        operand = v * w * (torch.abs(eb) ** (self.p - 1)) * torch.sign(eT) * self.response_model.dmapdf(dd)
        return dose, mapped, eT, eb, v, loss, operand

    def _grad(self, operand, loss_value):
        op = operand.float()
        if self.gradient == "exact":
            G = self.P.backward_adjoint(op)
        else:
            G = self.P.forward(op)
        if loss_value != 0.0:
            G = (self.q * loss_value ** ((self.q - self.p) / self.q)) * G
        return G.reshape(-1) if G.ndim != len(self.g0_shape) else G

    def evaluateNormMetrics(self, k, eT, eb, v):
        """BCLP.py:396-437."""
        lg = self.logs
        ab = torch.abs(eb)
        lg.l0[k] = float(torch.sum(v).item()) * self.dvol
        lg.l0_eps0[k] = float(torch.sum(torch.abs(eT) > 0).item()) * self.dvol
        lg.l1[k] = float(torch.sum(v * ab).item()) * self.dvol
        lg.l1_eps0[k] = float(torch.sum(torch.abs(eT)).item()) * self.dvol
        lg.l2[k] = float(torch.sqrt(torch.sum(v * eb ** 2) * self.dvol).item())
        lg.l2_eps0[k] = float(torch.sqrt(torch.sum(eT ** 2) * self.dvol).item())
        lg.lf[k] = float(torch.amax(v * ab).item())
        lg.lf_eps0[k] = float(torch.amax((torch.abs(eT) > 0) * torch.abs(eT)).item())

    def checkConvergence(self, k):
        """BCLP.py:548-590."""
        loss = self.logs.loss
        if loss[k] == 0.0:
            return True
        n = 5
        if k >= n and self.exit_param is not None:
            change = np.abs(np.diff(loss[k - n:k + 1]))
            return bool(np.mean(change) <= self.exit_param * loss[k])
        return False

    # -- loops ----------------------------------------------------------------
    def gradientDescent(self):
        if self.fused:
            return self._gradient_descent_fused()
        return self._gradient_descent_unfused()

    def _gradient_descent_unfused(self):
        g = self.g0.reshape(-1).clone()
        n_iter = self.logs.options.n_iter
        for k in range(n_iter):
            dose, mapped, eT, eb, v, loss, operand = self._eval_state(g)
            g = self._zero_dose(g)
            L = float(loss.item())                                 # host branch of BCLP.py:370-375
            self.logs.loss[k] = L
            G = self._grad(operand, L)
            g = self.imposeSinogramConstraints(self._cyl(g - self.learning_rate * G.reshape(-1))).reshape(-1)
            if self.log_metrics:
                self.evaluateNormMetrics(k, eT, eb, v)
            self.logs.curr_iter = k
            self.logs.recordIterTime()
            self.logs.curr_iter = k + 1
            if self.verbose in ("time", "plot", "plot_demo"):
                self.logger.info(f"Iteration {k: 4.0f} at time: {self.logs.iter_times[k]: 6.1f} s. Loss = {L}")
            if self.exit_param is not None and self.checkConvergence(k):
                self.logger.warning("Optimization converged or loss value reached zero.")
                self.logs.is_converged = True
                break
        self._finish(g, self.logs.curr_iter)
        return self._zero_dose(g)

    def _vamx_args(self):
        resp, tabs = self.response_model.vamx_struct(self.device)
        self._keep = tabs
        prm = _lib.BclpParams(
            p=self.p, q=self.q, eps=float(self.eps) if self.eps_t is None else 0.0,
            weight=float(self.weight) if self.w_t is None else 0.0, learning_rate=self.learning_rate,
            glb=-np.inf if self.glb is None else float(self.glb), gub=np.inf if self.gub is None else float(self.gub),
            dvol=float(self.dvol), grad_op=1 if self.gradient == "exact" else 0)
        if self.bit_depth is not None:
            # the step's update stays unbounded; imposeSinogramConstraints
            # (discretize, then the bounds) follows it, as in the unfused loop
            prm.glb, prm.gub = -np.inf, np.inf
        return resp, prm

    def _gradient_descent_fused(self):
        lib = _lib.load()
        dev = self.device
        n_iter = self.logs.options.n_iter
        if self.parallel:
            plan = self.P.plan
            handle = plan.handle
            g = self.g0.reshape(plan.sino_shape).contiguous().clone()
            dose = torch.empty(plan.vol_shape, device=dev, dtype=torch.float32)
            ws_fn, step_name = lib.vamx_bclp_workspace_bytes, "vamx_bclp_step"
        else:
            handle = self.P.handle
            g = self.g0.reshape(-1).contiguous().clone()
            dose = torch.empty(self.P.vol_shape, device=dev, dtype=torch.float32)
            ws_fn, step_name = lib.vamx_rt_bclp_workspace_bytes, "vamx_rt_bclp_step"
        step_fn = getattr(lib, step_name)
        if self.T.numel() != dose.numel():
            raise ValueError("the target and the dose grid differ in size")
        T = self.T.contiguous()
        loss_log = torch.full((n_iter + 1,), float("nan"), device=dev, dtype=torch.float64)
        counter = torch.zeros(1, device=dev, dtype=torch.int32)
        nbytes = ctypes.c_int64()
        _lib.check(ws_fn(handle, ctypes.byref(nbytes)), "workspace bytes")
        ws = torch.empty(int(nbytes.value), device=dev, dtype=torch.uint8)
        resp, prm = self._vamx_args()
        eps_p = 0 if self.eps_t is None else self.eps_t.data_ptr()
        w_p = 0 if self.w_t is None else self.w_t.data_ptr()
        # zero-dose rays: the unfused loop back-projects, updates and returns
        # the sinogram zeroed on them (see _zero_dose); here g is zeroed on
        # them before the first step and after every step, in place (a fill
        # kernel, so the step stays graph-capturable), which gives every back
        # projection the same input and the same final g
        zd = self.P._zero_dose() if (self.parallel and self.zero_dose) else None
        if zd is not None:
            zd = zd.reshape(g.shape)
            g.masked_fill_(zd, 0.0)

        def step():
            _lib.check(step_fn(handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(), T.data_ptr(), eps_p, w_p,
                               dose.data_ptr(), loss_log.data_ptr(), counter.data_ptr(), ws.data_ptr(),
                               _lib.stream_ptr(dev)), step_name)
            if self.bit_depth is not None:
                g.copy_(self.imposeSinogramConstraints(g))
            if zd is not None:
                g.masked_fill_(zd, 0.0)

        per_iter_host = self.exit_param is not None or self.log_metrics or self.verbose in ("time", "plot")
        # (discretize's reductions stay out of graphs)
        use_graph = self.hip_graph and not per_iter_host and n_iter > 1 and self.bit_depth is None
        k = 0
        if use_graph:
            step()  # first iteration eagerly (also warms up the kernels)
            graph = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream(device=dev)
            s.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(s):
                with torch.cuda.graph(graph, stream=s):
                    step()
            torch.cuda.current_stream(dev).wait_stream(s)
            # the capture did not execute the step: replay n_iter - 1 times
            for _ in range(n_iter - 1):
                graph.replay()
            k = n_iter
        else:
            for k in range(n_iter):
                step()
                if per_iter_host:
                    L = float(loss_log[k].item())
                    self.logs.loss[k] = L
                    self.logs.curr_iter = k
                    self.logs.recordIterTime()
                    if self.log_metrics:
                        # metrics of iterate k need its dose, which the step just computed
                        _, _, eT, eb, v, _, _ = self._state_from_dose(dose)
                        self.evaluateNormMetrics(k, eT, eb, v)
                    if self.exit_param is not None and self.checkConvergence(k):
                        self.logs.is_converged = True
                        k += 1
                        break
            else:
                k = n_iter
        self.logs.curr_iter = k
        torch.cuda.synchronize(dev)
        self.logs.loss[:k] = loss_log[:k].cpu().numpy()
        self.logs.recordIterTime()
        self._finish(g.reshape(-1), k)
        return g.reshape(-1)

    def _state_from_dose(self, dose):
        dd = dose.double()
        mapped = self.response_model.map(dd)
        eT = mapped - self.T.double()
        eps = self.eps if self.eps_t is None else self.eps_t.double()
        eb = torch.abs(eT) - eps
        v = eb > 0
        return dose, mapped, eT, eb, v, None, None

    def _finish(self, g, k):
        """Final evaluation of the last iterate (BCLP.py:502-505)."""
        dose, mapped, eT, eb, v, loss, _ = self._eval_state(g)
        self.dose = dose
        self.mapped_dose = mapped
        self.loss = float(loss.item())
        self.logs.loss[k] = self.loss
        self.evaluateNormMetrics(k, eT, eb, v)


def minimizeBCLP(target_geo, proj_geo, options, output="packaged"):
    """BCLP.py:593-641."""
    bclp = BCLPNorm(target_geo, proj_geo, options)
    g_opt = bclp.gradientDescent()
    g_cyl = g_opt.reshape(bclp.g0_shape)
    g_np = g_cyl.double().cpu().numpy()
    dose_np = bclp.dose.double().cpu().numpy()
    if output == "packaged":
        return (geometry.Sinogram(g_np, proj_geo, options), geometry.Reconstruction(dose_np, proj_geo, options),
                bclp.logs.loss)
    if output == "full":
        return (geometry.Sinogram(g_np, proj_geo, options), geometry.Reconstruction(dose_np, proj_geo, options),
                geometry.Reconstruction(bclp.mapped_dose.cpu().numpy(), proj_geo, options), bclp.logs)
    if output == "components":
        return g_np, dose_np, bclp.mapped_dose.cpu().numpy(), bclp.logs
    raise ValueError("unknown output %r" % (output,))
