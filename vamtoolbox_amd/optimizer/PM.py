"""Projection optimisation by L-BFGS-B (vamtoolbox/optimizer/PM.py:25-229).

The quasi-Newton iteration stays on the host (scipy's L-BFGS-B, as in the
reference); every loss / gradient evaluation runs the projections on the GPU:

    dose = (pi/180) B b
    R1 = dose >= d_h, R2 = dose <= d_l, V1 = core & !R1, V2 = background & !R2
    loss = sum(rho_1 (dose - d_l)^p V2 + rho_2 (d_h - dose)^p V1)
    grad = (pi/180) A [rho_1 V2 p (dose - d_l)^(p-1) - rho_2 V1 p (d_h - dose)^(p-1)]

with the reference's gradient through the forward projector (PM.py:117-140)
and bounds 0 <= b <= 1.  ``np.product`` of the reference (removed in NumPy 2)
is ``np.prod`` here.
"""

from __future__ import annotations

import time

import numpy as np
import torch
from scipy import optimize

from .. import geometry, metrics
from ..projectorconstructor import projectorconstructor
from ..util import data


def minimizePM(target_geo, proj_geo, options):
    A = projectorconstructor(target_geo, proj_geo)
    dev = A.device
    factor = np.pi / 180
    d_h, d_l, p = float(options.d_h), float(options.d_l), float(options.p)
    rho_1, rho_2 = float(options.rho_1), float(options.rho_2)
    arr = np.asarray(target_geo.array)
    core = np.zeros(arr.shape, dtype=bool)
    core[target_geo.gel_inds] = True
    background = np.zeros(arr.shape, dtype=bool)
    background[target_geo.void_inds] = True
    core_t = torch.as_tensor(core, device=dev)
    back_t = torch.as_tensor(background, device=dev)
    n_iter = options.n_iter
    error = np.full(n_iter + 1, np.nan)
    iter_times = np.zeros(n_iter + 1)
    losses = []
    state = {"dose": None, "ver": np.nan, "it": 0}
    t0 = time.perf_counter()

    def regions(dose):
        V1 = core_t & ~(dose >= d_h)
        V2 = back_t & ~(dose <= d_l)
        return V1, V2

    def loss(b_flat):
        b = b_flat
        if options.bit_depth is not None:
            b = data.discretize(b, options.bit_depth, [0.0, np.max(b)])
        bt = torch.as_tensor(np.reshape(b, b0_shape), dtype=torch.float32, device=dev)
        dose = A.backward(bt).double() * factor                          # PM.py:85-86
        state["dose"] = dose
        state["ver"] = metrics.calcVER(target_geo, dose)
        V1, V2 = regions(dose)
        integrand = rho_1 * torch.pow(dose - d_l, p) * V2 + rho_2 * torch.pow(d_h - dose, p) * V1
        L = float(torch.sum(integrand).item())
        losses.append(L)
        return L

    def loss_gradient(b_flat):
        dose = state["dose"]                       # the dose of the last loss evaluation (PM.py:111-114)
        V1, V2 = regions(dose)
        g = rho_1 * V2.double() * p * torch.pow(dose - d_l, p - 1) - \
            rho_2 * V1.double() * p * torch.pow(d_h - dose, p - 1)
        grad = A.forward(g.float()).double().reshape(-1).cpu().numpy()
        return grad * factor

    def callback(xk):
        state["it"] += 1
        iter_times[state["it"]] = time.perf_counter() - t0
        error[state["it"]] = state["ver"]

    T = torch.as_tensor(np.asarray(arr, dtype=np.float32), device=dev)
    b0 = data.filterSinogram(A.forward(T).double(), options.filter)
    b0 = torch.clamp(b0, min=0.0)
    b0 = b0 / torch.amax(b0)
    b0_shape = tuple(b0.shape)
    x0 = A.backward(b0.float())
    x0 = x0 / torch.amax(x0)
    error[0] = metrics.calcVER(target_geo, x0)
    iter_times[0] = time.perf_counter() - t0
    b0 = b0.reshape(-1).cpu().numpy()
    bounds = np.tile(np.array([[0.0, 1.0]]), (b0.size, 1))
    lbfgs_options = {"maxiter": n_iter, "ftol": 1e-12, "disp": False, "maxcor": 10, "maxls": 30}
    t0 = time.perf_counter()
    result = optimize.minimize(fun=loss, x0=b0, callback=callback, method="L-BFGS-B", jac=loss_gradient,
                               bounds=bounds, options=lbfgs_options)
    b_opt = np.reshape(result.x, b0_shape)
    if options.bit_depth is not None:
        b_opt = data.discretize(b_opt, options.bit_depth, [0.0, np.max(b_opt)])
    x_opt = A.backward(torch.as_tensor(b_opt, dtype=torch.float32, device=dev)).double() * factor
    options.iter_times = iter_times
    options.loss = np.array(losses)
    return (geometry.Sinogram(b_opt, proj_geo, options),
            geometry.Reconstruction(x_opt.cpu().numpy(), proj_geo, options),
            error)
