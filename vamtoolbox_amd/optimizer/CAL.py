"""Computed axial lithography gradient descent on the GPU (vamtoolbox/optimizer/CAL.py:10-218)."""

from __future__ import annotations

import time

import numpy as np
import torch

from .. import geometry, metrics
from ..projectorconstructor import projectorconstructor
from ..util import data
from .FBP import _zero_dose


def _threshold_reconstruction(x, target_geo, sigmoid_strength):
    """CAL.py:39-110: the threshold (of 100 from min to max) maximising the
    gel fraction above it minus the void fraction above it, then a sigmoid
    around it.  The thresholds and scores are the reference's float64 NumPy
    arithmetic on the host; the counts #(x > t) for all thresholds come from
    the sorted gel and void values on the device."""
    gel, void, n_gel, n_void = metrics.device_masks(target_geo, x.device)
    xd = x.double()
    ths = np.linspace(float(torch.amin(xd)), float(torch.amax(xd)), 100)
    t = torch.as_tensor(ths, device=x.device)
    c_gel = (n_gel - torch.searchsorted(torch.sort(xd[gel]).values, t, right=True)).cpu().numpy()
    c_void = (n_void - torch.searchsorted(torch.sort(xd[void]).values, t, right=True)).cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        score = c_gel / n_gel - c_void / n_void
    th = float(ths[int(np.argmax(score))])
    return 1.0 / (1.0 + torch.exp(-(float(sigmoid_strength) * (xd - th)))), th


def minimizeCAL(target_geo, proj_geo, options):
    A = projectorconstructor(target_geo, proj_geo)
    dev = A.device
    _error = np.full(options.n_iter, np.nan)
    iter_times = np.zeros(options.n_iter)
    T = torch.as_tensor(np.asarray(target_geo.array, dtype=np.float32), device=dev)
    t0 = time.perf_counter()
    b = A.forward(T)                                             # CAL.py:127-129
    b = data.filterSinogram(b, options.filter).to(torch.float32)
    b = torch.clamp(b, min=0.0)
    b0 = b
    delta_b_prev = torch.zeros_like(b)
    x0 = A.backward(b0)
    b0 = b = _zero_dose(A, b0)
    x0 = x0 / torch.amax(x0)
    x = x0
    _error[0] = metrics.calcVER(target_geo, x0)
    iter_times[0] = time.perf_counter() - t0
    b_opt = b
    m = float(options.momentum)
    curr_iter = 0
    for curr_iter in range(1, options.n_iter):
        if options.d_h is not None:
            x_thresh = data.sigmoid(x - options.d_h, options.sigmoid)
        else:
            x_thresh, _ = _threshold_reconstruction(x, target_geo, options.sigmoid)
            x_thresh = x_thresh.float()
        delta_x = x_thresh - T
        delta_b = A.forward(delta_x)
        grad = delta_b * (1 - m) + delta_b_prev * m / (1 - m ** (curr_iter + 1))
        b = torch.clamp(b - options.learning_rate * grad, min=0.0)
        if options.bit_depth is not None:
            b = data.discretize(b, options.bit_depth, [0.0, torch.amax(b)])
        x = A.backward(b)
        b = _zero_dose(A, b)
        x = x / torch.amax(x)
        delta_b_prev = delta_b
        b_opt = b
        _error[curr_iter] = metrics.calcVER(target_geo, x)
        iter_times[curr_iter] = time.perf_counter() - t0
        if options.verbose in ("time", "plot"):
            print("Iteration %4.0f at time: %6.1f s" % (curr_iter, iter_times[curr_iter]))
        if options.exit_param is not None and _error[curr_iter] <= options.exit_param:
            break
    b_opt = torch.clamp(b_opt, min=0.0)                          # CAL.py:375-378
    b_opt = b_opt / torch.amax(b_opt)
    x_opt = A.backward(b_opt)
    b_opt = _zero_dose(A, b_opt)
    x_opt = x_opt / torch.amax(x_opt)
    options.iter_times = iter_times
    return (geometry.Sinogram(b_opt.double().cpu().numpy(), proj_geo, options),
            geometry.Reconstruction(x_opt.double().cpu().numpy(), proj_geo, options),
            _error)
