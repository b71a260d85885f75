"""Object-space model optimisation on the GPU (vamtoolbox/optimizer/OSMO.py:9-178).

The iteration of the reference -- a void step and a gel step, each one
forward and one backward projection plus max normalisations -- runs on
GPU-resident float32 tensors; only the per-iteration VER and the final
arrays cross to the host.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .. import geometry, metrics
from ..projectorconstructor import projectorconstructor
from ..util import data
from .FBP import _zero_dose


def minimizeOSMO(target_geo, proj_geo, options):
    A = projectorconstructor(target_geo, proj_geo)
    dev = A.device
    _error = np.full(options.n_iter, np.nan)
    iter_times = np.zeros(options.n_iter)
    min_proj_val = float(options.inhibition)
    T = torch.as_tensor(np.asarray(target_geo.array, dtype=np.float64), device=dev)
    gel = torch.as_tensor(target_geo.gel_inds, device=dev)
    void = torch.as_tensor(target_geo.void_inds, device=dev)
    bit_depth = options.bit_depth

    def project(xm):
        b = A.forward(xm)
        b = b / torch.amax(b)
        if bit_depth is not None:
            b = data.discretize(b, bit_depth, [torch.amin(b), torch.amax(b)])
        # b < min_proj_val -> min_proj_val (OSMO.py:65, 103): one clamp,
        # which also keeps NaN as the comparison does
        return torch.clamp(b, min=min_proj_val)

    def reconstruct(b):
        x = A.backward(b).reshape(T.shape)      # the algebraic propagator returns flat arrays
        x = x / torch.amax(x)
        return torch.clamp(x, min=0.0)

    t0 = time.perf_counter()
    x_model = data.filterTargetOSMO(T, options.filter)
    x_model = (x_model.real if torch.is_complex(x_model) else x_model).to(torch.float32)
    b = torch.clamp(A.forward(x_model), min=0.0)                 # OSMO.py:133-134
    x = A.backward(b).reshape(T.shape)
    b = _zero_dose(A, b)
    x = x / torch.amax(x)
    _error[0] = metrics.calcVER(target_geo, x)
    iter_times[0] = time.perf_counter() - t0
    curr_iter = 0
    for curr_iter in range(1, options.n_iter):
        # void step (OSMO.py:36-71)
        void_diff = torch.clamp(x - options.d_l, min=0.0)
        x_model = torch.where(void, x_model - void_diff, x_model)
        b = project(x_model)
        x = reconstruct(b)
        b = _zero_dose(A, b)
        # gel step (OSMO.py:73-109)
        gel_diff = torch.clamp(options.d_h - x, min=0.0)
        x_model = torch.where(gel, x_model + gel_diff, x_model)
        b = project(x_model)
        x = reconstruct(b)
        b = _zero_dose(A, b)
        _error[curr_iter] = metrics.calcVER(target_geo, x)
        iter_times[curr_iter] = time.perf_counter() - t0
        if options.verbose in ("time", "plot"):
            print("Iteration %4.0f at time: %6.1f s" % (curr_iter, iter_times[curr_iter]))
        if options.exit_param is not None and _error[curr_iter] <= options.exit_param:
            break
    b_opt = data.discretize(b, bit_depth, [0.0, torch.amax(b)]) if bit_depth is not None else b
    options.iter_times = iter_times
    return (geometry.Sinogram(b_opt.double().cpu().numpy(), proj_geo, options),
            geometry.Reconstruction(x.double().cpu().numpy(), proj_geo, options),
            _error)
