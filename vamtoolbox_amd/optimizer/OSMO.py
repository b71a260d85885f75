"""Object-space model optimisation on the GPU (vamtoolbox/optimizer/OSMO.py:9-178).

The iteration of the reference -- a void step and a gel step, each one
forward and one backward projection plus max normalisations -- runs on
GPU-resident float32 tensors: the model update and each normalisation with
its bound are one library kernel each (vamx_osmo_model_update,
vamx_scale_clamp).  The VER of every iteration goes to a device log read at
the end, unless exit_param or verbose timing needs it on the host each
iteration; only the final arrays cross to the host otherwise.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .. import _lib, geometry, metrics
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
    gel = torch.as_tensor(target_geo.gel_inds, device=dev).contiguous()
    void = torch.as_tensor(target_geo.void_inds, device=dev).contiguous()
    bit_depth = options.bit_depth
    lib = _lib.load()
    st = _lib.stream_ptr(dev)

    def _native(*ts):
        return all(t.dtype == torch.float32 and t.is_cuda and t.is_contiguous() for t in ts)

    def normalise(v, lo):
        """v / max(v), then v[v < lo] = lo (one pass on float32 device arrays)."""
        m = torch.amax(v)
        if _native(v):
            _lib.check(lib.vamx_scale_clamp(v.data_ptr(), v.numel(), m.data_ptr(), lo, st), "vamx_scale_clamp")
            return v
        return torch.clamp(v / m, min=lo)

    def update_model(x, x_model, mask, d, is_gel):
        """OSMO.py:54-56 (void) and 92-94 (gel)."""
        if _native(x, x_model) and mask.shape == x_model.shape:
            _lib.check(lib.vamx_osmo_model_update(x.data_ptr(), x_model.data_ptr(), mask.data_ptr(), x_model.numel(),
                                                  float(d), int(is_gel), st), "vamx_osmo_model_update")
            return x_model
        if is_gel:
            return torch.where(mask, x_model + torch.clamp(d - x, min=0.0), x_model)
        return torch.where(mask, x_model - torch.clamp(x - d, min=0.0), x_model)

    def project(xm):
        b = A.forward(xm)
        if bit_depth is not None:
            b = b / torch.amax(b)
            b = data.discretize(b, bit_depth, [torch.amin(b), torch.amax(b)])
            # b < min_proj_val -> min_proj_val (OSMO.py:65, 103): one clamp,
            # which also keeps NaN as the comparison does
            return torch.clamp(b, min=min_proj_val)
        return normalise(b, min_proj_val)

    def reconstruct(b):
        x = A.backward(b).reshape(T.shape)      # the algebraic propagator returns flat arrays
        return normalise(x, 0.0)

    # the VER of every iteration is read once at the end unless a per-iteration
    # host decision needs it (exit_param, verbose timing); iteration times
    # then come from events
    per_iter_host = options.exit_param is not None or options.verbose in ("time", "plot")
    ver_log = None if per_iter_host else torch.zeros(options.n_iter, device=dev, dtype=torch.float64)
    events = [] if ver_log is not None else None

    def record(k, x):
        if ver_log is None:
            _error[k] = metrics.calcVER(target_geo, x)
            iter_times[k] = time.perf_counter() - t0
        else:
            metrics.calcVER_device(target_geo, x, out=ver_log[k])
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            events.append(ev)

    t0 = time.perf_counter()
    ev0 = None
    if events is not None:
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record()
    x_model = data.filterTargetOSMO(T, options.filter)
    x_model = (x_model.real if torch.is_complex(x_model) else x_model).to(torch.float32).contiguous()
    b = torch.clamp(A.forward(x_model), min=0.0)                 # OSMO.py:133-134
    x = A.backward(b).reshape(T.shape)
    b = _zero_dose(A, b)
    x = x / torch.amax(x)
    record(0, x)
    curr_iter = 0
    for curr_iter in range(1, options.n_iter):
        # void step (OSMO.py:36-71)
        x_model = update_model(x, x_model, void, options.d_l, False)
        b = project(x_model)
        x = reconstruct(b)
        b = _zero_dose(A, b)
        # gel step (OSMO.py:73-109)
        x_model = update_model(x, x_model, gel, options.d_h, True)
        b = project(x_model)
        x = reconstruct(b)
        b = _zero_dose(A, b)
        record(curr_iter, x)
        if options.verbose in ("time", "plot"):
            print("Iteration %4.0f at time: %6.1f s" % (curr_iter, iter_times[curr_iter]))
        if options.exit_param is not None and _error[curr_iter] <= options.exit_param:
            break
    if ver_log is not None:
        torch.cuda.synchronize(dev)
        n_rec = len(events)
        _error[:n_rec] = ver_log[:n_rec].cpu().numpy()
        iter_times[:n_rec] = [ev0.elapsed_time(ev) * 1e-3 for ev in events]
    b_opt = data.discretize(b, bit_depth, [0.0, torch.amax(b)]) if bit_depth is not None else b
    options.iter_times = iter_times
    return (geometry.Sinogram(b_opt.double().cpu().numpy(), proj_geo, options),
            geometry.Reconstruction(x.double().cpu().numpy(), proj_geo, options),
            _error)
