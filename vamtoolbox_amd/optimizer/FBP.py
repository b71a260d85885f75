"""Filtered back projection (vamtoolbox/optimizer/FBP.py:9-76) on the GPU.

No iteration: b = filter(A x_target), made non-negative (clip, or an offset
by the per-row magnitude of the minimum when ``options.offset``), normalised,
optionally discretised; x = B b normalised; error = VER of x.  The sinogram
and dose stay on the GPU until the returned arrays are built.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .. import geometry, metrics
from ..projectorconstructor import projectorconstructor
from ..util import data


def _zero_dose(A, b):
    """The sinogram as the reference's back projection leaves it: zero on the
    rays through the zero-dose body (parallel-beam projectors only)."""
    f = getattr(A, "imposeZeroDose", None)
    return b if f is None else f(b)


def minimizeFBP(target_geo, proj_geo, options):
    A = projectorconstructor(target_geo, proj_geo)
    dev = A.device
    _error = np.full(options.n_iter, np.nan)
    iter_times = np.zeros(options.n_iter)
    t0 = time.perf_counter()
    T = torch.as_tensor(np.asarray(target_geo.array, dtype=np.float32), device=dev)
    b = data.filterSinogram(A.forward(T).double(), options.filter)          # FBP.py:43-44
    if getattr(options, "offset", False):
        # offset by |min| over the detector axis (FBP.py:46-49)
        b = b + torch.abs(torch.amin(b, dim=0, keepdim=True)).expand_as(b)
    else:
        b = torch.clamp(b, min=0.0)                                           # FBP.py:51-53
    b = b / torch.amax(b)
    if options.bit_depth is not None:
        b = data.discretize(b, options.bit_depth, [0.0, torch.amax(b)])
    x = A.backward(b.float())
    b = _zero_dose(A, b)
    x = x / torch.amax(x)
    _error[0] = metrics.calcVER(target_geo, x)
    iter_times[0] = time.perf_counter() - t0
    options.iter_times = iter_times
    return (geometry.Sinogram(b.cpu().numpy(), proj_geo, options),
            geometry.Reconstruction(x.double().cpu().numpy(), proj_geo, options),
            _error)
