"""Projector dispatch (vamtoolbox/projectorconstructor.py:21-165).

Same signature and dispatch keys as the reference:
``projectorconstructor(target_geo, proj_geo, optical_params=None)`` returns
an object with ``forward(x)`` and ``backward(b)``.  Every propagator here
runs on the GPU through libvamx_hip.so:

* ``ray_type == "ray_trace"``  -> ``PyTorchRayTracingPropagator`` (HIP ray tracer)
* ``ray_type == "parallel"``   -> ``Projector3DParallelHIP`` for both ``CUDA``
  settings (the reference's Python-projector discretisation; ASTRA's
  ``line``/``parallel3d`` kernels are not reproduced)
* ``ray_type == "algebraic"``  -> ``PyTorchAlgebraicPropagator`` (CSR matrix
  from ``proj_geo.loading_path_for_matrix``, HIP CSR kernel)

As in the reference, an insert turns into infinite attenuation (62-70) and a
zero-dose body into ``proj_geo.zero_dose_sino`` (162-163).  With an
``attenuation_field`` the parallel projector applies the reference's binary
occlusion shadowing (voxels with attenuation > 0 occlude,
Projector3DParallel.py:72-111, 147-153, 192-216).
"""

from __future__ import annotations

from typing import Protocol

import numpy as np

from . import geometry


class CALopticalparams:
    def __init__(self, *args, **kwargs):
        raise NotImplementedError


class Projector(Protocol):
    def forward(self, x): ...

    def backward(self, b): ...


def projectorconstructor(target_geo: geometry.TargetGeometry, proj_geo: geometry.ProjectionGeometry,
                         optical_params=None) -> Projector:
    assert isinstance(target_geo, geometry.TargetGeometry), "target_geo should be of type: geometry.TargetGeometry"
    assert isinstance(proj_geo, geometry.ProjectionGeometry), \
        "proj_geo should be of type: geometry.ProjectionGeometry"
    if target_geo.insert is not None:
        if proj_geo.attenuation_field is not None:
            af = np.array(proj_geo.attenuation_field, dtype=np.float64, copy=True)
            af[np.where(target_geo.insert == 1, True, False)] = np.inf
            proj_geo.attenuation_field = af
        else:
            proj_geo.attenuation_field = np.where(target_geo.insert == 1, np.inf, 0)

    if proj_geo.ray_type == "algebraic":
        from .projector.pyTorchAlgebraicPropagation import PyTorchAlgebraicPropagator
        A = PyTorchAlgebraicPropagator(target_geo, proj_geo)
    elif proj_geo.ray_type == "ray_trace":
        from .projector.pyTorchRayTrace import PyTorchRayTracingPropagator
        A = PyTorchRayTracingPropagator(target_geo, proj_geo)
    elif proj_geo.ray_type == "parallel":
        from .projector.Projector3DParallel import Projector3DParallelHIP
        A = Projector3DParallelHIP(target_geo, proj_geo)
    else:
        raise NotImplementedError("ray_type %r is not supported" % (proj_geo.ray_type,))

    if target_geo.zero_dose is not None:
        proj_geo.calcZeroDoseSinogram(A, target_geo)
    return A
