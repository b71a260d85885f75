"""Parallel-beam projector on the GPU.

``Projector3DParallelHIP`` is the drop-in for the reference's parallel-beam
projector classes (``vamtoolbox/projector/Projector3DParallel.py:57-216``,
``Projector3DParallelCUDA.py:13-72``): ``forward(x) -> b`` and
``backward(b) -> x`` with the reference's layouts (volume (nY, nX, nZ),
sinogram (nR, nTheta, nZ)).  NumPy in -> NumPy out (float64, as the
reference returns); torch GPU tensor in -> float32 GPU tensor out (no host
copies).  The discretisation is the reference's Python projector (bilinear
rotate-and-sum forward, np.interp back projection, circle clip), evaluated
by the HIP kernels of libvamx_hip.so, with the reference's binary occlusion
shadowing when ``proj_geo.attenuation_field`` has positive entries.
Additionally ``forward_adjoint`` and ``backward_adjoint`` are the exact
transposes used for autograd (not available with occlusion).
"""

from __future__ import annotations

import numpy as np
import torch

from .._device import require_gpu
from ..parallel import ParallelPlan


class Projector3DParallelHIP:
    def __init__(self, target_geo, proj_geo, device=None):
        self.target_geo = target_geo
        self.proj_geo = proj_geo
        self.device = require_gpu(device)
        nY, nX = target_geo.array.shape[0], target_geo.array.shape[1]
        if nY != nX:
            raise ValueError("the parallel-beam projector needs square slices (nY == nX)")
        nZ = target_geo.array.shape[2] if np.ndim(target_geo.array) == 3 else 1
        self.angles = np.asarray(proj_geo.angles, dtype=np.float64)
        incl = proj_geo.inclination_angle or 0.0
        radius = proj_geo.absorption_radius
        # occlusion shadowing by the voxels of positive attenuation
        # (Projector3DParallel.py:72-75)
        occ = None
        if getattr(proj_geo, "attenuation_field", None) is not None:
            af = np.asarray(proj_geo.attenuation_field)
            occ = np.broadcast_to(af.reshape(af.shape[0], af.shape[1], -1), (nY, nX, nZ)) > 0
        self.plan = ParallelPlan(nY, nZ, self.angles, det_offset=proj_geo.det_offset, incl_deg=float(incl),
                                 mu=proj_geo.absorption_per_voxel, radius=radius, occlusion=occ,
                                 device=self.device)
        self.occluded = self.plan.occluded
        self.vol_shape = (nY, nX, nZ)
        self.sino_shape = (nY, self.angles.size, nZ)
        self._mask_t = None
        self._zd_t = None

    # -- helpers ----------------------------------------------------------
    def _to_dev(self, a, shape):
        if isinstance(a, torch.Tensor):
            t = a.to(self.device, torch.float32)
            return t.reshape(shape).contiguous(), True
        arr = np.asarray(a)
        return torch.as_tensor(np.ascontiguousarray(arr.reshape(shape), dtype=np.float32),
                               device=self.device), False

    def _out(self, t, torch_in, shape=None):
        if torch_in:
            return t
        out = t.double().cpu().numpy()
        return out if shape is None else out.reshape(shape)

    def _abs_mask(self):
        # reference ASTRA classes multiply by the radial absorption mask before
        # the forward and after the backward projection (Projector3DParallelCUDA.py:43-44, 66-67)
        if self.proj_geo.absorption_coeff is None or getattr(self.proj_geo, "absorption_mask", None) is None:
            return None
        if self._mask_t is None:
            m = np.broadcast_to(np.asarray(self.proj_geo.absorption_mask), self.vol_shape)
            self._mask_t = torch.as_tensor(np.ascontiguousarray(m, dtype=np.float32), device=self.device)
        return self._mask_t

    def _zero_dose(self):
        zd = self.proj_geo.zero_dose_sino
        if zd is None:
            return None
        if self._zd_t is None:
            self._zd_t = torch.as_tensor(np.array(np.broadcast_to(zd, self.sino_shape)), device=self.device)
        return self._zd_t

    def imposeZeroDose(self, b):
        """b with the rays through the zero-dose body set to 0 (a new tensor).
        The reference's back projection zeroes them in its input array, in
        place (Projector3DParallelCUDA.py:57-58, Projector3DParallel.py:43-44),
        so its optimisers carry on -- and return -- the zeroed sinogram; they
        call this after every back projection of their own sinogram."""
        zd = self._zero_dose()
        if zd is None:
            return b
        t = b if isinstance(b, torch.Tensor) else torch.as_tensor(np.asarray(b), device=self.device)
        return torch.where(zd.reshape(t.shape), torch.zeros((), device=t.device, dtype=t.dtype), t)

    # -- operators --------------------------------------------------------
    def forward(self, x):
        """Sinogram of ``x`` (Projector3DParallel.py:113-156)."""
        shape_in = np.shape(x)
        t, torch_in = self._to_dev(x, self.vol_shape)
        m = self._abs_mask()
        if m is not None:
            t = t * m
        b = self.plan.forward(t)
        if len(shape_in) == 2:
            return self._out(b, torch_in, (self.sino_shape[0], self.sino_shape[1]))
        return self._out(b, torch_in)

    def backward(self, b):
        """Reconstruction of ``b`` (Projector3DParallel.py:158-207); the input is not modified."""
        t, torch_in = self._to_dev(b, self.sino_shape)
        zd = self._zero_dose()
        if zd is not None:
            t = torch.where(zd, torch.zeros((), device=self.device), t)
        x = self.plan.backward(t)
        m = self._abs_mask()
        if m is not None:
            x = x * m
        return self._out(x, torch_in)

    def forward_adjoint(self, b):
        t, torch_in = self._to_dev(b, self.sino_shape)
        return self._out(self.plan.forward_adjoint(t), torch_in)

    def backward_adjoint(self, x):
        t, torch_in = self._to_dev(x, self.vol_shape)
        return self._out(self.plan.backward_adjoint(t), torch_in)
