"""Curved-ray tracing propagator on the GPU.

``PyTorchRayTracingPropagator`` keeps the reference's class name, constructor
and methods (``vamtoolbox/projector/pyTorchRayTrace.py:15-166``):
``forward(f) -> g`` (flattened sinogram, C order (t, theta, z)),
``backward(g) -> dose`` and ``inverseBackward(f, filter_option)``.  Rays are
traced by the HIP kernel of libvamx_hip.so, one lane per ray for the whole
trajectory; trajectories are recomputed in ``backward`` instead of stored.
Inputs are never modified (the reference multiplies its ``forward`` input by
the absorption coefficient in place, 515-519).  ``ray_trace_ode_solver`` may
be "forward_symplectic_euler", "forward_euler", "leapfrog" or "rk4" (RK4 is a
placeholder in the reference, 289-292).
"""

from __future__ import annotations

import ctypes
import logging

import numpy as np
import torch

from .. import _lib
from .._device import require_gpu
from ..util.data import filterSinogram

_SOLVERS = {"forward_symplectic_euler": _lib.RT_SYMPLECTIC_EULER, "forward_euler": _lib.RT_FORWARD_EULER,
            "leapfrog": _lib.RT_LEAPFROG, "rk4": _lib.RT_RK4}
_PARAMS = {"canonical": _lib.RT_CANONICAL, "physical_path_length": _lib.RT_PHYSICAL,
           "optical_path_length": _lib.RT_OPTICAL}


class PyTorchRayTracingPropagator:
    def __init__(self, target_geo, proj_geo, output_torch_tensor=False, device=None):
        self.logger = logging.getLogger(__name__)
        self.target_geo = target_geo
        self.proj_geo = proj_geo
        self.output_torch_tensor = output_torch_tensor
        self.device = require_gpu(device)
        self.lib = _lib.load()
        im = proj_geo.index_model
        if im is None:
            raise Exception("ray tracing needs proj_geo.index_model")
        self.index_model = im
        self.attenuation_model = proj_geo.attenuation_model
        self.absorption_model = proj_geo.absorption_model
        method = proj_geo.ray_trace_method or "eikonal"
        if method not in ("eikonal", "hybrid"):
            raise Exception("ray_trace_method = %s is not supported (eikonal, hybrid)" % method)
        param = proj_geo.eikonal_parametrization
        if param not in _PARAMS:
            raise Exception(f"eikonal_parametrization = {param} and is not one of the available options.")
        solver = proj_geo.ray_trace_ode_solver
        if solver not in _SOLVERS:
            raise Exception(f"ode_solver = {solver} and is not one of the available options.")
        cfg = proj_geo.ray_trace_ray_config
        if cfg == "cone":
            raise Exception("Ray setup for cone beam is not yet implemented")
        if cfg != "parallel":
            raise Exception(f"Ray config: {str(cfg)} is not one of the supported string: 'parallel', 'cone'.")
        self.domain_n_dim = len(np.squeeze(target_geo.array).shape)
        self.tracing_step_size = float(im.voxel_size[0] / 2)           # pyTorchRayTrace.py:44-46
        self.max_num_step = int(target_geo.array.shape[0] * (2 * 1.5))  # 47-49
        self.num_step_per_exit_check = 10
        rd = proj_geo.ray_density if proj_geo.ray_density is not None else 1
        nx, ny, nz = im.shape
        nt = round(max(nx, ny) * rd)
        nzr = nz if nz == 1 else max(round(nz * rd), 1)
        self.sino_shape = (nt, int(np.size(proj_geo.angles)), nzr)
        self.num_rays = int(np.prod(self.sino_shape))
        self.ray_energy_0 = rd ** (-1.0) if nz == 1 else rd ** (-2.0)
        ikind, iR, ip, insur = im.kernel_index_spec()
        self._field = None if ikind != _lib.RT_INDEX_GRID else im.device_field(self.device)
        att = self.attenuation_model
        if att is None:
            akind, aalpha, arad, agrid = _lib.RT_ATT_NONE, 0.0, 0.0, None
        else:
            akind, aalpha, arad, agrid = att.kernel_spec(self.device)
        self._att_grid = agrid
        ab = self.absorption_model
        self.alpha_abs = None if ab is None else ab.alpha_at_grid()
        self._abs = None if ab is None else torch.as_tensor(
            np.ascontiguousarray(self.alpha_abs, dtype=np.float32), device=self.device)
        self._angles = np.ascontiguousarray(np.asarray(proj_geo.angles, dtype=np.float64))
        incl = proj_geo.inclination_angle or 0.0
        g = _lib.RtGeom(
            nx=nx, ny=ny, nz=nz,
            xmin=float(im.xv.min()), ymin=float(im.yv.min()), zmin=float(im.zv.min()),
            dx=float(im.voxel_size[0]), dy=float(im.voxel_size[1]),
            dz=float(im.voxel_size[2]) if nz > 1 else 1.0,
            nt=nt, n_angles=self.sino_shape[1], nzr=nzr,
            angles_deg=self._angles.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            incl_deg=float(incl), step=self.tracing_step_size, max_steps=self.max_num_step,
            exit_every=self.num_step_per_exit_check, solver=_SOLVERS[solver], param=_PARAMS[param],
            ray_energy0=float(self.ray_energy_0), n_grad=0 if self._field is None else self._field.data_ptr(),
            index_kind=ikind, index_R=iR, index_p=ip, index_n_sur=insur, att_kind=akind,
            att_alpha=aalpha, att_radius=arad, att_grid=0 if agrid is None else agrid.data_ptr(),
            abs_grid=0 if self._abs is None else self._abs.data_ptr())
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_rt_create(ctypes.byref(g), ctypes.byref(h)), "vamx_rt_create")
        self.handle = h
        self.vol_shape = (nx, ny, nz)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.vamx_destroy(h)
            except Exception:
                pass
            self.handle = None

    def _vol_in(self, f):
        if isinstance(f, torch.Tensor):
            return f.to(self.device, torch.float32).reshape(self.vol_shape).contiguous(), True
        return torch.as_tensor(np.ascontiguousarray(np.asarray(f).reshape(self.vol_shape), dtype=np.float32),
                               device=self.device), False

    def _sino_in(self, g):
        if isinstance(g, torch.Tensor):
            return g.to(self.device, torch.float32).reshape(-1).contiguous(), True
        return torch.as_tensor(np.ascontiguousarray(np.asarray(g).ravel(), dtype=np.float32),
                               device=self.device), False

    def forward(self, f):
        """g = W (alpha_abs f) (pyTorchRayTrace.py:87-107)."""
        t, torch_in = self._vol_in(f)
        out = torch.empty(self.num_rays, device=self.device, dtype=torch.float32)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_rt_integrate(self.handle, t.data_ptr(), out.data_ptr(),
                                                  _lib.stream_ptr(self.device)), "vamx_rt_integrate")
        if self.output_torch_tensor or torch_in:
            return out
        return out.cpu().numpy()

    def backward(self, ray_energy_0):
        """dose = alpha_abs (W^T g) (pyTorchRayTrace.py:109-133)."""
        t, torch_in = self._sino_in(ray_energy_0)
        out = torch.empty(self.vol_shape, device=self.device, dtype=torch.float32)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_rt_deposit(self.handle, t.data_ptr(), out.data_ptr(),
                                                _lib.stream_ptr(self.device)), "vamx_rt_deposit")
        if self.output_torch_tensor or torch_in:
            return out
        return np.squeeze(out.cpu().numpy())

    def stats(self):
        """(total ray-steps, rays still active at max_num_step) of the last trace."""
        buf = (ctypes.c_int64 * 2)()
        _lib.check(self.lib.vamx_rt_stats(self.handle, buf), "vamx_rt_stats")
        if buf[1]:
            self.logger.error(
                f"Some rays are still active when max_num_step ({self.max_num_step}) is reached. "
                f"Rays active: {buf[1]}/{self.num_rays}.")
        return int(buf[0]), int(buf[1])

    def buildPropagationMatrix(self):
        """Propagation matrix P (rays x voxels) with P f = forward(f) and
        P^T g = backward(g) (pyTorchRayTrace.py:168-197, recordEnergyUntilExit
        573-690): every interaction tap of every ray is recorded on the GPU
        (two passes over the same traces: counts, then taps), equal
        (ray, voxel) pairs are summed on the host in float64, and column v is
        scaled by the absorption coefficient at voxel v.  Returned as a torch
        sparse COO tensor when ``output_torch_tensor`` is set, else as a
        ``scipy.sparse.coo_array``."""
        import scipy.sparse
        dev = self.device
        n = self.num_rays
        counts = torch.empty(n, dtype=torch.int64, device=dev)
        with torch.cuda.device(dev):
            _lib.check(self.lib.vamx_rt_record_counts(self.handle, counts.data_ptr(), _lib.stream_ptr(dev)),
                       "vamx_rt_record_counts")
            offsets = (torch.cumsum(counts, 0) - counts).contiguous()
            total = int(counts.sum().item())
            vox = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
            w = torch.empty(max(total, 1), dtype=torch.float32, device=dev)
            _lib.check(self.lib.vamx_rt_record(self.handle, offsets.data_ptr(), vox.data_ptr(), w.data_ptr(),
                                               _lib.stream_ptr(dev)), "vamx_rt_record")
        self.stats()
        rows = np.repeat(np.arange(n, dtype=np.int64), counts.cpu().numpy())
        nvox = int(np.prod(self.vol_shape))
        P = scipy.sparse.coo_matrix((w[:total].cpu().numpy().astype(np.float64), (rows, vox[:total].cpu().numpy())),
                                    shape=(n, nvox)).tocsr()            # sums equal (ray, voxel) pairs
        if self.alpha_abs is not None:
            P = P @ scipy.sparse.diags(np.asarray(self.alpha_abs, dtype=np.float64).ravel())
        P = scipy.sparse.coo_array(P)
        if self.output_torch_tensor:
            idx = torch.as_tensor(np.vstack((P.row, P.col)).astype(np.int64), device=dev)
            return torch.sparse_coo_tensor(idx, torch.as_tensor(P.data, dtype=torch.float32, device=dev),
                                           size=P.shape).coalesce()
        return P

    def inverseBackward(self, f, filter_option="ram-lak", **kwargs):
        """Initial sinogram: f / alpha_abs^2, forward, ramp filter, * pi/(2 n_angles) (135-166)."""
        t, torch_in = self._vol_in(f)
        t = t.double()
        if self._abs is not None:
            t = t / self._abs.double() ** 2.0
        t = torch.where(torch.isnan(t) | torch.isinf(t), torch.zeros((), dtype=t.dtype, device=t.device), t)
        g0 = self.forward(t.float()).reshape(self.sino_shape)
        g0 = filterSinogram(g0, filter_option).reshape(-1)
        g0 = g0 * (np.pi / (2 * self.proj_geo.n_angles))
        if self.output_torch_tensor or torch_in:
            return g0.float()
        return g0.cpu().numpy()

    def reshape(self, ray_quantity):
        return ray_quantity.reshape(self.sino_shape)
