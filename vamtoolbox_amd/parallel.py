"""Device-level parallel-beam projector plan (float32 torch tensors on the GPU).

Thin owner of a ``vamx_plan`` created by ``vamx_par_create``; every method
enqueues one HIP kernel sequence on torch's current stream through the C ABI.
Layouts follow the reference: volume (n, n, nz), sinogram (n, n_angles, nz),
z contiguous (vamtoolbox/geometry.py:223-234).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def outside_circle_mask(n: int) -> np.ndarray:
    """clipToCircle's mask (vamtoolbox/util/data.py:29-38): True outside the
    inscribed circle, computed with the reference's float64 formula."""
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n))
    return circle_x ** 2 + circle_y ** 2 > 1


class ParallelPlan:
    """Parallel-beam geometry bound to one GPU.

    Parameters mirror ``vamx_par_geom``: ``n`` (square slice size), ``nz``,
    ``angles_deg``, ``det_offset`` (voxels), ``incl_deg``, ``mu`` (absorption
    per voxel), ``radius`` (cylinder radius, default (n-1)/2), and the two
    clip switches (both on = reference behaviour).  ``per_lane_geometry``
    selects the kernels that evaluate the geometry in every lane instead of
    the wavefront-uniform ones (same results to float32 rounding).
    ``occlusion`` (n, n, nz) or (n, n, 1), non-zero where the attenuation field
    is > 0, turns on the reference's binary occlusion shadowing
    (Projector3DParallel.py:72-111, 147-153, 192-216).
    """

    def __init__(self, n: int, nz: int, angles_deg, det_offset: float = 0.0, incl_deg: float = 0.0,
                 mu: float = 0.0, radius: float | None = None, warp_clip: bool = True,
                 circle_clip: bool = True, per_lane_geometry: bool = False, occlusion=None, device=None):
        self.lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.n, self.nz = int(n), int(nz)
        self.angles_deg = np.ascontiguousarray(np.atleast_1d(np.asarray(angles_deg, dtype=np.float64)))
        self.na = int(self.angles_deg.size)
        self.det_offset, self.incl_deg, self.mu = float(det_offset), float(incl_deg), float(mu)
        self.radius = float(radius) if radius is not None else (self.n - 1) / 2.0
        mask = np.ascontiguousarray(outside_circle_mask(self.n).astype(np.uint8))
        flags = (0 if warp_clip else _lib.PAR_NO_WARP_CLIP) | (0 if circle_clip else _lib.PAR_NO_CIRCLE_CLIP) | \
            (_lib.PAR_PER_LANE_GEOMETRY if per_lane_geometry else 0)
        occ = None
        if occlusion is not None:
            occ = np.ascontiguousarray(np.broadcast_to(np.asarray(occlusion).reshape(
                (self.n, self.n, -1)), (self.n, self.n, self.nz)) > 0, dtype=np.uint8)
            if not occ.any():
                occ = None            # an empty insert shadows nothing
        self.occluded = occ is not None
        geom = _lib.ParGeom(
            n=self.n, nz=self.nz, n_angles=self.na,
            angles_deg=self.angles_deg.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            det_offset=self.det_offset, incl_deg=self.incl_deg, mu=self.mu, radius=self.radius,
            outside_mask=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), flags=flags,
            occlusion=None if occ is None else occ.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_par_create(ctypes.byref(geom), ctypes.byref(handle)), "vamx_par_create")
        self.handle = handle
        self.vol_shape = (self.n, self.n, self.nz)
        self.sino_shape = (self.n, self.na, self.nz)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.vamx_destroy(h)
            except Exception:
                pass
            self.handle = None

    # -- helpers ---------------------------------------------------------
    def _in(self, x: torch.Tensor, shape) -> torch.Tensor:
        if not isinstance(x, torch.Tensor):
            raise TypeError("expected a torch tensor")
        if x.device != self.device:
            raise ValueError("tensor on %s, plan on %s" % (x.device, self.device))
        if tuple(x.shape) != tuple(shape):
            if x.numel() == int(np.prod(shape)):
                x = x.reshape(shape)
            else:
                raise ValueError("expected shape %s, got %s" % (shape, tuple(x.shape)))
        if x.dtype != torch.float32:
            x = x.float()
        return x.contiguous()

    def _run(self, name, src, dst):
        with torch.cuda.device(self.device):
            _lib.check(getattr(self.lib, name)(self.handle, src.data_ptr(), dst.data_ptr(),
                                               _lib.stream_ptr(self.device)), name)
        return dst

    # -- operators -------------------------------------------------------
    def forward(self, vol: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        vol = self._in(vol, self.vol_shape)
        out = torch.empty(self.sino_shape, device=self.device, dtype=torch.float32) if out is None else out
        return self._run("vamx_par_forward", vol, out)

    def backward(self, sino: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        sino = self._in(sino, self.sino_shape)
        out = torch.empty(self.vol_shape, device=self.device, dtype=torch.float32) if out is None else out
        return self._run("vamx_par_backward", sino, out)

    def forward_adjoint(self, sino: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        sino = self._in(sino, self.sino_shape)
        out = torch.empty(self.vol_shape, device=self.device, dtype=torch.float32) if out is None else out
        return self._run("vamx_par_forward_adjoint", sino, out)

    def backward_adjoint(self, vol: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        vol = self._in(vol, self.vol_shape)
        out = torch.empty(self.sino_shape, device=self.device, dtype=torch.float32) if out is None else out
        return self._run("vamx_par_backward_adjoint", vol, out)

    # -- work-order tuning -----------------------------------------------------
    def get_tuning(self) -> dict:
        t = _lib.ParTuning()
        _lib.check(self.lib.vamx_par_get_tuning(self.handle, ctypes.byref(t)), "vamx_par_get_tuning")
        return {f: int(getattr(t, f)) for f, _ in _lib.ParTuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update({k: int(v) for k, v in kw.items()})
        t = _lib.ParTuning(**cur)
        _lib.check(self.lib.vamx_par_set_tuning(self.handle, ctypes.byref(t)), "vamx_par_set_tuning")

    # (angles -- or opposite-angle pairs -- per tile, rays per tile, z values
    # per lane, register tap reuse, opposite-angle pairing)
    FP_CANDIDATES = [(10, 4, 4, 1, 1), (10, 8, 4, 1, 1), (8, 8, 4, 1, 1), (12, 4, 4, 1, 1), (16, 8, 4, 1, 1),
                     (8, 16, 4, 1, 1), (16, 4, 4, 1, 1), (4, 16, 4, 1, 1),
                     (32, 8, 4, 1, 1), (8, 8, 2, 1, 1), (16, 16, 4, 0, 1), (16, 8, 4, 1, 0), (8, 16, 4, 1, 0),
                     (64, 8, 4, 1, 0), (32, 32, 4, 1, 0), (64, 16, 2, 1, 0), (32, 32, 4, 0, 0), (32, 32, 1, 0, 0)]
    # (tile edge in 2x2 voxel quads, z values per lane)
    BP_CANDIDATES = [(8, 4), (16, 4), (4, 4), (32, 4), (8, 2), (16, 2)]

    def autotune(self, vol: torch.Tensor | None = None, sino: torch.Tensor | None = None, reps: int = 3) -> dict:
        """Time the forward and back projection for each candidate work order
        (tile shape x z values per lane) on this plan's shape and keep the
        fastest of each; returns {setting: time_ms} plus the choice.  Results
        are cached per (n, nz, angles, absorption on/off, inclination on/off,
        detector offset) -- the angles and the offset decide the pairing."""
        key = (self.n, self.nz, self.na, self.mu != 0.0, self.incl_deg != 0.0, self.det_offset,
               hash(self.angles_deg.tobytes()))
        if key in _TUNING_CACHE:
            self.set_tuning(**_TUNING_CACHE[key]["choice"])
            return _TUNING_CACHE[key]
        dev = self.device
        if vol is None:
            vol = torch.rand(self.vol_shape, device=dev)
        if sino is None:
            sino = torch.rand(self.sino_shape, device=dev)
        out_s = torch.empty(self.sino_shape, device=dev)
        out_v = torch.empty(self.vol_shape, device=dev)

        def timed(fn):
            fn()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                fn()
            b.record()
            torch.cuda.synchronize(dev)
            return a.elapsed_time(b) / reps

        report = {"forward": {}, "backward": {}}
        for ta, tk, zv, ru, pr in self.FP_CANDIDATES:
            self.set_tuning(fp_tile_angles=ta, fp_tile_rays=tk, fp_zv=zv, fp_reuse=ru, fp_pair=pr)
            report["forward"][(ta, tk, zv, ru, pr)] = timed(lambda: self.forward(vol, out=out_s))
        for tile, zv in self.BP_CANDIDATES:
            self.set_tuning(bp_tile=tile, bp_zv=zv)
            report["backward"][(tile, zv)] = timed(lambda: self.backward(sino, out=out_v))
        fbest = min(report["forward"], key=report["forward"].get)
        bbest = min(report["backward"], key=report["backward"].get)
        choice = dict(fp_tile_angles=fbest[0], fp_tile_rays=fbest[1], fp_zv=fbest[2], fp_reuse=fbest[3],
                      fp_pair=fbest[4], bp_tile=bbest[0], bp_zv=bbest[1])
        self.set_tuning(**choice)
        report["choice"] = choice
        _TUNING_CACHE[key] = report
        return report


_TUNING_CACHE: dict = {}
