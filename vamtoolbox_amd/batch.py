"""Many independent parallel-beam problems of uneven size, projected in one
persistent launch, and the OSMO iteration over all of them at once.

VAMToolbox solves one target at a time (vamtoolbox/optimize.py:174-217).  A
batch holds k problems -- each its own ``ParallelPlan`` with its own size,
angle set and absorption -- with all volumes in one flat float32 buffer and
all sinograms in another.  ``ParallelBatch.forward`` / ``backward`` project
every problem with ``vamx_batch_forward`` / ``vamx_batch_backward`` on a batch
handle (``vamx_batch_create``) that owns the batch's device tables, tile
counter and side streams: one launch whose blocks take work tiles from a
device-side queue, largest problem first, so small problems fill the gaps the
large ones leave.

``minimizeOSMOBatched`` runs the OSMO iteration of
vamtoolbox/optimizer/OSMO.py:9-178 for every problem (the same steps as
``vamtoolbox_amd.optimizer.OSMO.minimizeOSMO`` on each problem alone) with
the projections batched and the element-wise steps as single launches over
the flat buffers (``vamx_osmo_model_update`` for the model updates,
``vamx_scale_clamp_segments`` for every problem's max normalisation at
once); the iteration is captured once in a HIP graph and replayed.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib, geometry, metrics
from .parallel import ParallelPlan
from .util import data


class ParallelBatch:
    """k parallel-beam plans on one GPU sharing flat volume / sinogram buffers.

    Problem i's volume is ``vol_flat[vol_off[i] : vol_off[i] + vol_size[i]]``
    in its plan's (n, n, nz) layout, and likewise for the sinogram.
    """

    def __init__(self, plans):
        self.plans = list(plans)
        if not self.plans:
            raise ValueError("ParallelBatch needs at least one plan")
        self.device = self.plans[0].device
        for p in self.plans:
            if p.device != self.device:
                raise ValueError("all plans of a batch must live on one device")
        if len({id(p) for p in self.plans}) != len(self.plans):
            raise ValueError("a plan may appear only once in a batch (its slice min/max scratch serves one problem)")
        self.k = len(self.plans)
        self.lib = _lib.load()
        self.vol_size = [int(np.prod(p.vol_shape)) for p in self.plans]
        self.sino_size = [int(np.prod(p.sino_shape)) for p in self.plans]
        self.vol_off = np.concatenate([[0], np.cumsum(self.vol_size)[:-1]]).astype(np.int64)
        self.sino_off = np.concatenate([[0], np.cumsum(self.sino_size)[:-1]]).astype(np.int64)
        self.n_vol = int(sum(self.vol_size))
        self.n_sino = int(sum(self.sino_size))
        self._handles = (ctypes.c_void_p * self.k)(*[p.handle.value for p in self.plans])
        self._ptrs = {}
        # the batch's device tables, tile counter and side streams (released
        # in close(), before the plans)
        self.handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_batch_create(self._handles, self.k, ctypes.byref(self.handle)),
                       "vamx_batch_create")

    def close(self):
        h, self.handle = getattr(self, "handle", None), None
        if h and h.value:
            self.lib.vamx_batch_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers -----------------------------------------------------------
    def empty_vol(self) -> torch.Tensor:
        return torch.empty(self.n_vol, device=self.device, dtype=torch.float32)

    def empty_sino(self) -> torch.Tensor:
        return torch.empty(self.n_sino, device=self.device, dtype=torch.float32)

    def vol_views(self, flat: torch.Tensor):
        return [flat[o:o + s].view(p.vol_shape) for o, s, p in zip(self.vol_off, self.vol_size, self.plans)]

    def sino_views(self, flat: torch.Tensor):
        return [flat[o:o + s].view(p.sino_shape) for o, s, p in zip(self.sino_off, self.sino_size, self.plans)]

    def pack_vols(self, vols) -> torch.Tensor:
        flat = self.empty_vol()
        for v, dst in zip(vols, self.vol_views(flat)):
            dst.copy_(torch.as_tensor(np.asarray(v) if not isinstance(v, torch.Tensor) else v,
                                      dtype=torch.float32).reshape(dst.shape))
        return flat

    def _pointer_table(self, t: torch.Tensor, offs):
        key = (t.data_ptr(), id(offs))
        tab = self._ptrs.get(key)
        if tab is None:
            base = t.data_ptr()
            tab = (ctypes.c_void_p * self.k)(*[base + 4 * int(o) for o in offs])
            if len(self._ptrs) > 64:
                self._ptrs.clear()
            self._ptrs[key] = tab
        return tab

    def _check(self, t: torch.Tensor, n: int, what: str):
        if not isinstance(t, torch.Tensor) or t.device != self.device or t.dtype != torch.float32 or \
                not t.is_contiguous() or t.numel() != n:
            raise ValueError("%s must be a contiguous float32 tensor of %d elements on %s" % (what, n, self.device))

    # -- operators -----------------------------------------------------------
    def forward(self, vol_flat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Sinograms of every problem (one persistent launch; problems that fill
        the GPU alone get their own forward launch)."""
        self._check(vol_flat, self.n_vol, "vol_flat")
        out = self.empty_sino() if out is None else out
        self._check(out, self.n_sino, "out")
        _lib.check(self.lib.vamx_batch_forward(self.handle, self._pointer_table(vol_flat, self.vol_off),
                                               self._pointer_table(out, self.sino_off), _lib.stream_ptr(self.device)),
                   "vamx_batch_forward")
        return out

    def backward(self, sino_flat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Back projections of every problem (one persistent launch)."""
        self._check(sino_flat, self.n_sino, "sino_flat")
        out = self.empty_vol() if out is None else out
        self._check(out, self.n_vol, "out")
        _lib.check(self.lib.vamx_batch_backward(self.handle, self._pointer_table(sino_flat, self.sino_off),
                                                self._pointer_table(out, self.vol_off), _lib.stream_ptr(self.device)),
                   "vamx_batch_backward")
        return out


def _plan_for(target_geo, proj_geo, device):
    arr = np.asarray(target_geo.array)
    n = arr.shape[0]
    if arr.shape[1] != n:
        raise ValueError("the parallel-beam projector needs square slices (nY == nX)")
    nz = arr.shape[2] if arr.ndim == 3 else 1
    if getattr(proj_geo, "zero_dose_sino", None) is not None or \
            (proj_geo.absorption_coeff is not None and getattr(proj_geo, "absorption_mask", None) is not None):
        raise NotImplementedError("batched OSMO: zero-dose constraints and absorption masks are per-problem "
                                  "options of minimizeOSMO; use absorption_per_voxel for absorption")
    occ = None
    if getattr(proj_geo, "attenuation_field", None) is not None:
        af = np.asarray(proj_geo.attenuation_field)
        occ = np.broadcast_to(af.reshape(af.shape[0], af.shape[1], -1), (n, n, nz)) > 0
    return ParallelPlan(n, nz, np.asarray(proj_geo.angles, dtype=np.float64), det_offset=proj_geo.det_offset,
                        incl_deg=float(proj_geo.inclination_angle or 0.0), mu=proj_geo.absorption_per_voxel,
                        radius=proj_geo.absorption_radius, occlusion=occ, device=device)


class OSMOBatch:
    """State of the batched OSMO iteration (OSMO.py:9-178 for every problem).

    ``step()`` enqueues one iteration (void step then gel step, each one
    model update, one batched forward and one batched back projection with
    the per-problem max normalisations) on the current stream without host
    synchronisation; ``capture()`` records it in a HIP graph that
    ``graph.replay()`` runs again.
    """

    def __init__(self, target_geos, proj_geos, options, device=None, autotune=False):
        from ._device import require_gpu
        self.device = require_gpu(device)
        self.target_geos = list(target_geos)
        self.proj_geos = list(proj_geos)
        if len(self.target_geos) != len(self.proj_geos):
            raise ValueError("one projection geometry per target")
        self.options = options
        plans = [_plan_for(t, p, self.device) for t, p in zip(self.target_geos, self.proj_geos)]
        if autotune:
            # each problem's work order (tile shape, z values per lane,
            # pairing) timed on its own shape; cached per shape
            for p in plans:
                p.autotune(reps=1)
        self.batch = ParallelBatch(plans)
        B = self.batch
        dev = self.device
        self.d_h, self.d_l = float(options.d_h), float(options.d_l)
        self.min_proj_val = float(options.inhibition)
        if options.bit_depth is not None:
            raise NotImplementedError("batched OSMO: bit_depth discretisation is a per-problem option of minimizeOSMO")
        gel = torch.zeros(B.n_vol, device=dev, dtype=torch.bool)
        void = torch.zeros(B.n_vol, device=dev, dtype=torch.bool)
        for t, g_, v_ in zip(self.target_geos, B.vol_views(gel), B.vol_views(void)):
            g_.copy_(torch.as_tensor(np.asarray(t.gel_inds).reshape(g_.shape)))
            v_.copy_(torch.as_tensor(np.asarray(t.void_inds).reshape(v_.shape)))
        self.gel, self.void = gel, void
        self._lib = _lib.load()
        # segment tables of the flat buffers for the per-problem maxima
        self._vol_seg = torch.as_tensor(np.append(B.vol_off, B.n_vol), dtype=torch.int64, device=dev)
        self._sino_seg = torch.as_tensor(np.append(B.sino_off, B.n_sino), dtype=torch.int64, device=dev)
        self._seg_work = torch.empty(B.k, dtype=torch.int32, device=dev)
        # initialisation (OSMO.py:128-140), per problem
        self.x_model = B.empty_vol()
        for t, xm in zip(self.target_geos, B.vol_views(self.x_model)):
            T = torch.as_tensor(np.asarray(t.array, dtype=np.float64), device=dev)
            f = data.filterTargetOSMO(T, options.filter)
            xm.copy_((f.real if torch.is_complex(f) else f).reshape(xm.shape).to(torch.float32))
        self.b = B.forward(self.x_model)
        self.b.clamp_(min=0.0)
        self.x = B.backward(self.b)
        self._normalise_vol(self.x, float("-inf"))
        self.x_views = B.vol_views(self.x)
        self.b_views = B.sino_views(self.b)
        self.graph = None

    def _normalise(self, flat, seg, lo):
        """Every problem's v / max(v), then v[v < lo] = lo (OSMO.py:60-69)."""
        # the current stream at each call: a capture runs step() on its own
        st = _lib.stream_ptr(self.device)
        _lib.check(self._lib.vamx_scale_clamp_segments(flat.data_ptr(), flat.numel(), seg.data_ptr(), self.batch.k,
                                                       lo, self._seg_work.data_ptr(), st),
                   "vamx_scale_clamp_segments")

    def _normalise_vol(self, flat, lo):
        self._normalise(flat, self._vol_seg, lo)

    def _project(self):
        B = self.batch
        B.forward(self.x_model, out=self.b)
        # b / max(b), b < inhibition -> inhibition (OSMO.py:60-66)
        self._normalise(self.b, self._sino_seg, self.min_proj_val)
        B.backward(self.b, out=self.x)
        self._normalise_vol(self.x, 0.0)

    def _update_model(self, d, mask, is_gel):
        _lib.check(self._lib.vamx_osmo_model_update(self.x.data_ptr(), self.x_model.data_ptr(), mask.data_ptr(),
                                                    self.x.numel(), d, int(is_gel), _lib.stream_ptr(self.device)),
                   "vamx_osmo_model_update")

    def step(self):
        """One OSMO iteration for every problem (OSMO.py:36-109)."""
        self._update_model(self.d_l, self.void, False)    # x_model -= max(x - d_l, 0) on void
        self._project()
        self._update_model(self.d_h, self.gel, True)      # x_model += max(d_h - x, 0) on gel
        self._project()

    def capture(self):
        """Capture one iteration in a HIP graph (after one eager iteration)."""
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.step()
        torch.cuda.current_stream(self.device).wait_stream(s)
        self.graph = g
        return g

    def ver(self):
        return np.array([metrics.calcVER(t, x) for t, x in zip(self.target_geos, self.x_views)])

    def results(self):
        opts = self.options
        return [(geometry.Sinogram(b.double().cpu().numpy(), p, opts),
                 geometry.Reconstruction(x.double().cpu().numpy(), p, opts))
                for b, x, p in zip(self.b_views, self.x_views, self.proj_geos)]


def minimizeOSMOBatched(target_geos, proj_geos, options, hip_graph: bool = True, device=None, autotune=False):
    """OSMO on k independent problems at once.

    Returns a list of (Sinogram, Reconstruction, error) per problem, with the
    same meaning as ``minimizeOSMO``'s return value; ``error`` is the VER per
    iteration (iteration 0 = the initialisation).  With ``hip_graph`` the
    first iteration runs eagerly and the rest replay a captured graph.  With
    ``autotune`` every problem's plan first times the candidate work orders
    on its own shape (``ParallelPlan.autotune``).
    """
    st = OSMOBatch(target_geos, proj_geos, options, device=device, autotune=autotune)
    n_iter = options.n_iter
    err = np.full((st.batch.k, n_iter), np.nan)
    t0 = time.perf_counter()
    err[:, 0] = st.ver()
    iter_times = np.zeros(n_iter)
    for it in range(1, n_iter):
        if hip_graph and it >= 2:
            if st.graph is None:
                st.capture()
            st.graph.replay()
        else:
            st.step()
        err[:, it] = st.ver()
        iter_times[it] = time.perf_counter() - t0
    options.iter_times = iter_times
    return [(s, r, e) for (s, r), e in zip(st.results(), err)]
