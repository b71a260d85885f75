"""Optimisation metrics (vamtoolbox/metrics.py): the volumetric error rate
``calcVER`` (47-87), the coefficient of variation ``calcCV`` (6-44), the
process window ``calcPW`` (90-127) and the in-part dose range ``calcIPDR``
(130-167), on NumPy arrays or device tensors.

``layerwise``: per z slice with the slice's gel / void voxels.  The
reference indexes each 2-D slice with the 3-D masks there, which NumPy
rejects for volumes, except in ``calcCV``, whose layerwise branch returns the
whole-volume value for every slice; ``calcCV`` keeps that."""

from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def device_masks(target, device):
    """The target's gel / void masks as device tensors and their voxel counts,
    cached on the target (built once per device)."""
    g = getattr(target, "_gel_t", None)
    if g is None or g.device != torch.device(device):
        target._gel_t = torch.as_tensor(np.asarray(target.gel_inds), device=device)
        target._void_t = torch.as_tensor(np.asarray(target.void_inds), device=device)
        target._n_gel = int(np.sum(target.gel_inds))
        target._n_void = int(np.sum(target.void_inds))
        target._num_gv = target._n_gel + target._n_void
    return target._gel_t, target._void_t, target._n_gel, target._n_void


def calcVER(target, recon, layerwise=False):
    """Volumetric error rate: #(void voxels >= min gel dose) / (#gel + #void)."""
    gel_inds, void_inds = target.gel_inds, target.void_inds
    if torch is not None and isinstance(recon, torch.Tensor):
        # masked reductions on the device (no compaction, one host sync for
        # the result); the gel / void masks and their count are cached on the
        # target
        gt, vt, _, _ = device_masks(target, recon.device)
        num = target._num_gv
        if target._n_gel == 0:
            raise ValueError("zero-size array to reduction operation minimum which has no identity")
        inf = float("inf")
        if not layerwise:
            return float(calcVER_device(target, recon).item())
        gt3 = gt if gt.ndim == 3 else gt[..., None]
        vt3 = vt if vt.ndim == 3 else vt[..., None]
        min_gel = torch.amin(torch.where(gt3, recon, inf), dim=(0, 1))          # per slice
        if bool(torch.isinf(min_gel).any()):
            raise ValueError("zero-size array to reduction operation minimum which has no identity")
        counts = torch.count_nonzero(vt3 & (recon >= min_gel), dim=(0, 1))
        return counts.double().cpu().numpy() / num
    num_gel_void = np.sum(gel_inds[:]) + np.sum(void_inds[:])
    if not layerwise:
        min_gel_dose = np.min(recon[gel_inds])
        return np.sum(recon[void_inds] >= min_gel_dose) / num_gel_void
    nX, nY, nZ = recon.shape
    out = np.zeros(nZ)
    for z in range(nZ):
        sl = recon[:, :, z]
        min_gel_dose = np.min(sl[gel_inds[:, :, z]])
        out[z] = np.sum(sl[void_inds[:, :, z]] >= min_gel_dose) / num_gel_void
    return out


_VER_WORK = {}


def calcVER_device(target, recon, out=None):
    """calcVER of a device tensor (not layerwise) as a 0-d float64 device
    tensor (written into ``out`` when given), without a host
    synchronisation: the same counts and division as calcVER, by the
    library's vamx_ver for a float32 dose."""
    gt, vt, _, _ = device_masks(target, recon.device)
    if target._n_gel == 0:
        raise ValueError("zero-size array to reduction operation minimum which has no identity")
    out = torch.empty((), dtype=torch.float64, device=recon.device) if out is None else out
    if recon.dtype == torch.float32 and recon.is_contiguous() and gt.numel() == recon.numel() and \
            gt.is_contiguous() and vt.is_contiguous():
        import ctypes
        from vamtoolbox_amd import _lib
        lib = _lib.load()
        ws = _VER_WORK.get(recon.device)
        if ws is None:
            nb = ctypes.c_int64()
            _lib.check(lib.vamx_ver_workspace_bytes(ctypes.byref(nb)), "vamx_ver_workspace_bytes")
            ws = _VER_WORK[recon.device] = torch.empty(int(nb.value), dtype=torch.uint8, device=recon.device)
        _lib.check(lib.vamx_ver(recon.data_ptr(), gt.data_ptr(), vt.data_ptr(), recon.numel(), target._num_gv,
                                out.data_ptr(), ws.data_ptr(), _lib.stream_ptr(recon.device)), "vamx_ver")
        return out
    min_gel = torch.amin(torch.where(gt, recon, float("inf")))
    out.copy_(torch.count_nonzero(vt & (recon >= min_gel)).double() / target._num_gv)
    return out


def _gel_void_values(target, recon):
    """(gel doses, void doses) as 1-D arrays or tensors."""
    if torch is not None and isinstance(recon, torch.Tensor):
        gt, vt, _, _ = device_masks(target, recon.device)
        return recon[gt.reshape(recon.shape)], recon[vt.reshape(recon.shape)]
    r = np.asarray(recon)
    return r[np.asarray(target.gel_inds).reshape(r.shape)], r[np.asarray(target.void_inds).reshape(r.shape)]


def _per_slice(target, recon, fn):
    nz = recon.shape[2]
    gel = np.asarray(target.gel_inds).reshape(recon.shape)
    void = np.asarray(target.void_inds).reshape(recon.shape)
    out = np.zeros(nz)
    for z in range(nz):
        sl = recon[:, :, z]
        if torch is not None and isinstance(recon, torch.Tensor):
            g = torch.as_tensor(gel[:, :, z], device=recon.device)
            v = torch.as_tensor(void[:, :, z], device=recon.device)
        else:
            g, v = gel[:, :, z], void[:, :, z]
        out[z] = fn(sl[g], sl[v])
    return out


def _f(x):
    return float(x.item()) if hasattr(x, "item") else float(x)


def calcCV(target, recon, layerwise=False):
    """Coefficient of variation of the gel doses: std / mean (metrics.py:6-44;
    population standard deviation, as np.std)."""
    g, _ = _gel_void_values(target, recon)
    if torch is not None and isinstance(g, torch.Tensor):
        gd = g.double()
        cv = _f(torch.sqrt(torch.mean((gd - torch.mean(gd)) ** 2)) / torch.mean(gd))
    else:
        cv = float(np.std(g) / np.mean(g))
    if not layerwise:
        return cv
    return np.full(recon.shape[2], cv)


def calcPW(target, recon, layerwise=False):
    """Process window: max void dose - min gel dose (metrics.py:90-127)."""
    fn = lambda g, v: _f(v.max()) - _f(g.min())   # noqa: E731
    if not layerwise:
        return fn(*_gel_void_values(target, recon))
    return _per_slice(target, recon, fn)


def calcIPDR(target, recon, layerwise=False):
    """In-part dose range: max gel dose - min gel dose (metrics.py:130-167)."""
    fn = lambda g, v: _f(g.max()) - _f(g.min())   # noqa: E731
    if not layerwise:
        return fn(*_gel_void_values(target, recon))
    return _per_slice(target, recon, fn)
