"""Optimisation metrics: ``calcVER`` (vamtoolbox/metrics.py:47-87)."""

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
            min_gel = torch.amin(torch.where(gt, recon, inf))
            return float(torch.count_nonzero(vt & (recon >= min_gel)).item()) / num
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
