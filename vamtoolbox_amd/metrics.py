"""Optimisation metrics: ``calcVER`` (vamtoolbox/metrics.py:47-87)."""

from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def calcVER(target, recon, layerwise=False):
    """Volumetric error rate: #(void voxels >= min gel dose) / (#gel + #void)."""
    gel_inds, void_inds = target.gel_inds, target.void_inds
    if torch is not None and isinstance(recon, torch.Tensor):
        # masked reductions on the device (no compaction, one host sync for
        # the result); the gel / void masks and their count are cached on the
        # target
        g = getattr(target, "_gel_t", None)
        if g is None or g.device != recon.device:
            target._gel_t = torch.as_tensor(np.asarray(gel_inds), device=recon.device)
            target._void_t = torch.as_tensor(np.asarray(void_inds), device=recon.device)
            target._num_gv = int(np.sum(gel_inds) + np.sum(void_inds))
            target._n_gel = int(np.sum(gel_inds))
        gt, vt, num = target._gel_t, target._void_t, target._num_gv
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
