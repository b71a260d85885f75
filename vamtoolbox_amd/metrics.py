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
        g = getattr(target, "_gel_t", None)
        if g is None or g.device != recon.device:
            target._gel_t = torch.as_tensor(gel_inds, device=recon.device)
            target._void_t = torch.as_tensor(void_inds, device=recon.device)
        gt, vt = target._gel_t, target._void_t
        num = int(gel_inds.sum() + void_inds.sum())
        if not layerwise:
            min_gel = torch.amin(recon[gt])
            return float(torch.sum(recon[vt] >= min_gel).item()) / num
        out = np.zeros(recon.shape[2])
        for z in range(recon.shape[2]):
            sl = recon[:, :, z]
            min_gel = torch.amin(sl[gt[:, :, z] if gt.ndim == 3 else gt])
            out[z] = float(torch.sum(sl[vt[:, :, z] if vt.ndim == 3 else vt] >= min_gel).item()) / num
        return out
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
