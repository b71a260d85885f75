"""Propagation matrix from impulse responses (vamtoolbox/util/build_sparse_matrix.py:12-77).

Row r of P is the back projection of the unit sinogram element r, so
P x = B^T x for the projector that ``projectorconstructor`` builds.  One back
projection per row on the GPU; meant for small problems (the reference's
method), the ray tracer records its own matrix directly with
``PyTorchRayTracingPropagator.buildPropagationMatrix``.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse
import torch


def buildMatrix(target_geo, proj_geo, sparse_format="csr", save_path: str | None = None,
                block_insert: bool | None = False):
    from ..projectorconstructor import projectorconstructor
    if block_insert:
        raise Exception("Block insertsion is not yet implemented.")
    A = projectorconstructor(target_geo, proj_geo)
    dev = A.device
    vol = torch.zeros(tuple(np.shape(target_geo.array)), dtype=torch.float32, device=dev)
    sino_shape = tuple(A.forward(vol).shape)
    n_rows, n_cols = int(np.prod(sino_shape)), int(vol.numel())
    unit = torch.zeros(n_rows, dtype=torch.float32, device=dev)
    rows, cols, vals = [], [], []
    for row in range(n_rows):
        unit.zero_()
        unit[row] = 1.0
        resp = A.backward(unit.reshape(sino_shape)).reshape(-1)
        nz = torch.nonzero(resp).reshape(-1)
        rows.append(np.full(nz.numel(), row, dtype=np.int64))
        cols.append(nz.cpu().numpy())
        vals.append(resp[nz].double().cpu().numpy())
    P = scipy.sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                shape=(n_rows, n_cols)).asformat(sparse_format)
    if save_path is not None:
        scipy.sparse.save_npz(save_path, P)
    return P
