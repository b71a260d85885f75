"""Algebraic (sparse-matrix) propagation on the GPU.

``PyTorchAlgebraicPropagator`` keeps the reference's class name, constructor
and methods (``vamtoolbox/projector/pyTorchAlgebraicPropagation.py:13-182``):
the propagation matrix P is read from ``proj_geo.loading_path_for_matrix``
(``.npz``: a SciPy sparse matrix; ``.pt``: a torch sparse tensor, loaded with
``weights_only=True``), kept on the GPU in CSR form together with its explicit
CSR transpose, and applied with z-tiling: P encodes one z layer, a volume of
``P.shape[1] * z_tiling`` voxels is the (n_cols, z_tiling) matrix X (z
fastest) and the sinogram is P X, again z fastest.  ``forward`` and
``backward`` run the CSR kernel of libvamx_hip.so (``vamx_csr_spmm``).
"""

from __future__ import annotations

import logging
import os

import numpy as np
import scipy.sparse
import torch

from .. import _lib
from .._device import require_gpu


class _DeviceCSR:
    """CSR matrix on the GPU: int64 row pointers, int32 columns, float32 values."""

    def __init__(self, m: scipy.sparse.csr_matrix, device):
        m = scipy.sparse.csr_matrix(m)
        m.sum_duplicates()
        m.sort_indices()
        if m.shape[1] >= 2 ** 31:
            raise ValueError("sparse matrix has too many columns for int32 indices")
        self.shape = m.shape
        self.row_ptr = torch.as_tensor(m.indptr.astype(np.int64), device=device)
        self.col = torch.as_tensor(m.indices.astype(np.int32), device=device)
        self.val = torch.as_tensor(m.data.astype(np.float32), device=device)


def _load_matrix(path) -> scipy.sparse.csr_matrix:
    ext = os.path.splitext(str(path))[-1].lower()
    if ext == ".npz":
        try:
            return scipy.sparse.csr_matrix(scipy.sparse.load_npz(path))
        except OSError:
            raise Exception(f'Sparse matrix file at "{path}" does not exist or cannot be read.')
    if ext == ".pt":
        try:
            t = torch.load(path, map_location="cpu", weights_only=True)
        except OSError:
            raise Exception(f'PyTorch tensor file at "{path}" does not exist or cannot be read.')
        t = t.to_sparse_coo().coalesce()
        idx = t.indices().numpy()
        return scipy.sparse.csr_matrix((t.values().double().numpy(), (idx[0], idx[1])), shape=tuple(t.shape))
    raise Exception("Invalid file format. Supported formats: .npz (SciPy) or .pt (PyTorch)")


class PyTorchAlgebraicPropagator:
    def __init__(self, target_geo, proj_geo, output_torch_tensor=False, dtype=torch.float32, matrix=None,
                 device=None):
        self.logger = logging.getLogger(__name__)
        self.device = require_gpu(device)
        self.lib = _lib.load()
        self.output_torch_tensor = output_torch_tensor
        if dtype != torch.float32:
            raise ValueError("the HIP CSR kernel computes in float32")
        self.dtype = dtype
        m = _load_matrix(proj_geo.loading_path_for_matrix) if matrix is None else scipy.sparse.csr_matrix(matrix)
        self.P = _DeviceCSR(m, self.device)
        self.P_H = _DeviceCSR(m.T.tocsr(), self.device)          # explicit transpose for backward
        self.internal_matrix_shape = m.shape
        self.n_row_internal, self.n_col_internal = m.shape
        self.n_col_eff = int(np.asarray(target_geo.array).size)
        if self.n_col_eff % self.n_col_internal != 0:
            raise Exception(
                f"The imported sparse matrix has {self.n_col_internal} columns, and it cannot match or be tiled to "
                f"match the total number of voxels ({self.n_col_eff}) in the real space target.")
        self.z_tiling = self.n_col_eff // self.n_col_internal
        self.n_row_eff = self.n_row_internal * self.z_tiling
        self.shape = (self.n_row_eff, self.n_col_eff)

    def _apply(self, M: _DeviceCSR, v, n_in, n_out):
        torch_in = isinstance(v, torch.Tensor)
        t = torch.as_tensor(v, dtype=torch.float32, device=self.device).reshape(-1).contiguous()
        if t.numel() != n_in * self.z_tiling:
            raise ValueError("input has %d elements, expected %d" % (t.numel(), n_in * self.z_tiling))
        out = torch.empty(n_out * self.z_tiling, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.vamx_csr_spmm(M.row_ptr.data_ptr(), M.col.data_ptr(), M.val.data_ptr(),
                                              M.shape[0], M.shape[1], self.z_tiling, t.data_ptr(), out.data_ptr(),
                                              _lib.stream_ptr(self.device)), "vamx_csr_spmm")
        if self.output_torch_tensor or torch_in:
            return out
        return out.cpu().numpy()

    def forward(self, x):
        """b = P x with z-tiling (pyTorchAlgebraicPropagation.py:152-166)."""
        return self._apply(self.P, x, self.n_col_internal, self.n_row_internal)

    def backward(self, b):
        """x = P^T b with z-tiling (pyTorchAlgebraicPropagation.py:168-182)."""
        return self._apply(self.P_H, b, self.n_row_internal, self.n_col_internal)
