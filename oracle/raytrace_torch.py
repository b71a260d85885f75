"""The reference's CPU ray-tracing path restated on CPU torch, as the timed
CPU baseline of the ray tracer.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by
``bench.py``'s ``cpu_baseline`` leg and by ``tests/`` only.

It follows the tensor program of ``vamtoolbox/projector/pyTorchRayTrace.py``
on ``torch.device("cpu")`` step for step, so that its cost is the
reference's cost on the host cores:

* ``integrateEnergyUntilExit`` / ``depositEnergyUntilExit`` (378-571): a
  Python loop over steps on a boolean active set, the exit test every
  ``exit_every`` steps (348-363, 1343-1424);
* ``_forwardSymplecticEuler`` (738-775): v_{i+1} = v_i + h dv(x_i),
  x_{i+1} = x_i + h dx(v_{i+1}), attenuance at x_i, E = E0 exp(-A);
* the index field and its gradient through
  ``torch.nn.functional.grid_sample`` (bilinear, align_corners=True, border
  padding) on (N, C, Z, Y, X) fields (``medium.py:146-211``);
* interactions at x_{i+1}: floor / ceil taps, weight
  (h / voxel) * prod(1 - |idx - tap|) * E (939-1143, 1253-1280), integrate by
  gather, deposit by uncoalesced sparse COO tensors added to the grid
  (1039-1043, 1323-1335).

Ray state in float32 (the reference's default ``tensor_dtype``), fields in
float32.
"""

from __future__ import annotations

import numpy as np
import torch


class TorchCPUTracer:
    def __init__(self, cv, n_field, grad=None, att_alpha=0.0, att_radius=np.inf, abs_grid=None, step=None,
                 max_steps=None, exit_every=10, param="canonical", dtype=torch.float32):
        self.dtype = dtype
        self.xv, self.yv, self.zv = [torch.as_tensor(np.asarray(c), dtype=torch.float64) for c in cv]
        n = np.asarray(n_field, dtype=np.float64)
        self.shape = n.shape
        self.voxel = torch.tensor([float(self.xv[1] - self.xv[0]), float(self.yv[1] - self.yv[0]),
                                   float(self.zv[1] - self.zv[0]) if self.zv.numel() > 1 else np.inf],
                                  dtype=dtype)
        span = [float(c.max() - c.min()) for c in (self.xv, self.yv, self.zv)]
        self.span = torch.tensor(span, dtype=dtype)
        self.norm = torch.tensor([2.0 / s if s > 0 else 0.0 for s in span], dtype=dtype)
        self.vmax = torch.tensor([float(c.max()) for c in (self.xv, self.yv, self.zv)], dtype=dtype)
        self.vmin = torch.tensor([float(c.min()) for c in (self.xv, self.yv, self.zv)], dtype=dtype)
        if grad is None:
            g = np.zeros(n.shape + (3,))
            vox = [float(self.voxel[0]), float(self.voxel[1]), float(self.voxel[2])]
            g[1:-1, :, :, 0] = (n[2:] - n[:-2]) / (2 * vox[0])
            g[:, 1:-1, :, 1] = (n[:, 2:] - n[:, :-2]) / (2 * vox[1])
            if n.shape[2] > 1:
                g[:, :, 1:-1, 2] = (n[:, :, 2:] - n[:, :, :-2]) / (2 * vox[2])
            grad = g
        # (N, C, Z, Y, X) as medium.py:146-158
        self.n5 = torch.as_tensor(n, dtype=dtype).permute(2, 1, 0)[None, None].contiguous()
        self.g5 = torch.as_tensor(np.asarray(grad), dtype=dtype).permute(3, 2, 1, 0)[None].contiguous()
        self.att_alpha = float(att_alpha)
        self.att_radius = float(att_radius)
        self.abs_grid = None if abs_grid is None else torch.as_tensor(np.asarray(abs_grid), dtype=dtype)
        self.h = float(self.voxel[0]) / 2 if step is None else float(step)
        self.max_steps = int(n.shape[0] * 3) if max_steps is None else int(max_steps)
        self.exit_every = int(exit_every)
        self.param = param

    # medium.py:161-211
    def _sample(self, field5, x):
        xn = x * self.norm
        out = torch.nn.functional.grid_sample(field5, xn[None, None, None, :, :], mode="bilinear",
                                              padding_mode="border", align_corners=True)
        return out[0, :, 0, 0, :].T

    def _dv(self, x, v):
        n = self._sample(self.n5, x)
        g = self._sample(self.g5, x)
        if self.param == "canonical":
            return n * g
        if self.param == "physical_path_length":
            return g
        return g / n

    def _dx(self, x, v):
        if self.param == "canonical":
            return v
        n = self._sample(self.n5, x)
        if self.param == "physical_path_length":
            return v / n
        return v / n ** 2

    def _alpha(self, x):
        r = torch.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2)
        return torch.where(r < self.att_radius, torch.full_like(r, self.att_alpha), torch.zeros_like(r))

    def _exited(self, x, v):
        pos = (x > self.vmax + self.voxel) & (v > 0)
        neg = (x < self.vmin - self.voxel) & (v < 0)
        return torch.any(pos | neg, dim=1)

    def run(self, x0, v0, e0, f=None, g=None):
        """Trace all rays; integrate f (forward) or deposit g (backward)."""
        dt = self.dtype
        x_ip1 = torch.as_tensor(x0, dtype=dt).clone()
        v_ip1 = torch.as_tensor(v0, dtype=dt).clone()
        nr = x_ip1.shape[0]
        E0 = torch.full((nr,), float(e0), dtype=dt) if g is None else torch.as_tensor(np.ravel(g), dtype=dt)
        A = torch.zeros(nr, dtype=dt)
        E = E0.clone()
        integral = torch.zeros(nr, dtype=dt)
        grid = torch.zeros(self.shape, dtype=dt)
        feff = None
        if f is not None:
            feff = torch.as_tensor(np.asarray(f), dtype=dt)
            if self.abs_grid is not None:
                feff = feff * self.abs_grid
        active = torch.ones(nr, dtype=torch.bool)
        h = self.h
        hw = h / float(self.voxel[0])
        sizes = torch.tensor(self.shape)
        n_adj = 4 if self.shape[2] == 1 else 8
        step = 0
        while step < self.max_steps and bool(active.any()):
            # _forwardSymplecticEuler (738-775)
            x_i = x_ip1[active]
            v_i = v_ip1[active]
            v_new = v_i + self._dv(x_i, v_i) * h
            dx = self._dx(x_i, v_new) * h
            x_new = x_i + dx
            ds = torch.linalg.norm(dx, dim=1)
            A[active] += self._alpha(x_i) * ds
            E[active] = E0[active] * torch.exp(-A[active])
            x_ip1[active] = x_new
            v_ip1[active] = v_new
            # interactions at x_{i+1} (835-1143)
            idx = (x_new - (-self.span / 2)) * (1 / self.voxel)
            if self.shape[2] == 1:
                idx[:, 2] = 0.0
            lo = torch.floor(idx).to(torch.long)
            hi = torch.ceil(idx).to(torch.long)
            Ea = E[active]
            for c in range(n_adj):
                sel = torch.stack([hi[:, k] if (c >> k) & 1 else lo[:, k] for k in range(3)], dim=1)
                valid = torch.all((sel >= 0) & (sel < sizes), dim=1)
                if not bool(valid.any()):
                    continue
                coef = torch.prod(1 - torch.abs(idx[valid] - sel[valid].to(dt)), dim=1)
                val = hw * coef * Ea[valid]
                sv = sel[valid]
                if feff is not None:
                    ia = torch.nonzero(active)[:, 0][valid]
                    integral[ia] += val * feff[sv[:, 0], sv[:, 1], sv[:, 2]]
                else:
                    grid += torch.sparse_coo_tensor(sv.T, val, size=self.shape)
            if step % self.exit_every == 0:
                active &= ~self._exited(x_ip1, v_ip1)
            step += 1
        if feff is not None:
            return integral.numpy()
        if self.abs_grid is not None:
            grid = grid * self.abs_grid
        return grid.numpy()
