"""Float64 restatement of the reference's curved-ray tracing propagator.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference: ``vamtoolbox/projector/pyTorchRayTrace.py`` and ``vamtoolbox/medium.py``.

* rays        RayState.setupRaysParallel (1515-1666) and resetRaysIterateToInitial (1668-1700)
* stepping    _forwardSymplecticEuler (738-775), _forwardEuler (777-814),
              _leapfrog_init (816-828), parametrisations (692-735); "rk4" is a
              placeholder in the reference (289-292) and is defined here as the
              classical RK4 on the (x, v) system of the parametrisation
* attenuance  A += alpha_att(x_old) * |x_new - x_old|, E = E0 exp(-A) (763-774)
* interaction at x_new: index coordinate (x + span/2)/voxel (1253-1260), taps
              floor / floor+1 with weight (h/voxel_x) * prod(1 - |idx - tap|) * E
              (939-1143).  The reference takes floor / ceil, which counts a tap
              twice when a coordinate is exactly integral (1268-1273); this
              oracle follows the build, which uses standard trilinear taps.
* exit test   every ``exit_every`` steps, starting at step 0 (348-363, 1391-1420)
* forward     g = W (alpha_abs f)            (integrateEnergyUntilExit, 471-571)
* backward    dose = alpha_abs (W^T g)       (depositEnergyUntilExit, 378-469)
* media       grid_sample trilinear, align_corners=True, "border" padding for
              n and grad n, "zeros" padding for an attenuation grid
              (medium.py:161-211); central differences with zero boundary
              faces for grad n (224-273); analytic cylinder / ball attenuation
              (895-915); Luneburg lens (486-549), presampled on the grid
              (luneburg_n) or in closed form at every sample (LuneburgMedium).
"""

from __future__ import annotations

import numpy as np


# --------------------------------------------------------------------------
# media
# --------------------------------------------------------------------------
def coord_vec(nx, ny, nz, rate=500.0):
    """Volume.constructCoordVec (geometry.py:365-430); 2-D when nz == 0."""
    xv = np.linspace(-(nx - 1) / (2 * rate), (nx - 1) / (2 * rate), nx)
    yv = np.linspace(-(ny - 1) / (2 * rate), (ny - 1) / (2 * rate), ny)
    zv = np.atleast_1d(0.0) if nz == 0 else np.linspace(-(nz - 1) / (2 * rate), (nz - 1) / (2 * rate), nz)
    return [xv, yv, zv]


def central_difference(n, voxel):
    """MediumModel.centralFiniteDifference (medium.py:224-273)."""
    g = np.zeros(n.shape + (3,))
    g[1:-1, :, :, 0] = (n[2:] - n[:-2]) / (2 * voxel[0])
    g[:, 1:-1, :, 1] = (n[:, 2:] - n[:, :-2]) / (2 * voxel[1])
    if n.shape[2] > 1:
        g[:, :, 1:-1, 2] = (n[:, :, 2:] - n[:, :, :-2]) / (2 * voxel[2])
    return g


class GridMedium:
    """Index field sampled like IndexModel(type='interpolation') (medium.py:395-471)."""

    def __init__(self, cv, n, grad=None):
        self.xv, self.yv, self.zv = [np.asarray(c, dtype=np.float64) for c in cv]
        self.n = np.asarray(n, dtype=np.float64)
        self.vmax = np.array([self.xv.max(), self.yv.max(), self.zv.max()])
        self.vmin = np.array([self.xv.min(), self.yv.min(), self.zv.min()])
        if self.zv.size == 1:
            self.voxel = np.array([self.xv[1] - self.xv[0], self.yv[1] - self.yv[0], np.inf])
        else:
            self.voxel = np.array([self.xv[1] - self.xv[0], self.yv[1] - self.yv[0], self.zv[1] - self.zv[0]])
        self.span = self.vmax - self.vmin
        # 2 / span, and 0 for the zero z span of a 2-D grid (medium.py:69-76)
        self.norm = np.divide(2.0, self.span, out=np.zeros(3), where=self.span > 0)
        self.grad = central_difference(self.n, self.voxel) if grad is None else np.asarray(grad, dtype=np.float64)
        self.field = np.concatenate([self.n[..., None], self.grad], axis=3)  # (nx, ny, nz, 4)

    def _index(self, x):
        xn = x * self.norm                                           # normalizePosition
        sizes = np.array(self.n.shape, dtype=np.float64)
        return (xn + 1.0) / 2.0 * (sizes - 1.0)                     # align_corners=True

    def sample(self, x, field=None, padding="border"):
        """Trilinear grid_sample of ``field`` (default n and grad n) at x (m, 3)."""
        f = self.field if field is None else field
        if f.ndim == 3:
            f = f[..., None]
        idx = self._index(x)
        shp = np.array(f.shape[:3])
        if padding == "border":
            idx = np.clip(idx, 0.0, shp - 1.0)
        fl = np.floor(idx)
        w = idx - fl
        i0 = fl.astype(np.int64)
        out = np.zeros((x.shape[0], f.shape[3]))
        for c in range(8):
            o = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
            ii = i0 + o
            ww = np.prod(np.where(o == 1, w, 1.0 - w), axis=1)
            ok = np.all((ii >= 0) & (ii < shp), axis=1) & (ww != 0)
            iic = np.clip(ii, 0, shp - 1)
            out += np.where(ok[:, None], ww[:, None] * f[iic[:, 0], iic[:, 1], iic[:, 2]], 0.0)
        return out


def luneburg_n(cv, R=1.0, p=2.0, n_sur=1.0):
    """IndexModel._n_lune / _grad_n_lune presampled on the grid (medium.py:486-549)."""
    xv, yv, zv = cv
    Rd = min((xv.max() - xv.min()) / 2, (yv.max() - yv.min()) / 2,
             *([] if zv.size == 1 else [(zv.max() - zv.min()) / 2]))
    Rp = R * Rd
    X, Y, Z = np.meshgrid(xv, yv, zv, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    n = np.ones_like(r)
    inside = r < Rp
    n[inside] = n_sur * (2 - (r[inside] / Rp) ** p) ** (1 / p)
    if n_sur != 1.0:
        n[~inside] = n_sur
    g = np.zeros(r.shape + (3,))
    c = (n_sur / Rp) ** p
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = -c * (n ** (1 - p)) * (r ** (p - 2))
    fac = np.where(inside, fac, 0.0)
    g[..., 0] = fac * X
    g[..., 1] = fac * Y
    g[..., 2] = fac * Z
    return n, g, Rp


class LuneburgMedium(GridMedium):
    """IndexModel(type="analytical", form="luneburg_lens") (medium.py:486-549):
    n and grad n in closed form at every sample position (no grid sampling);
    the grid of ``cv`` serves the interaction taps and the exit box."""

    def __init__(self, cv, R=1.0, p=2.0, n_sur=1.0):
        nf, g, Rp = luneburg_n(cv, R, p, n_sur)
        super().__init__(cv, nf, grad=g)
        self.Rp, self.p, self.n_sur = Rp, float(p), float(n_sur)

    def sample(self, x, field=None, padding="border"):
        if field is not None:
            return super().sample(x, field, padding)
        r = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2)
        n = np.full(r.shape, self.n_sur)
        inside = r < self.Rp
        n[inside] = self.n_sur * (2 - (r[inside] / self.Rp) ** self.p) ** (1 / self.p)
        g = np.zeros(x.shape)
        c = (self.n_sur / self.Rp) ** self.p
        g[inside] = (-c * (n[inside] ** (1 - self.p)) * (r[inside] ** (self.p - 2)))[:, None] * x[inside]
        return np.concatenate([n[:, None], g], axis=1)


def att_alpha(x, kind, alpha=0.0, radius=np.inf, grid=None, medium=None):
    """AttenuationModel.alpha (medium.py:895-915, 161-185 zeros padding)."""
    if kind == "none":
        return np.zeros(x.shape[0])
    if kind == "cylinder":
        return np.where(np.linalg.norm(x[:, :2], axis=1) < radius, alpha, 0.0)
    if kind == "ball":
        return np.where(np.linalg.norm(x, axis=1) < radius, alpha, 0.0)
    if kind == "grid":
        return medium.sample(x, field=grid, padding="zeros")[:, 0]
    raise ValueError(kind)


# --------------------------------------------------------------------------
# rays
# --------------------------------------------------------------------------
def setup_rays(cv, angles_deg, incl_deg=0.0, ray_density=1.0):
    """setupRaysParallel (pyTorchRayTrace.py:1515-1666) in float64."""
    xv, yv, zv = cv
    nX, nY, nZ = xv.size, yv.size, zv.size
    if nZ == 1:
        shape = (round(max(nX, nY) * ray_density), len(angles_deg), nZ)
    else:
        shape = (round(max(nX, nY) * ray_density), len(angles_deg), max(round(nZ * ray_density), 1))
    n0min = min(xv.min(), yv.min())
    n0max = max(xv.max(), yv.max())
    n0 = np.linspace(n0min, n0max, shape[0])
    n2 = np.linspace(zv.min(), zv.max(), shape[2])
    th = np.asarray(angles_deg, dtype=np.float64) * np.pi / 180.0
    phi = incl_deg * np.pi / 180.0
    D = np.sqrt(n0max ** 2 + zv.max() ** 2)
    G0, G1, G2 = np.meshgrid(n0, th, n2, indexing="ij")
    G0, G1, G2 = G0.ravel(), G1.ravel(), G2.ravel()
    x0 = np.empty((G0.size, 3))
    x0[:, 0] = D * np.cos(G1) * np.cos(phi) + (-G0 * np.sin(G1) - G2 * np.cos(G1) * np.sin(phi))
    x0[:, 1] = D * np.sin(G1) * np.cos(phi) + (G0 * np.cos(G1) - G2 * np.sin(G1) * np.sin(phi))
    x0[:, 2] = D * np.sin(phi) + G2 * np.cos(phi)
    v0 = np.empty((G0.size, 3))
    v0[:, 0] = -np.cos(G1) * np.cos(phi)
    v0[:, 1] = -np.sin(G1) * np.cos(phi)
    v0[:, 2] = -np.sin(phi)
    e0 = ray_density ** (-1.0) if n2.size == 1 else ray_density ** (-2.0)
    return x0, v0, shape, e0


# --------------------------------------------------------------------------
# solver
# --------------------------------------------------------------------------
class Tracer:
    def __init__(self, medium: GridMedium, param="canonical", solver="forward_symplectic_euler",
                 step=None, max_steps=None, exit_every=10, att=("none", 0.0, np.inf, None),
                 abs_grid=None):
        self.m = medium
        self.param = param
        self.solver = solver
        self.h = medium.voxel[0] / 2 if step is None else step
        self.max_steps = int(medium.n.shape[0] * 3) if max_steps is None else int(max_steps)
        self.exit_every = exit_every
        self.att = att
        self.abs_grid = abs_grid

    def _rhs(self, x, v):
        f = self.m.sample(x)
        n, gn = f[:, 0:1], f[:, 1:4]
        if self.param == "canonical":
            return v, n * gn
        if self.param == "physical_path_length":
            return v / n, gn
        return v * n ** (-2), gn / n

    def _dv(self, x):
        f = self.m.sample(x)
        n, gn = f[:, 0:1], f[:, 1:4]
        if self.param == "canonical":
            return n * gn, n
        if self.param == "physical_path_length":
            return gn, n
        return gn / n, n

    def _dx(self, v, n):
        if self.param == "canonical":
            return v
        if self.param == "physical_path_length":
            return v / n
        return v * n ** (-2)

    def advance(self, x, v):
        h = self.h
        if self.solver == "rk4":
            k1x, k1v = self._rhs(x, v)
            k2x, k2v = self._rhs(x + 0.5 * h * k1x, v + 0.5 * h * k1v)
            k3x, k3v = self._rhs(x + 0.5 * h * k2x, v + 0.5 * h * k2v)
            k4x, k4v = self._rhs(x + h * k3x, v + h * k3v)
            return x + h / 6.0 * ((k1x + 2.0 * k2x) + (2.0 * k3x + k4x)), \
                v + h / 6.0 * ((k1v + 2.0 * k2v) + (2.0 * k3v + k4v))
        dv, n = self._dv(x)
        vnew = v + h * dv
        vx = v if self.solver == "forward_euler" else vnew
        return x + h * self._dx(vx, n), vnew

    def _taps(self, x):
        m = self.m
        idx = (x - (-m.span / 2)) * (1.0 / m.voxel)                 # 1253-1260
        if m.zv.size == 1:
            idx[:, 2] = 0.0
        fl = np.floor(idx)
        w = idx - fl
        return fl.astype(np.int64), w

    def _exited(self, x, v):
        m = self.m
        pos = (x > m.vmax + m.voxel) & (v > 0.0)
        neg = (x < m.vmin - m.voxel) & (v < 0.0)
        return pos.any(axis=1) | neg.any(axis=1)

    def _alpha_att(self, x):
        kind, a, r, grid = self.att
        return att_alpha(x, kind, a, r, grid, self.m)

    def run(self, x0, v0, e0, f=None, g=None, return_steps=False):
        """Trace every ray; integrate f (forward) or deposit g (backward)."""
        m = self.m
        shp = np.array(m.n.shape)
        nz1 = m.zv.size == 1
        x = x0.copy()
        v = v0.copy()
        nr = x.shape[0]
        E0 = np.full(nr, float(e0)) if g is None else np.asarray(g, dtype=np.float64).ravel().copy()
        if self.solver == "leapfrog":
            dv, _ = self._dv(x)
            v = v - dv * self.h / 2.0
        A = np.zeros(nr)
        integral = np.zeros(nr)
        dose = np.zeros(m.n.shape)
        feff = None
        if f is not None:
            feff = np.asarray(f, dtype=np.float64) * (1.0 if self.abs_grid is None else self.abs_grid)
        active = np.ones(nr, dtype=bool)
        steps = 0
        hw = self.h / m.voxel[0]
        for step in range(self.max_steps):
            ia = np.nonzero(active)[0]
            if ia.size == 0:
                break
            xo = x[ia]
            xn, vn = self.advance(xo, v[ia])
            x[ia], v[ia] = xn, vn
            ds = np.linalg.norm(xn - xo, axis=1)
            A[ia] += self._alpha_att(xo) * ds
            E = E0[ia] * np.exp(-A[ia])
            i0, w = self._taps(xn)
            steps += ia.size
            for c in range(8):
                o = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
                if nz1 and o[2]:
                    continue
                ii = i0 + o
                ww = np.prod(np.where(o == 1, w, 1.0 - w)[:, :(2 if nz1 else 3)], axis=1)
                ok = np.all((ii >= 0) & (ii < shp), axis=1) & (ww != 0)
                if not ok.any():
                    continue
                sel = np.nonzero(ok)[0]
                wt = hw * ww[sel] * E[sel]
                iic = ii[sel]
                if f is not None:
                    integral[ia[sel]] += feff[iic[:, 0], iic[:, 1], iic[:, 2]] * wt
                else:
                    np.add.at(dose, (iic[:, 0], iic[:, 1], iic[:, 2]), wt)
            if step % self.exit_every == 0:
                active &= ~self._exited(x, v)
        still = int(active.sum())
        if f is not None:
            out = integral
        else:
            out = dose * (1.0 if self.abs_grid is None else self.abs_grid)
        if return_steps:
            return out, steps, still
        return out
