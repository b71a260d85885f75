"""Propagation media for the curved-ray tracer.

Mirrors ``vamtoolbox/medium.py``: ``IndexModel`` (276-549),
``AttenuationModel`` (753-915), ``AbsorptionModel`` (993-996) with the same
constructor arguments (``coord_vec``, ``type``, ``form``, keyword parameters
R, length_unit_of_R, p, n_sur, alpha_internal, n_x, grad_n_x, alpha_x,
interpolation_padding_mode).  Fields are held as float64 NumPy arrays on the
(x, y, z) grid ('ij' order, axis 0 = x) and uploaded to the GPU as the
float32 (n, dn/dx, dn/dy, dn/dz) field the ray-trace kernel samples with
grid_sample semantics (trilinear, align_corners=True, border padding).
"""

from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


class MediumModel:
    """Coordinate grid shared by the index and attenuation models (medium.py:13-84)."""

    def __init__(self, coord_vec):
        self.xv = np.asarray(coord_vec[0], dtype=np.float64)
        self.yv = np.asarray(coord_vec[1], dtype=np.float64)
        self.zv = np.asarray(coord_vec[2], dtype=np.float64)
        self.xv_yv_zv_max = np.array([self.xv.max(), self.yv.max(), self.zv.max()])
        self.xv_yv_zv_min = np.array([self.xv.min(), self.yv.min(), self.zv.min()])
        if self.zv.size == 1:
            self.voxel_size = np.array([self.xv[1] - self.xv[0], self.yv[1] - self.yv[0], np.inf])
        else:
            self.voxel_size = np.array([self.xv[1] - self.xv[0], self.yv[1] - self.yv[0], self.zv[1] - self.zv[0]])
        self.grid_span = self.xv_yv_zv_max - self.xv_yv_zv_min
        self.shape = (self.xv.size, self.yv.size, self.zv.size)
        self.interpolation_padding_mode = "border"

    def _plot_mid_slice(self, values, title, fig, ax, block):
        """Middle z slice of a field on the grid (medium.py:570-599, 918-947)."""
        import matplotlib.pyplot as plt
        sl = np.asarray(values).reshape(self.shape)[:, :, self.shape[2] // 2]
        if ax is None:
            fig, ax = plt.subplots()
        mappable = ax.imshow(sl, cmap="gray", vmin=np.amin(sl), vmax=np.amax(sl))
        ax.set_title(title)
        ax.set_xlabel("x")
        ax.set_ylabel("y")
        plt.colorbar(mappable)
        if block:
            plt.show(block=True)
        return fig, ax

    def getPositionVectorsAtGridPoints(self):
        """(num_points, 3) positions of the grid points (medium.py:218-222)."""
        xg, yg, zg = np.meshgrid(self.xv, self.yv, self.zv, indexing="ij")
        return np.vstack((xg.ravel(), yg.ravel(), zg.ravel())).T

    def _fractional_radius(self, params):
        if params["length_unit_of_R"] == "fractional_domain_radius":
            r0 = (self.xv.max() - self.xv.min()) / 2
            r1 = (self.yv.max() - self.yv.min()) / 2
            r2 = (self.zv.max() - self.zv.min()) / 2
            rmin = min(r0, r1) if self.zv.size == 1 else min(r0, r1, r2)
            return params["R"] * rmin
        if params["length_unit_of_R"] == "physical_coordinate":
            return params["R"]
        raise Exception('"length_unit" should be either "fractional_domain_radius" or "physical_coordinate"')

    @staticmethod
    def centralFiniteDifference(field, voxel_size):
        """Central differences with zero boundary faces (medium.py:224-273)."""
        f = np.asarray(field, dtype=np.float64)
        g = np.zeros(f.shape + (3,))
        g[1:-1, :, :, 0] = (f[2:] - f[:-2]) / (2 * voxel_size[0])
        g[:, 1:-1, :, 1] = (f[:, 2:] - f[:, :-2]) / (2 * voxel_size[1])
        if f.shape[2] > 1 and not np.isnan(voxel_size[2]):
            g[:, :, 1:-1, 2] = (f[:, :, 2:] - f[:, :, :-2]) / (2 * voxel_size[2])
        return g


class IndexModel(MediumModel):
    """Real refractive index n(x) and its gradient (medium.py:276-549)."""

    _default_analytical = {"R": 1.0, "length_unit_of_R": "fractional_domain_radius", "p": 2.0, "n_sur": 1.0}
    _default_interpolation = {"interpolation_padding_mode": "border"}

    def __init__(self, coord_vec, type: str = "analytical", form: str = "homogeneous", **kwargs):
        super().__init__(coord_vec)
        self.type = type
        self.form = form
        self.params = self._default_analytical.copy()
        self.params.update(self._default_interpolation)
        self.params.update(kwargs)
        self.params["R_physical"] = self._fractional_radius(self.params)
        if type not in ("analytical", "interpolation"):
            raise Exception('"type" should be either "analytical" or "interpolation".')
        if form in ("maxwell_lens", "eaton_lens"):
            raise Exception("To be implemented.")
        x = self.getPositionVectorsAtGridPoints()
        if form == "homogeneous":
            n = self._n_homo(x)
            grad = np.zeros_like(x)
        elif form == "luneburg_lens":
            # analytical: the kernel evaluates the closed form at every sample;
            # interpolation: the closed form presampled on the grid
            n = self._n_lune(x)
            grad = self._grad_n_lune(x)
        elif form == "freeform":
            if type != "interpolation":
                raise Exception("Form: Other analytical functions are not supported yet.")
            n = np.asarray(self.params.get("n_x"), dtype=np.float64).ravel()
            g = self.params.get("grad_n_x", None)
            grad = None if g is None else np.asarray(g, dtype=np.float64).reshape(-1, 3)
        else:
            raise Exception("Other interpolation functions are not supported yet.")
        self.presampled_scalar_field_3D = n.reshape(self.shape)
        if self.presampled_scalar_field_3D.shape != self.shape:
            raise Exception("Size mismatch between the constructed coordinate grid and imported scalar field.")
        if grad is None:
            self.presampled_vector_field_4D = self.centralFiniteDifference(self.presampled_scalar_field_3D,
                                                                           self.voxel_size)
        else:
            self.presampled_vector_field_4D = grad.reshape(self.shape + (3,))
        self._dev = {}

    def _n_homo(self, x):
        return self.params["n_sur"] * np.ones(x.shape[0])

    def _n_lune(self, x=None, r_known=None):
        r = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2) if r_known is None else r_known
        n = np.ones_like(r)
        rp, p, ns = self.params["R_physical"], self.params["p"], self.params["n_sur"]
        inside = r < rp
        n[inside] = ns * (2 - (r[inside] / rp) ** p) ** (1 / p)
        if ns != 1.0:
            n[r >= rp] = ns
        return n

    def _grad_n_lune(self, x):
        r = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2)
        n = self._n_lune(r_known=r)
        grad = np.zeros_like(x)
        rp, p, ns = self.params["R_physical"], self.params["p"], self.params["n_sur"]
        c = (ns / rp) ** p
        inside = r < rp
        with np.errstate(divide="ignore", invalid="ignore"):
            grad[inside, :] = -c * ((n[inside] ** (1 - p)) * (r[inside] ** (p - 2)))[:, None] * x[inside, :]
        return grad

    def n(self, x):
        """n at positions x (num, 3): the closed form for the analytical
        models, trilinear grid sampling (border padding) otherwise."""
        x = np.asarray(x, dtype=np.float64)
        if self.type == "analytical":
            return self._n_homo(x) if self.form == "homogeneous" else self._n_lune(x)
        return _grid_sample(self, self.presampled_scalar_field_3D[..., None], x, "border")[:, 0]

    def plotIndex(self, fig=None, ax=None, block=False):
        return self._plot_mid_slice(self.n(self.getPositionVectorsAtGridPoints()), "Index distribution", fig, ax,
                                    block)

    def grad_n(self, x):
        x = np.asarray(x, dtype=np.float64)
        if self.type == "analytical":
            return np.zeros_like(x) if self.form == "homogeneous" else self._grad_n_lune(x)
        return _grid_sample(self, self.presampled_vector_field_4D, x, "border")

    def kernel_index_spec(self):
        """(kind, R, p, n_sur) of the ray-trace kernel's index model: the
        analytical Luneburg lens is evaluated in closed form; every other model
        (the homogeneous one included, whose grid samples are exact) is
        sampled from the float4 grid."""
        from . import _lib
        if self.type == "analytical" and self.form == "luneburg_lens":
            return (_lib.RT_INDEX_LUNEBURG, float(self.params["R_physical"]), float(self.params["p"]),
                    float(self.params["n_sur"]))
        return _lib.RT_INDEX_GRID, 0.0, 0.0, 0.0

    def device_field(self, device):
        """float32 (nx, ny, nz, 4) tensor (n, dn/dx, dn/dy, dn/dz) on ``device``."""
        key = str(device)
        t = self._dev.get(key)
        if t is None:
            f = np.concatenate([self.presampled_scalar_field_3D[..., None], self.presampled_vector_field_4D], axis=3)
            t = torch.as_tensor(np.ascontiguousarray(f, dtype=np.float32), device=device)
            self._dev[key] = t
        return t


class AttenuationModel(MediumModel):
    """Attenuation / absorption coefficient alpha(x) (medium.py:753-915)."""

    _default_analytical = {"R": 1.0, "length_unit_of_R": "fractional_domain_radius", "alpha_internal": 1e-3}
    _default_interpolation = {"interpolation_padding_mode": "zeros"}

    def __init__(self, coord_vec, type: str = "analytical", form: str = "homogeneous_cylinder", **kwargs):
        super().__init__(coord_vec)
        self.type = type
        self.form = form
        self.params = self._default_analytical.copy()
        self.params.update(self._default_interpolation)
        self.params.update(kwargs)
        self.params["R_physical"] = self._fractional_radius(self.params)
        if type == "analytical":
            if form not in ("homogeneous_cylinder", "homogeneous_ball"):
                raise Exception("Form: Other analytical functions are not supported yet.")
            self.presampled_scalar_field_3D = None
        elif type == "interpolation":
            x = self.getPositionVectorsAtGridPoints()
            if form == "homogeneous_cylinder":
                a = self._alpha_homo_cylinder(x)
            elif form == "homogeneous_ball":
                a = self._alpha_homo_ball(x)
            elif form == "freeform":
                a = np.asarray(self.params.get("alpha_x"), dtype=np.float64).ravel()
            else:
                raise Exception("Other interpolation functions are not supported yet.")
            self.presampled_scalar_field_3D = a.reshape(self.shape)
        else:
            raise Exception('"type" should be either "analytical" or "interpolation".')
        self._dev = {}

    def _alpha_homo_cylinder(self, x):
        r = np.linalg.norm(x[:, 0:2], axis=1)
        return np.where(r < self.params["R_physical"], self.params["alpha_internal"], 0.0)

    def _alpha_homo_ball(self, x):
        r = np.linalg.norm(x, axis=1)
        return np.where(r < self.params["R_physical"], self.params["alpha_internal"], 0.0)

    def alpha(self, x):
        x = np.asarray(x, dtype=np.float64)
        if self.type == "analytical":
            return self._alpha_homo_cylinder(x) if self.form == "homogeneous_cylinder" else self._alpha_homo_ball(x)
        return _grid_sample(self, self.presampled_scalar_field_3D[..., None], x, "zeros")[:, 0]

    def plotAlpha(self, fig=None, ax=None, block=False):
        return self._plot_mid_slice(self.alpha_at_grid(), "Attenuation coefficient", fig, ax, block)

    def alpha_at_grid(self):
        """alpha at the grid points, (nx, ny, nz) float64."""
        return self.alpha(self.getPositionVectorsAtGridPoints()).reshape(self.shape)

    def kernel_spec(self, device):
        """(kind, alpha, radius, grid tensor or None) for the ray-trace kernel."""
        from . import _lib
        if self.type == "analytical":
            kind = _lib.RT_ATT_CYLINDER if self.form == "homogeneous_cylinder" else _lib.RT_ATT_BALL
            return kind, float(self.params["alpha_internal"]), float(self.params["R_physical"]), None
        key = str(device)
        t = self._dev.get(key)
        if t is None:
            t = torch.as_tensor(np.ascontiguousarray(self.presampled_scalar_field_3D, dtype=np.float32), device=device)
            self._dev[key] = t
        return _lib.RT_ATT_GRID, 0.0, 0.0, t


class AbsorptionModel(AttenuationModel):
    """Alias of AttenuationModel (medium.py:993-996)."""


class ScatteringModel(AttenuationModel):
    """Reserved (medium.py:999-1002)."""


def _grid_sample(model: MediumModel, field, x, padding):
    """grid_sample(mode='bilinear', align_corners=True) on the host in float64."""
    norm = 2.0 / model.grid_span
    if np.isinf(norm[2]):
        norm[2] = 0.0
    sizes = np.array(field.shape[:3], dtype=np.float64)
    idx = ((x * norm) + 1.0) / 2.0 * (sizes - 1.0)
    if padding == "border":
        idx = np.clip(idx, 0.0, sizes - 1.0)
    fl = np.floor(idx)
    w = idx - fl
    i0 = fl.astype(np.int64)
    shp = np.array(field.shape[:3])
    out = np.zeros((x.shape[0], field.shape[3]))
    for c in range(8):
        o = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
        ii = i0 + o
        ww = np.prod(np.where(o == 1, w, 1.0 - w), axis=1)
        ok = np.all((ii >= 0) & (ii < shp), axis=1) & (ww != 0)
        iic = np.clip(ii, 0, shp - 1)
        out += np.where(ok[:, None], ww[:, None] * field[iic[:, 0], iic[:, 1], iic[:, 2]], 0.0)
    return out
