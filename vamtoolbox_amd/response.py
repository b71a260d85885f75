"""Material response model (optical dose -> material response).

Mirrors ``vamtoolbox/response.py:12-399`` (``ResponseModel`` with
``type`` "analytical" / "interpolation" and ``form`` "gen_log_fun" /
"linear" / "identity" / "freeform", same parameter names and defaults).
``map``/``dmapdf``/``map_inv`` accept NumPy arrays (float64, as the
reference) or torch tensors (evaluated on the tensor's device in float64
and returned in the tensor's dtype).  ``vamx_struct`` packs the model for
the fused BCLP kernels.
"""

from __future__ import annotations

import warnings

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _is_t(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


def _interp_t(x, xp, fp, left, right):
    """np.interp on torch (float64): same branch order as numpy."""
    xd = x.to(torch.float64)
    xp_t = torch.as_tensor(xp, device=x.device, dtype=torch.float64)
    fp_t = torch.as_tensor(fp, device=x.device, dtype=torch.float64)
    n = xp_t.numel()
    j = torch.searchsorted(xp_t, xd, right=True) - 1           # xp[j] <= x < xp[j+1]
    j = j.clamp(0, n - 2)
    slope = (fp_t[j + 1] - fp_t[j]) / (xp_t[j + 1] - xp_t[j])
    val = slope * (xd - xp_t[j]) + fp_t[j]
    val = torch.where(xd == xp_t[j], fp_t[j], val)
    val = torch.where(xd == xp_t[-1], fp_t[-1], val)
    val = torch.where(xd > xp_t[-1], torch.as_tensor(float(right), device=x.device, dtype=torch.float64), val)
    val = torch.where(xd < xp_t[0], torch.as_tensor(float(left), device=x.device, dtype=torch.float64), val)
    return val.to(x.dtype)


class ResponseModel:
    _default_gen_log_fun = {"A": 0, "K": 1, "B": 25, "M": 0.5, "nu": 1}
    _default_linear = {"M": 1, "C": 0}
    _default_interpolation = {"interp_min": 0, "interp_max": 1, "n_pts": 512}

    def __init__(self, type: str = "interpolation", form: str = "gen_log_fun", **kwargs):
        self.type = type
        self.form = form
        if self.type == "analytical":
            if self.form == "gen_log_fun":
                self.params = self._default_gen_log_fun.copy()
                self.params.update(kwargs)
                self.map, self.dmapdf, self.map_inv = self._map_glf, self._dmapdf_glf, self._map_inv_glf
            elif self.form == "linear":
                self.params = self._default_linear.copy()
                self.params.update(kwargs)
                self.map, self.dmapdf, self.map_inv = self._map_lin, self._dmapdf_lin, self._map_inv_lin
            elif self.form == "identity":
                self.params = {}
                self.map, self.dmapdf, self.map_inv = self._map_id, self._dmapdf_id, self._map_inv_id
            else:
                raise Exception("Form: Other analytical functions are not supported yet.")
        elif self.type == "interpolation":
            self.map, self.dmapdf, self.map_inv = self._map_interp, self._dmapdf_interp, self._map_inv_interp
            self.params = self._default_interpolation.copy()
            if self.form in ("gen_log_fun", "linear", "identity"):
                if self.form == "gen_log_fun":
                    self.params.update(self._default_gen_log_fun)
                elif self.form == "linear":
                    self.params.update(self._default_linear)
                self.params.update(kwargs)
                self.interp_f_0 = np.linspace(self.params["interp_min"], self.params["interp_max"],
                                              self.params["n_pts"])
                fmap = {"gen_log_fun": (self._map_glf, self._dmapdf_glf),
                        "linear": (self._map_lin, self._dmapdf_lin),
                        "identity": (self._map_id, self._dmapdf_id)}[self.form]
                self.interp_map_0 = np.asarray(fmap[0](self.interp_f_0), dtype=np.float64).copy()
                self.interp_dmapdf_0 = np.asarray(fmap[1](self.interp_f_0), dtype=np.float64).copy()
            elif self.form == "freeform":
                self.interp_f_0 = kwargs.get("interp_f_0", None)
                self.interp_map_0 = kwargs.get("interp_map_0", None)
                if (len(self.interp_f_0.shape) > 1) or (len(self.interp_map_0.shape) > 1):
                    raise Exception('Imported data for material response curve should be 1D. Check "interp_f_0" '
                                    'and "interp_map_0".')
                if self.interp_f_0.shape != self.interp_map_0.shape:
                    raise Exception('Size mismatch between "interp_f_0" and "interp_map_0".')
                self.interp_dmapdf_0 = np.diff(
                    self.interp_map_0, n=1,
                    append=(self.interp_map_0[-1] + (self.interp_map_0[-1] - self.interp_map_0[-2])))
            else:
                raise Exception("Other interpolation functions are not supported yet.")
        else:
            raise Exception('Mapping type ("type") should be either "analytical" or "interpolation".')
        self._dev_tables = {}

    # ---- generalised logistic function (response.py:191-229) -------------
    def _map_glf(self, f):
        p = self.params
        if _is_t(f):
            fd = f.to(torch.float64)
            e = torch.exp(-p["B"] * (fd - p["M"]))
            return (p["A"] + (p["K"] - p["A"]) / (1 + e) ** (1 / p["nu"])).to(f.dtype)
        e = np.exp(-p["B"] * (f - p["M"]))
        return p["A"] + (p["K"] - p["A"]) / (1 + e) ** (1 / p["nu"])

    def _dmapdf_glf(self, f, use_cached_result: bool = False):
        p = self.params
        c1 = (1 / (p["K"] - p["A"])) ** p["nu"]
        c2 = p["B"] / p["nu"]
        if _is_t(f):
            fd = f.to(torch.float64)
            c3 = (self._map_glf(fd) - p["A"]) ** (p["nu"] + 1)
            return (c1 * c2 * c3 * torch.exp(-p["B"] * (fd - p["M"]))).to(f.dtype)
        c3 = (self._map_glf(f) - p["A"]) ** (p["nu"] + 1)
        return c1 * c2 * c3 * np.exp(-p["B"] * (f - p["M"]))

    def _map_inv_glf(self, mapped):
        p = self.params
        if _is_t(mapped):
            md = mapped.to(torch.float64)
            num = -torch.log(((p["K"] - p["A"]) / (md - p["A"])) ** p["nu"] - 1)
            return (num / p["B"] + p["M"]).to(mapped.dtype)
        with np.errstate(divide="ignore", invalid="ignore"):
            num = -np.log(((p["K"] - p["A"]) / (mapped - p["A"])) ** p["nu"] - 1)
        return num / p["B"] + p["M"]

    # ---- linear / identity (response.py:233-253) --------------------------
    def _map_lin(self, f):
        return self.params["M"] * f + self.params["C"]

    def _dmapdf_lin(self, f, use_cached_result: bool = False):
        if _is_t(f):
            return torch.ones_like(f) * self.params["M"]
        return np.ones_like(f) * self.params["M"]

    def _map_inv_lin(self, mapped):
        return (mapped - self.params["C"]) / self.params["M"]

    def _map_id(self, f):
        return f

    def _dmapdf_id(self, f, use_cached_result: bool = False):
        if _is_t(f):
            return torch.ones_like(f)
        return np.ones_like(f)

    def _map_inv_id(self, mapped):
        return mapped

    # ---- interpolation (response.py:256-291) ------------------------------
    def _map_interp(self, f):
        if _is_t(f):
            return _interp_t(f, self.interp_f_0, self.interp_map_0, self.interp_map_0[0], self.interp_map_0[-1])
        return np.interp(f, self.interp_f_0, self.interp_map_0, left=self.interp_map_0[0],
                         right=self.interp_map_0[-1])

    def _dmapdf_interp(self, f):
        if _is_t(f):
            return _interp_t(f, self.interp_f_0, self.interp_dmapdf_0, self.interp_dmapdf_0[0],
                             self.interp_dmapdf_0[-1])
        return np.interp(f, self.interp_f_0, self.interp_dmapdf_0, left=self.interp_dmapdf_0[0],
                         right=self.interp_dmapdf_0[-1])

    def _map_inv_interp(self, mapped):
        if _is_t(mapped):
            return _interp_t(mapped, self.interp_map_0, self.interp_f_0, self.interp_f_0[0], self.interp_f_0[-1])
        return np.interp(mapped, self.interp_map_0, self.interp_f_0, left=self.interp_f_0[0],
                         right=self.interp_f_0[-1])

    # ---- utilities ---------------------------------------------------------
    def checkResponseTarget(self, f_T) -> bool:
        """response.py:363-396."""
        f_T = f_T.detach().cpu().numpy() if _is_t(f_T) else np.asarray(f_T)
        f_T_min, f_T_max = np.amin(f_T), np.amax(f_T)
        validity = True
        if self.form == "gen_log_fun" and f_T_max > self.params["K"]:
            warnings.warn("Maximum response target is greater than right asymptotic value of response function.")
            validity = False
        m0 = self.map(np.float64(0))
        if (f_T_min < m0) and not np.isclose(f_T_min, m0, atol=0.01 * f_T_max):
            warnings.warn("Minimum response target is lower than response at zero optical dose.")
            validity = False
        if np.isinf(f_T).any():
            warnings.warn("Response target contains infinite value(s).")
            validity = False
        if np.isnan(f_T).any():
            warnings.warn("Response target contains nan value(s).")
            validity = False
        return validity

    # ---- plots (response.py:294-361) ---------------------------------------
    def _plot(self, fn, ylabel, fig, ax, lb, ub, n_pts, block, plot_kwargs, xlabel="Optical dose"):
        import matplotlib.pyplot as plt
        f = np.linspace(lb, ub, n_pts)
        if ax is None:
            fig, ax = plt.subplots()
        ax.plot(f, fn(f), **plot_kwargs)
        ax.set_xlabel(xlabel)
        ax.set_ylabel(ylabel)
        if block:
            if "label" in plot_kwargs:
                ax.legend()
            plt.show(block=True)
        return fig, ax

    def plotMap(self, fig=None, ax=None, lb=0, ub=1, n_pts=512, block=True, **plot_kwargs):
        return self._plot(self.map, "Material response (mapped dose)", fig, ax, lb, ub, n_pts, block, plot_kwargs)

    def plotDmapDf(self, fig=None, ax=None, lb=0, ub=1, n_pts=512, block=True, **plot_kwargs):
        return self._plot(self.dmapdf, "Derivative of the response", fig, ax, lb, ub, n_pts, block, plot_kwargs)

    def plotMapInv(self, fig=None, ax=None, lb=0, ub=1, n_pts=512, block=True, **plot_kwargs):
        return self._plot(self.map_inv, "Optical dose", fig, ax, lb, ub, n_pts, block, plot_kwargs,
                          xlabel="Material response (mapped dose)")

    def vamx_struct(self, device):
        """Pack the model for the fused BCLP kernels (``vamx_response``)."""
        from . import _lib
        r = _lib.Response()
        if self.type == "analytical":
            if self.form == "gen_log_fun":
                p = self.params
                r.kind = _lib.RESP_GLF
                r.A, r.K, r.B, r.M, r.nu = float(p["A"]), float(p["K"]), float(p["B"]), float(p["M"]), float(p["nu"])
            elif self.form == "linear":
                r.kind = _lib.RESP_LINEAR
                r.lin_M, r.lin_C = float(self.params["M"]), float(self.params["C"])
            else:
                r.kind = _lib.RESP_IDENTITY
            return r, None
        key = str(device)
        tabs = self._dev_tables.get(key)
        if tabs is None:
            tabs = tuple(torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
                         for a in (self.interp_f_0, self.interp_map_0, self.interp_dmapdf_0))
            self._dev_tables[key] = tabs
        r.kind = _lib.RESP_TABLE
        r.n_pts = int(tabs[0].numel())
        r.table_f, r.table_map, r.table_dmap = (t.data_ptr() for t in tabs)
        return r, tabs

    def __repr__(self):
        return str(self.params)
