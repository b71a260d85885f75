"""Float64 restatement of the fan-beam rebinning (vamtoolbox/geometry.py:736-944).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

A parallel-beam sinogram is resampled where a non-telecentric projector's
rays, refracted at the air/vial wall, actually sample Radon space (Orth et
al., Opt. Express 29, 11037 (2021)), then scaled by the change of the
differential area and by the inverse Fresnel transmission.  The resampling is
scipy's ``interpolate.interpn(method="linear", bounds_error=False,
fill_value=0)``, the routine the reference calls.
"""

from __future__ import annotations

import numpy as np
from scipy import interpolate


def sampling(angles, vial_width, n_u, n_write, throw_ratio):
    """(x_samp, theta_samp, dxv tiled, 1/T tiled, xp) of geometry.py:823-918."""
    angles = np.asarray(angles, dtype=np.float64)
    na = angles.size
    w = int(vial_width)
    tr = throw_ratio * n_u                                     # throw ratio in pixels
    Rv = (w / 2) * np.sqrt(1 + (w / (2 * tr)) ** 2)            # true vial radius in pixels
    xp = np.linspace(-w / 2, w / 2, w)                         # projector pixel coordinate
    phi = np.arctan(xp / tr)                                   # divergence of pixel xp
    xps = (xp - xp * np.sqrt(1 - (1 + (xp / tr) ** 2) * (1 - (Rv / tr) ** 2))) / (1 + (xp / tr) ** 2)
    theta10 = np.arcsin(xps / Rv)                              # wall normal at the hit point
    thetai = np.arcsin(xps / Rv) + phi                         # incidence angle
    thetat = np.arcsin((1.0 / n_write) * np.sin(thetai))       # Snell, n_air = 1
    thetav = theta10 - thetat                                  # ray deviation inside the vial
    Ts = 1 - ((np.cos(thetai) - n_write * np.cos(thetat)) / (np.cos(thetai) + n_write * np.cos(thetat))) ** 2
    Tp = 1 - ((np.cos(thetat) - n_write * np.cos(thetai)) / (np.cos(thetat) + n_write * np.cos(thetai))) ** 2
    T = (Ts + Tp) / 2
    T[T < 0] = 0
    xv = xps * np.cos(thetav) - np.sin(thetav) * np.sqrt(Rv ** 2 - xps ** 2)
    dxv = np.gradient(xv)
    dxv[T < 0] = 0
    dxv[dxv < 0] = 0
    theta_samp, x_samp = np.meshgrid(angles, xv)               # (w, na)
    theta_samp = theta_samp + np.transpose(np.tile((180 / np.pi) * thetav, (na, 1)))
    lo, hi = angles[0], angles[-1]
    step = abs(angles[1] - angles[0])
    theta_samp[theta_samp > hi] = theta_samp[theta_samp > hi] - 360
    theta_samp[theta_samp < lo] = 360 + theta_samp[theta_samp < lo]
    theta_samp[theta_samp > hi - step] = lo
    theta_samp[(theta_samp <= hi - step) & (theta_samp > hi)] = hi
    return x_samp, theta_samp, np.transpose(np.tile(dxv, (na, 1))), np.transpose(np.tile(1 / T, (na, 1))), xp


def rebin_fan_beam(sino, angles, vial_width, n_screen, n_write, throw_ratio):
    """Rebinned sinogram (vial_width, n_angles, nz) of geometry.py:920-944."""
    sino = np.asarray(sino, dtype=np.float64)
    w = int(vial_width)
    x_samp, theta_samp, dxv, t_inv, xp = sampling(angles, w, n_screen[0], n_write, throw_ratio)
    pd = int((w - sino.shape[0]) / 2)
    if pd > 0:
        sino = np.pad(sino, ((pd, pd), (0, 0), (0, 0)), mode="constant")
    if sino.shape[0] < w:
        sino = np.pad(sino, ((0, 1), (0, 0), (0, 0)), mode="constant")
    out = np.zeros_like(sino)
    for z in range(sino.shape[2]):
        b = interpolate.interpn((xp, np.asarray(angles, dtype=np.float64)), sino[..., z], (x_samp, theta_samp),
                                method="linear", bounds_error=False, fill_value=0)
        out[..., z] = t_inv * (b * dxv)
    return out
