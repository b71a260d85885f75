"""Write golden vectors for the D1 projector with occlusion shadowing.

Run with the interpreter that has scikit-image 0.18.3:

    /opt/conda/bin/python3.9 tests/golden/make_occlusion_vectors.py

The loops restate the reference's occlusion path of the CPU projector, which
it takes whenever ``proj_geo.attenuation_field`` is set
(``vamtoolbox/projector/Projector3DParallel.py``):

* insert = attenuation_field > 0 (73-74).  The reference casts it to int;
  scikit-image 0.18.3's int64-to-float conversion maps 0 to 5.4e-20, so
  every in-grid sample would count as occluding and each column's first hit
  would be the grid's first row.  These vectors pass the insert as float
  0.0 / 1.0, which gives the shadow of the occluding voxels themselves --
  the binary shadow the reference's loops are written for;
* occlusion sinogram (77-111): per slice and angle, the insert is rotated
  with ``skimage.transform.warp(..., clip=True)`` and, per detector column,
  the smallest row coordinate ``y = r - N//2`` with a positive rotated value
  is kept (``np.nanmin``; NaN when the column never meets the insert);
* forward (129-156): the rotated slice is zeroed where ``y > occ`` for that
  column, but only when the angle's occlusion column holds at least one
  finite non-zero entry (``count_nonzero - isnan != 0``), then summed;
* backward (175-216): each angle's ``np.interp`` gather is zeroed where
  ``s > floor(np.interp(t, proj_t, occ_column, left=nan, right=nan))``,
  then the inscribed-circle clip.

Only library routines (scikit-image's ``warp``, NumPy's ``interp`` and
``nanmin``) are executed; the output is data in ``d1_occlusion_vectors.npz``.
"""

import os
import warnings

import numpy as np
from skimage.transform import warp


def rot_matrix(angle, center):
    cos_a, sin_a = np.cos(angle), np.sin(angle)
    return np.array([[cos_a, sin_a, -center * (cos_a + sin_a - 1)],
                     [-sin_a, cos_a, -center * (cos_a - sin_a - 1)],
                     [0, 0, 1]])


def ref_occ_sinogram(insert, angles):
    n, _, nz = insert.shape
    center = n // 2
    y = (np.mgrid[:n, :n] - n // 2)[0]
    occ = np.zeros((n, angles.shape[0], nz))
    for z_i in range(nz):
        for i, angle in enumerate(np.deg2rad(angles)):
            rotated = warp(insert[:, :, z_i], rot_matrix(angle, center), clip=True)
            s_occ = np.where(rotated > 0, y, np.nan)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)   # all-NaN columns
                occ[:, i, z_i] = np.nanmin(s_occ, axis=0)
    return occ


def ref_forward(target, angles, occ):
    n = target.shape[0]
    center = n // 2
    y = (np.mgrid[:n, :n] - n // 2)[0]
    out = np.zeros((n, angles.shape[0], target.shape[2]))
    for z_i in range(target.shape[2]):
        for i, angle in enumerate(np.deg2rad(angles)):
            rotated = warp(target[:, :, z_i], rot_matrix(angle, center), clip=True)
            col = occ[:, i, z_i]
            if np.count_nonzero(col) - np.sum(np.isnan(col)) != 0:
                rotated = rotated * np.logical_not(y > col[np.newaxis, :])
            out[:, i, z_i] = rotated.sum(0)
    return out


def clip_circle(x):
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, x.shape[0]), np.linspace(-1, 1, x.shape[1]))
    mask = np.broadcast_to((circle_x ** 2 + circle_y ** 2 > 1)[..., np.newaxis], x.shape)
    x = x.copy()
    x[mask] = 0
    return x


def ref_backward(sino, angles, occ):
    n = sino.shape[0]
    y, x = np.mgrid[:n, :n] - n // 2
    proj_t = np.arange(n) - n // 2
    out = np.zeros((n, n, sino.shape[2]))
    for z_i in range(sino.shape[2]):
        rec = np.zeros((n, n))
        for i, (curr_proj, angle) in enumerate(zip(sino[:, :, z_i].T, np.deg2rad(angles))):
            cos_a, sin_a = np.cos(angle), np.sin(angle)
            t = x * cos_a - y * sin_a
            s = x * sin_a + y * cos_a
            back = np.interp(t, proj_t, curr_proj, left=0, right=0)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)   # comparisons with NaN
                shadow = s > np.floor(np.interp(t, proj_t, occ[:, i, z_i], left=np.nan, right=np.nan))
            rec += back * np.logical_not(shadow)
        out[:, :, z_i] = rec
    return clip_circle(out)


def main():
    rng = np.random.default_rng(20240602)
    cases = {}
    angles_a = np.array([0.0, 7.5, 30.0, 45.0, 90.0, 133.0, 180.0, 251.0, 270.0, 359.0])
    angles_b = np.arange(0.0, 360.0, 24.0)
    for n in (16, 15):
        yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
        rod = ((xx - 0.15 * n) ** 2 + (yy + 0.1 * n) ** 2 <= (0.12 * n) ** 2)
        bar = (np.abs(xx + 0.2 * n) <= 0.08 * n) & (np.abs(yy) <= 0.3 * n)
        empty = np.zeros((n, n), dtype=bool)
        centre = np.zeros((n, n), dtype=bool)
        centre[n // 2, n // 2] = True                 # first hits at s = 0 for the central column
        fields = {
            "rod": np.stack([rod, rod | bar, empty], axis=2),
            "centre": np.stack([centre, rod], axis=2),
        }
        for name, field in fields.items():
            nz = field.shape[2]
            att = np.where(field, np.inf, 0.0)
            insert = np.where(att > 0, 1.0, 0.0)              # Projector3DParallel.py:73-74 (as float)
            vol = rng.uniform(0.2, 1.0, (n, n, nz))
            for aname, ang in (("a", angles_a), ("b", angles_b)):
                key = "n%d_%s_%s" % (n, name, aname)
                occ = ref_occ_sinogram(insert, ang)
                sino = rng.normal(size=(n, ang.size, nz))
                cases[key + "_insert"] = insert.astype(np.uint8)
                cases[key + "_ang"] = ang
                cases[key + "_occ"] = occ
                cases[key + "_vol"] = vol
                cases[key + "_fwd"] = ref_forward(vol, ang, occ)
                cases[key + "_sino"] = sino
                cases[key + "_bwd"] = ref_backward(sino, ang, occ)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "d1_occlusion_vectors.npz"), **cases)
    print("wrote", len(cases), "arrays")


if __name__ == "__main__":
    main()
