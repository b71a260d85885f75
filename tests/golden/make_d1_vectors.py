"""Write golden vectors for the D1 parallel-beam projector.

Run with the interpreter that has scikit-image 0.18.3 (the routine the
reference's CPU projector calls for its forward rotation):

    /opt/conda/bin/python3.9 tests/golden/make_d1_vectors.py

The forward loop below is the reference's per-(z, angle) rotate-and-sum
(``vamtoolbox/projector/Projector3DParallel.py:129-156``) calling the library
routine itself (``skimage.transform.warp(..., clip=True)``); the backward loop
is the reference's per-(z, angle) ``np.interp`` gather followed by the
inscribed-circle clip (``Projector3DParallel.py:175-207``,
``vamtoolbox/util/data.py:14-40``).  Only library routines are executed; the
output is data (inputs and expected outputs) in ``d1_vectors.npz``.
"""

import os
from functools import partial

import numpy as np
from skimage.transform import warp


def ref_forward(target, angles):
    n = target.shape[0]
    center = n // 2
    out = np.zeros((n, angles.shape[0], target.shape[2]))
    for z_i in range(target.shape[2]):
        for i, angle in enumerate(np.deg2rad(angles)):
            cos_a, sin_a = np.cos(angle), np.sin(angle)
            R = np.array([[cos_a, sin_a, -center * (cos_a + sin_a - 1)],
                          [-sin_a, cos_a, -center * (cos_a - sin_a - 1)],
                          [0, 0, 1]])
            rotated = warp(target[:, :, z_i], R, clip=True)
            out[:, i, z_i] = rotated.sum(0)
    return out


def clip_circle(x):
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, x.shape[0]), np.linspace(-1, 1, x.shape[1]))
    mask = np.broadcast_to((circle_x ** 2 + circle_y ** 2 > 1)[..., np.newaxis], x.shape)
    x = x.copy()
    x[mask] = 0
    return x


def ref_backward(sino, angles):
    n = sino.shape[0]
    radius = n // 2
    y, x = np.mgrid[:n, :n] - radius
    proj_t = np.arange(n) - n // 2
    out = np.zeros((n, n, sino.shape[2]))
    for z_i in range(sino.shape[2]):
        rec = np.zeros((n, n))
        for curr_proj, angle in zip(sino[:, :, z_i].T, np.deg2rad(angles)):
            cos_a, sin_a = np.cos(angle), np.sin(angle)
            t = x * cos_a - y * sin_a
            rec += partial(np.interp, xp=proj_t, fp=curr_proj, left=0, right=0)(t)
        out[:, :, z_i] = rec
    return clip_circle(out)


def main():
    rng = np.random.default_rng(20240601)
    cases = {}
    angles_a = np.array([0.0, 7.5, 30.0, 45.0, 90.0, 133.0, 180.0, 251.0, 270.0, 359.0])
    angles_b = np.arange(0.0, 360.0, 24.0)
    for n in (16, 15):
        yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
        ell = ((xx / (0.35 * n)) ** 2 + (yy / (0.22 * n)) ** 2 <= 1).astype(np.float64)
        vols = {
            "pos": rng.uniform(0.2, 1.0, (n, n, 2)),          # no zeros: exercises the warp clip
            "signed": rng.normal(size=(n, n, 2)),
            "ellipse": np.stack([ell, np.roll(ell, 2, axis=1)], axis=2),
        }
        for name, vol in vols.items():
            for aname, ang in (("a", angles_a), ("b", angles_b)):
                key = "n%d_%s_%s" % (n, name, aname)
                cases[key + "_vol"] = vol
                cases[key + "_ang"] = ang
                cases[key + "_fwd"] = ref_forward(vol, ang)
                sino = rng.normal(size=(n, ang.size, 2))
                cases[key + "_sino"] = sino
                cases[key + "_bwd"] = ref_backward(sino, ang)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "d1_vectors.npz"), **cases)
    print("wrote", len(cases), "arrays")


if __name__ == "__main__":
    main()
