"""Pinning of the D1 oracle (CPU): golden vectors from the reference's own
third-party routines, brute-force matrices, closed forms and invariants."""

import os

import numpy as np
import pytest

import oracle.d1 as d1

GOLD = os.path.join(os.path.dirname(__file__), "golden", "d1_vectors.npz")


def test_golden_skimage_and_interp():
    """scikit-image 0.18.3 warp (forward) and numpy.interp (backward) golden vectors."""
    z = np.load(GOLD)
    keys = sorted(set(k.rsplit("_", 1)[0] for k in z.files))
    assert len(keys) == 12
    for k in keys:
        fwd = d1.forward(z[k + "_vol"], z[k + "_ang"])
        bwd = d1.backward(z[k + "_sino"], z[k + "_ang"])
        np.testing.assert_allclose(fwd, z[k + "_fwd"], rtol=1e-12, atol=1e-12, err_msg=k)
        np.testing.assert_allclose(bwd, z[k + "_bwd"], rtol=1e-12, atol=1e-12, err_msg=k)


def test_warp_clip_matters():
    """The per-slice output clip of warp(clip=True) changes positive-only slices (pins it)."""
    z = np.load(GOLD)
    k = "n16_pos_a"
    with_clip = d1.forward(z[k + "_vol"], z[k + "_ang"])
    no_clip = d1.forward(z[k + "_vol"], z[k + "_ang"], clip_output=False)
    assert np.abs(with_clip - no_clip).max() > 1e-3
    np.testing.assert_allclose(with_clip, z[k + "_fwd"], rtol=1e-12, atol=1e-12)


def _matrix(op, shape_in):
    n = int(np.prod(shape_in))
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(op(e.reshape(shape_in)).ravel())
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("kw", [{}, {"det_offset": 1.5}, {"mu": 0.1}, {"incl_deg": 9.0},
                                {"mu": 0.05, "det_offset": -2.0, "incl_deg": -6.0}])
def test_adjoints_are_transposes(kw):
    """Brute-force matrices: forward_adjoint == F^T and backward_adjoint == B^T."""
    n, nz = 7, 3
    ang = np.array([0.0, 23.0, 90.0, 141.0, 270.0])
    F = _matrix(lambda v: d1.forward(v, ang, clip_output=False, **kw), (n, n, nz))
    B = _matrix(lambda s: d1.backward(s, ang, **kw), (n, len(ang), nz))
    FT = _matrix(lambda s: d1.forward_adjoint(s, ang, **kw), (n, len(ang), nz))
    BT = _matrix(lambda v: d1.backward_adjoint(v, ang, **kw), (n, n, nz))
    np.testing.assert_allclose(FT, F.T, atol=1e-12)
    np.testing.assert_allclose(BT, B.T, atol=1e-12)


def test_forward_brute_force_scalar_loop():
    """Independent scalar restatement of one rotated column sum (no vectorisation)."""
    rng = np.random.default_rng(3)
    n = 9
    img = rng.random((n, n))
    ang = 37.0
    c = n // 2
    th = np.deg2rad(ang)
    co, sn = np.cos(th), np.sin(th)
    ref = np.zeros(n)
    for k in range(n):
        acc = 0.0
        for r in range(n):
            rho = c - sn * (k - c) + co * (r - c)
            gam = c + co * (k - c) + sn * (r - c)
            r0, c0 = int(np.floor(rho)), int(np.floor(gam))
            val = 0.0
            for rr, wr in ((r0, 1 - (rho - r0)), (r0 + 1, rho - r0)):
                for cc, wc in ((c0, 1 - (gam - c0)), (c0 + 1, gam - c0)):
                    if 0 <= rr < n and 0 <= cc < n:
                        val += wr * wc * img[rr, cc]
            acc += val
        ref[k] = acc
    out = d1.forward(img[:, :, None], [ang], clip_output=False)[:, 0, 0]
    np.testing.assert_allclose(out, ref, atol=1e-12)


def test_rotation_by_90_degrees_shifts_angles():
    """f rotated by 90 degrees <-> sinogram shifted by 90 one-degree angle samples."""
    rng = np.random.default_rng(5)
    n = 16
    f = np.zeros((n, n, 1))
    f[3:13, 4:12, 0] = rng.random((10, 8))
    ang = np.arange(0.0, 360.0, 1.0)
    s = d1.forward(f, ang, clip_output=False)
    # rotating the image content by +90 degrees about the centre c = n//2
    fr = np.zeros_like(f)
    c = n // 2
    for r in range(n):
        for k in range(n):
            rr, kk = c + (k - c), c - (r - c)
            if 0 <= rr < n and 0 <= kk < n:
                fr[rr, kk, 0] = f[r, k, 0]
    sr = d1.forward(fr, ang, clip_output=False)
    shifted = np.roll(s, -90, axis=1)
    assert (np.allclose(sr, shifted, atol=1e-9) or np.allclose(sr, np.roll(s, 90, axis=1), atol=1e-9))


def test_integer_offset_is_index_shift():
    rng = np.random.default_rng(6)
    n = 16
    f = rng.random((n, n, 2))
    ang = np.arange(0.0, 180.0, 15.0)
    s0 = d1.forward(f, ang)
    s3 = d1.forward(f, ang, det_offset=3.0)
    np.testing.assert_allclose(s3[3:], s0[:-3], atol=1e-12)
    b = rng.random((n, len(ang), 2))
    b3 = np.zeros_like(b)
    b3[3:] = b[:-3]
    # voxels whose detector coordinate stays inside both shifted detectors (|t| < 4)
    c = n // 2
    y, x = np.mgrid[:n, :n] - c
    inner = x ** 2 + y ** 2 < 4 ** 2
    np.testing.assert_allclose(d1.backward(b3, ang, det_offset=3.0)[inner], d1.backward(b, ang)[inner], atol=1e-12)


def test_mu_zero_and_phi_zero_reduce_exactly():
    rng = np.random.default_rng(7)
    f = rng.random((12, 12, 3))
    ang = np.arange(0.0, 360.0, 30.0)
    base = d1.forward(f, ang)
    assert np.array_equal(d1.forward(f, ang, mu=0.0, incl_deg=0.0), base)


def test_absorption_closed_form_uniform_cylinder():
    """Per-ray exp absorption in a uniform medium: b(t) = int_{-L}^{L} e^{-mu (L - s)} ds,
    L = sqrt(R^2 - t^2); with a constant slice the discrete sum approaches the closed
    form (1 - e^{-2 mu L}) / mu (getExpabs' geometry)."""
    n = 129
    f = np.ones((n, n, 1))
    mu, R = 0.02, 40.0
    f = d1.clip_to_circle(f)
    # restrict the object to the cylinder of radius R so the ray sum sees a uniform medium
    yy, xx = np.mgrid[:n, :n] - n // 2
    f[..., 0] *= (xx ** 2 + yy ** 2 <= R ** 2)
    s = d1.forward(f, [0.0], mu=mu, radius=R, clip_output=False)[:, 0, 0]
    for t in (0, 10, 25):
        L = np.sqrt(R ** 2 - t ** 2)
        closed = (1 - np.exp(-2 * mu * L)) / mu
        assert abs(s[n // 2 + t] - closed) / closed < 0.03
    # orientation: the light enters at s = +sqrt(R^2 - t^2) and runs toward -s
    # (at 0 degrees output row r samples input row r, s = r - n // 2), so a
    # voxel at s = +20 is dimmed by exp(-mu (L - 20)), one at s = -20 by
    # exp(-mu (L + 20)); getExpabs' convention, astra3Dabsorption.py:130-139
    c = n // 2
    one = np.zeros((n, n, 1))
    one[c + 20, c, 0] = 1.0
    two = np.zeros((n, n, 1))
    two[c - 20, c, 0] = 1.0
    a = d1.forward(one, [0.0], mu=mu, radius=R, clip_output=False)[c, 0, 0]
    b = d1.forward(two, [0.0], mu=mu, radius=R, clip_output=False)[c, 0, 0]
    assert abs(a - np.exp(-mu * (R - 20))) < 1e-12 and abs(b - np.exp(-mu * (R + 20))) < 1e-12


def test_constant_sinogram_backprojection():
    """BP of b == 1: each voxel inside the circle receives the number of angles
    whose detector coordinate lies inside the detector."""
    n = 32
    ang = np.arange(0.0, 180.0, 3.0)
    x = d1.backward(np.ones((n, len(ang), 1)), ang)[:, :, 0]
    c = n // 2
    y, xx = np.mgrid[:n, :n] - c
    expect = np.zeros((n, n))
    for a in np.deg2rad(ang):
        t = xx * np.cos(a) - y * np.sin(a)
        expect += ((t >= -c) & (t <= n - 1 - c))
    expect[d1.outside_circle_mask(n, n)] = 0
    np.testing.assert_allclose(x, expect, atol=1e-12)
    assert np.isclose(x[c, c], len(ang))


def test_mass_conservation_axis_aligned():
    """At 0 and 90 degrees the rotated grid is the pixel grid: detector sum == image sum."""
    rng = np.random.default_rng(9)
    n = 20
    f = np.zeros((n, n, 2))
    f[5:15, 5:15] = rng.random((10, 10, 2))
    s = d1.forward(f, [0.0, 90.0, 180.0], clip_output=False)
    for i in range(3):
        np.testing.assert_allclose(s[:, i, :].sum(axis=0), f.sum(axis=(0, 1)), rtol=1e-9)


def test_clip_to_circle_counts():
    m = d1.outside_circle_mask(256, 256)
    # symmetric and strictly inside-out
    assert np.array_equal(m, m.T) and np.array_equal(m, m[::-1]) and m[0, 0] and not m[128, 128]
    assert m.sum() == 256 * 256 - d1.inside_circle_mask(256, 256).sum()
    # points exactly on the circle are kept (util/data.py:29-33 zeroes r^2 > 1):
    # at odd n the four axis end points lie on it
    m17 = d1.outside_circle_mask(17, 17)
    assert not m17[0, 8] and not m17[8, 0] and not m17[16, 8] and not m17[8, 16] and m17[0, 0]


# --------------------------------------------------------------------------
# occlusion shadowing (Projector3DParallel.py:77-111, 147-153, 192-199, 209-216)
# --------------------------------------------------------------------------
def _occ_cases():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "d1_occlusion_vectors.npz"))
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files})
    return g, keys


def test_occlusion_sinogram_matches_golden():
    g, keys = _occ_cases()
    assert len(keys) == 8
    for k in keys:
        occ = d1.occlusion_sinogram(g[k + "_insert"], g[k + "_ang"])
        ref = g[k + "_occ"]
        assert np.array_equal(np.isnan(occ), np.isnan(ref)), k
        assert np.array_equal(occ[~np.isnan(occ)], ref[~np.isnan(ref)]), k


def test_occluded_forward_backward_match_golden():
    g, keys = _occ_cases()
    for k in keys:
        ang, occ = g[k + "_ang"], g[k + "_occ"]
        f = d1.forward(g[k + "_vol"], ang, occ=occ)
        b = d1.backward(g[k + "_sino"], ang, occ=occ)
        assert np.max(np.abs(f - g[k + "_fwd"])) <= 1e-12 * max(1.0, np.abs(g[k + "_fwd"]).max()), k
        assert np.max(np.abs(b - g[k + "_bwd"])) <= 1e-12 * max(1.0, np.abs(g[k + "_bwd"]).max()), k


def test_occlusion_without_occluder_is_neutral():
    rng = np.random.default_rng(5)
    n, nz = 13, 2
    ang = np.arange(0.0, 360.0, 17.0)
    occ = d1.occlusion_sinogram(np.zeros((n, n, nz)), ang)
    assert np.isnan(occ).all()
    vol, sino = rng.uniform(0.1, 1, (n, n, nz)), rng.normal(size=(n, ang.size, nz))
    assert np.array_equal(d1.forward(vol, ang, occ=occ), d1.forward(vol, ang))
    assert np.array_equal(d1.backward(sino, ang, occ=occ), d1.backward(sino, ang))


def test_occlusion_first_hit_brute_force():
    """At 0 degrees the rotation is the identity: the first hit of column k is
    the first occluding row r (minus N//2), found here by a plain scan."""
    rng = np.random.default_rng(8)
    n = 11
    ins = (rng.uniform(size=(n, n, 1)) > 0.85).astype(np.uint8)
    occ = d1.occlusion_sinogram(ins, [0.0])[:, 0, 0]
    for k in range(n):
        rows = [r for r in range(n) if ins[r, k, 0]]
        if rows:
            assert occ[k] == rows[0] - n // 2
        else:
            assert np.isnan(occ[k])
