"""Independent checks of the ray-tracer oracle (``oracle/raytrace.py``) (CPU).

* ``GridMedium.sample`` against ``torch.nn.functional.grid_sample`` called as
  the reference calls it (``vamtoolbox/medium.py:161-211``: field permuted to
  (Z, Y, X), positions times 2 / span, ``align_corners=True``, "border" or
  "zeros" padding), on random fields and points inside, on and beyond the
  grid's faces, in 3-D and 2-D.
* ``central_difference`` (``medium.py:224-273``): exact on quadratic fields
  in the interior, zero on the boundary faces, no z component in 2-D.
* The three parametrisations (``pyTorchRayTrace.py:692-735``) trace the same
  geometric ray when v0 = n(x0) t0 (the ray equation d/ds (n dx/ds) = grad n
  in arc length, optical-path and canonical time), with RK4 and with
  symplectic Euler.
* Straight rays (vacuum): a one-voxel f integrated along a ray, and the
  dose one ray deposits in that voxel, equal the closed form
  sum_k (h / dx) exp(-A_k) hat(x_k - x_voxel), with x_k = x0 + (k + 1) h v
  the positions after each step and A_k = sum_{m <= k} alpha(x0 + m h v) h
  the attenuance collected at the step starts (``pyTorchRayTrace.py:763-774,
  939-1143, 1253-1260``).
* The exit step count of a vacuum ray, checked every 10 steps starting at
  step 0 (``pyTorchRayTrace.py:348-363, 1391-1420``), equals its closed form.
* ``setup_rays`` (``pyTorchRayTrace.py:1515-1666``): each ray starts at
  distance D along -v0 with lateral coordinates (G0, G2) in the orthonormal
  frame of its angle and inclination; unit v0; ray energy 1 / rho^2 (3-D).
* ``luneburg_n`` (``medium.py:486-549``): n(0) = sqrt(2) n_sur, n(R) = n_sur,
  grad n radial with the analytic derivative.
"""

import numpy as np
import pytest

import oracle.raytrace as rt


def _grid_sample_ref(field, cv, x, padding):
    """The reference's grid_sample call (medium.py:161-211) on a (nx, ny, nz, C) field."""
    torch = pytest.importorskip("torch")
    xv, yv, zv = cv
    span = np.array([xv.max() - xv.min(), yv.max() - yv.min(), zv.max() - zv.min()])
    with np.errstate(divide="ignore"):
        norm = 2.0 / span
    norm[np.isinf(norm)] = 0.0
    inp = torch.tensor(np.transpose(field, (3, 2, 1, 0))[None], dtype=torch.float64)   # [1, C, Z, Y, X]
    grid = torch.tensor((x * norm)[None, None, None], dtype=torch.float64)
    out = torch.nn.functional.grid_sample(inp, grid, mode="bilinear", padding_mode=padding, align_corners=True)
    return out[0, :, 0, 0, :].numpy().T


def _points(rng, cv, m):
    lo = np.array([c.min() for c in cv])
    hi = np.array([c.max() for c in cv])
    ext = hi - lo
    x = lo + rng.uniform(-0.2, 1.2, (m, 3)) * ext           # inside and beyond the faces
    x[: m // 4] = lo + rng.integers(0, 2, (m // 4, 3)) * ext  # exactly on corners
    x[m // 4: m // 2, 0] = hi[0]                              # exactly on a face
    return x


def check_grid_sample_parity():
    rng = np.random.default_rng(0)
    for shape in ((7, 9, 5), (6, 8, 1)):
        cv = rt.coord_vec(shape[0], shape[1], 0 if shape[2] == 1 else shape[2])
        n = 1.5 + 0.1 * rng.standard_normal(shape)
        med = rt.GridMedium(cv, n)
        x = _points(rng, cv, 400)
        if shape[2] == 1:
            x[:, 2] = rng.uniform(-1, 1, 400) * 1e-3           # any z: 2-D normalisation is 0
        for padding in ("border", "zeros"):
            ours = med.sample(x, padding=padding)
            ref = _grid_sample_ref(med.field, cv, x, padding)
            np.testing.assert_allclose(ours, ref, rtol=1e-12, atol=1e-12)
        a = rng.uniform(0, 1, shape)
        ours = rt.att_alpha(x, "grid", grid=a, medium=med)
        ref = _grid_sample_ref(a[..., None], cv, x, "zeros")[:, 0]
        np.testing.assert_allclose(ours, ref, rtol=1e-12, atol=1e-12)


def test_grid_sample_parity():
    check_grid_sample_parity()


def check_central_difference():
    cv = rt.coord_vec(9, 7, 6)
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    vox = np.array([cv[0][1] - cv[0][0], cv[1][1] - cv[1][0], cv[2][1] - cv[2][0]])
    n = 1.2 + 30 * X ** 2 - 17 * Y ** 2 + 11 * Z ** 2 + 40 * X * Y - 23 * Y * Z + 5 * X
    g = rt.central_difference(n, vox)
    ex = np.stack([60 * X + 40 * Y + 5, -34 * Y + 40 * X - 23 * Z, 22 * Z - 23 * Y], axis=-1)
    np.testing.assert_allclose(g[1:-1, :, :, 0], ex[1:-1, :, :, 0], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(g[:, 1:-1, :, 1], ex[:, 1:-1, :, 1], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(g[:, :, 1:-1, 2], ex[:, :, 1:-1, 2], rtol=1e-9, atol=1e-9)
    for ax in range(3):
        face = [slice(None)] * 3
        face[ax] = 0
        assert np.all(g[tuple(face) + (ax,)] == 0)
        face[ax] = -1
        assert np.all(g[tuple(face) + (ax,)] == 0)
    cv2 = rt.coord_vec(9, 7, 0)
    X2, Y2 = np.meshgrid(cv2[0], cv2[1], indexing="ij")
    g2 = rt.central_difference((X2 ** 2 + Y2)[..., None], np.array([vox[0], vox[1], np.inf]))
    assert np.all(g2[..., 2] == 0)
    np.testing.assert_allclose(g2[1:-1, :, 0, 0], 2 * X2[1:-1], rtol=1e-9, atol=1e-12)


def test_central_difference():
    check_central_difference()


def _affine_medium(n=24):
    cv = rt.coord_vec(n, n, n)
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    a, b, c = 15.0, -9.0, 6.0
    nf = 1.5 + a * X + b * Y + c * Z
    g = np.zeros(nf.shape + (3,))
    g[..., 0], g[..., 1], g[..., 2] = a, b, c
    return cv, rt.GridMedium(cv, nf, grad=g)


def _plane_crossing(tr, fine, x, v, plane_x, max_steps=200000):
    """Trace (x, v) with ``tr`` until x[0] falls below plane_x, re-trace the
    last step with the 256 times finer ``fine`` and return (y, z) at the
    crossing by linear interpolation (error ~ (h / 256)^2 curvature)."""
    for _ in range(max_steps):
        xn, vn = tr.advance(x, v)
        if xn[0, 0] <= plane_x:
            for _ in range(512):
                xn, vn = fine.advance(x, v)
                if xn[0, 0] <= plane_x:
                    s = (x[0, 0] - plane_x) / (x[0, 0] - xn[0, 0])
                    return x[0, 1:] + s * (xn[0, 1:] - x[0, 1:])
                x, v = xn, vn
            raise AssertionError("fine re-trace missed the plane")
        x, v = xn, vn
    raise AssertionError("ray never reached the plane")


def check_parametrisations_same_geometry():
    cv, med = _affine_medium()
    vox = med.voxel[0]
    x0 = np.array([[cv[0].max() * 0.9, 0.3 * cv[1].max(), -0.2 * cv[2].max()]])
    t0 = np.array([-1.0, 0.25, 0.1])
    t0 /= np.linalg.norm(t0)
    n0 = med.sample(x0)[0, 0]
    plane = cv[0].min() * 0.8
    for solver, h, tol in (("rk4", vox / 4, 1e-9), ("forward_symplectic_euler", vox / 64, 2e-3)):
        ends = {}
        for param, speed in (("canonical", n0), ("physical_path_length", 1.0), ("optical_path_length", 1.0 / n0)):
            # same arc length per step for every parametrisation: h_param = h_s / |dx/dtau|
            tr = rt.Tracer(med, param=param, solver=solver, step=h / speed)
            fine = rt.Tracer(med, param=param, solver=solver, step=h / speed / 256)
            ends[param] = _plane_crossing(tr, fine, x0.copy(), (n0 * t0)[None], plane)
        span = cv[1].max() - cv[1].min()
        ref = ends["physical_path_length"]
        for param in ("canonical", "optical_path_length"):
            assert np.max(np.abs(ends[param] - ref)) < tol * span, (solver, param, ends[param], ref)
        # the field bends the ray: the crossing is far from the straight line's
        straight = x0[0, 1:] + (plane - x0[0, 0]) / t0[0] * t0[1:]
        assert np.max(np.abs(ref - straight)) > 0.02 * span


def test_parametrisations_same_geometry():
    check_parametrisations_same_geometry()


def _vacuum(n=16, nz=5):
    cv = rt.coord_vec(n, n, nz)
    return cv, rt.GridMedium(cv, np.ones((n, n, nz)))


def _hat(u):
    return np.maximum(0.0, 1.0 - np.abs(u))


def _closed_form_straight(med, x0, v0, h, voxel_idx, alpha, radius, e0=1.0):
    """sum_k (h / dx) E_k hat(idx(x_k) - voxel) for a straight ray, E_k = e0 exp(-A_k)."""
    vox = med.voxel
    k = np.arange(int(4 * med.n.shape[0] / (h / vox[0])) + 10)
    xs = x0[None] + (k[:, None] + 1) * h * v0[None]              # position after step k
    xo = x0[None] + k[:, None] * h * v0[None]                     # position at the start of step k
    inside = np.linalg.norm(xo[:, :2], axis=1) < radius
    A = np.cumsum(np.where(inside, alpha, 0.0) * h)
    idx = (xs + med.span / 2) / vox
    w = _hat(idx[:, 0] - voxel_idx[0]) * _hat(idx[:, 1] - voxel_idx[1]) * _hat(idx[:, 2] - voxel_idx[2])
    return float(np.sum((h / vox[0]) * e0 * np.exp(-A) * w))


def check_straight_ray_attenuation_closed_form():
    n, nz = 16, 5
    cv, med = _vacuum(n, nz)
    ang = [0.0, 90.0, 33.0]
    x0, v0, shape, e0 = rt.setup_rays(cv, ang)
    R = 0.6 * cv[0].max()
    alpha = 30.0                     # per unit length: 0.06 per voxel
    tr = rt.Tracer(med, param="physical_path_length", solver="forward_symplectic_euler",
                   att=("cylinder", alpha, R, None))
    absg = np.full((n, n, nz), 2.0)
    tr_abs = rt.Tracer(med, param="physical_path_length", solver="forward_symplectic_euler",
                       att=("cylinder", alpha, R, None), abs_grid=absg)
    rays = np.arange(x0.shape[0]).reshape(shape)
    checked = 0
    for (t, a, z) in ((5, 0, 2), (9, 0, 1), (7, 1, 3), (8, 2, 2)):
        r = rays[t, a, z]
        for vi in ((3, t, z), (8, t, z), (12, t, z), (t, 4, z), (6, 9, 2)):
            f = np.zeros((n, n, nz))
            f[vi] = 1.0
            expect = _closed_form_straight(med, x0[r], v0[r], tr.h, np.array(vi, dtype=float), alpha, R, e0)
            got = tr.run(x0[r:r + 1], v0[r:r + 1], e0, f=f)[0]
            assert abs(got - expect) <= 1e-12 * max(1.0, abs(expect)), (t, a, z, vi, got, expect)
            got_abs = tr_abs.run(x0[r:r + 1], v0[r:r + 1], e0, f=f)[0]
            assert abs(got_abs - 2.0 * expect) <= 1e-12 * max(1.0, abs(expect))
            dose = tr.run(x0[r:r + 1], v0[r:r + 1], e0, g=np.array([0.7]))
            assert abs(dose[vi] - 0.7 * expect / e0) <= 1e-12 * max(1.0, abs(expect))
            checked += expect > 0.1
    assert checked >= 6


def test_straight_ray_attenuation_closed_form():
    check_straight_ray_attenuation_closed_form()


def _exit_steps_closed_form(med, x0, v0, h, every, max_steps):
    lo = med.vmin - med.voxel
    hi = med.vmax + med.voxel
    first = np.inf
    for d in range(3):
        if v0[d] > 0 and np.isfinite(hi[d]):
            q = (hi[d] - x0[d]) / (h * v0[d])
        elif v0[d] < 0 and np.isfinite(lo[d]):
            q = (lo[d] - x0[d]) / (h * v0[d])
        else:
            continue
        first = min(first, np.floor(q) if q > 0 else 0.0)
    k = every * np.ceil(first / every)
    return (int(k) + 1, 0) if k < max_steps else (max_steps, 1)


def check_exit_step_count():
    cv, med = _vacuum(16, 6)
    x0, v0, shape, e0 = rt.setup_rays(cv, [0.0, 17.0, 45.0, 90.0, 200.0], incl_deg=7.0)
    for every in (10, 3):
        tr = rt.Tracer(med, param="physical_path_length", solver="forward_symplectic_euler", exit_every=every)
        for r in range(0, x0.shape[0], 7):
            exp_steps, exp_left = _exit_steps_closed_form(med, x0[r], v0[r], tr.h, every, tr.max_steps)
            _, steps, left = tr.run(x0[r:r + 1], v0[r:r + 1], e0, f=np.zeros(med.n.shape), return_steps=True)
            assert (steps, left) == (exp_steps, exp_left), (every, r, steps, exp_steps)
    # a step cap below the exit leaves the ray active after exactly max_steps steps
    tr = rt.Tracer(med, param="physical_path_length", solver="forward_symplectic_euler", max_steps=20)
    _, steps, left = tr.run(x0[:1], v0[:1], e0, f=np.zeros(med.n.shape), return_steps=True)
    assert (steps, left) == (20, 1)


def test_exit_step_count():
    check_exit_step_count()


def check_setup_rays_frame():
    cv = rt.coord_vec(12, 10, 7)
    ang = np.array([0.0, 30.0, 135.0, 270.0])
    for phi in (0.0, 8.0):
        x0, v0, shape, e0 = rt.setup_rays(cv, ang, incl_deg=phi)
        assert shape == (12, 4, 7) and e0 == 1.0
        n0 = np.linspace(min(cv[0].min(), cv[1].min()), max(cv[0].max(), cv[1].max()), 12)
        n2 = np.linspace(cv[2].min(), cv[2].max(), 7)
        D = np.hypot(max(cv[0].max(), cv[1].max()), cv[2].max())
        p = np.deg2rad(phi)
        xs = x0.reshape(shape + (3,))
        vs = v0.reshape(shape + (3,))
        for ia, a in enumerate(np.deg2rad(ang)):
            d = np.array([np.cos(a) * np.cos(p), np.sin(a) * np.cos(p), np.sin(p)])
            u = np.array([-np.sin(a), np.cos(a), 0.0])
            w = np.array([-np.cos(a) * np.sin(p), -np.sin(a) * np.sin(p), np.cos(p)])
            np.testing.assert_allclose(vs[:, ia], np.broadcast_to(-d, vs[:, ia].shape), atol=1e-15)
            np.testing.assert_allclose(xs[:, ia] @ d, D, rtol=1e-13)
            np.testing.assert_allclose(xs[:, ia] @ u, np.broadcast_to(n0[:, None], (12, 7)), atol=1e-15)
            np.testing.assert_allclose(xs[:, ia] @ w, np.broadcast_to(n2[None, :], (12, 7)), atol=1e-15)
    _, _, shape2, e2 = rt.setup_rays(cv, ang, ray_density=2.0)
    assert shape2 == (24, 4, 14) and e2 == 0.25
    cv2 = rt.coord_vec(12, 10, 0)
    _, _, shape3, e3 = rt.setup_rays(cv2, ang, ray_density=2.0)
    assert shape3 == (24, 4, 1) and e3 == 0.5


def test_setup_rays_frame():
    check_setup_rays_frame()


def check_luneburg_profile():
    n = 41
    cv = rt.coord_vec(n, n, n)
    nf, g, R = rt.luneburg_n(cv)
    c = n // 2
    assert abs(nf[c, c, c] - np.sqrt(2.0)) < 1e-12
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    inside = (r < R) & (r > 0)
    np.testing.assert_allclose(nf[inside], np.sqrt(2 - (r[inside] / R) ** 2), rtol=1e-13)
    assert np.all(nf[r >= R] == 1.0)
    # grad n = dn/dr r_hat with dn/dr = -r / (R^2 n)
    dndr = -r[inside] / (R ** 2 * nf[inside])
    for k, C in enumerate((X, Y, Z)):
        np.testing.assert_allclose(g[..., k][inside], dndr * C[inside] / r[inside], rtol=1e-12, atol=1e-12)
    assert np.all(g[r >= R] == 0)


def test_luneburg_profile():
    check_luneburg_profile()


def check_coord_vec_and_attenuation_models():
    cv = rt.coord_vec(5, 4, 3, rate=500.0)
    for c, m in zip(cv, (5, 4, 3)):
        np.testing.assert_allclose(np.diff(c), 1 / 500.0, rtol=1e-12)
        np.testing.assert_allclose(c, -c[::-1], atol=1e-18)
        assert c.size == m
    assert np.array_equal(rt.coord_vec(5, 4, 0)[2], [0.0])
    x = np.array([[0.3, 0.4, 5.0], [0.3, 0.41, 0.0], [0.1, 0.1, 0.49], [0.1, 0.1, 0.5]])
    np.testing.assert_array_equal(rt.att_alpha(x, "cylinder", 2.0, 0.5), [0.0, 0.0, 2.0, 2.0])
    np.testing.assert_array_equal(rt.att_alpha(x, "ball", 2.0, 0.5), [0.0, 0.0, 0.0, 0.0])
    # |x| = 5.02, 0.508, 0.510, 0.519
    np.testing.assert_array_equal(rt.att_alpha(x, "ball", 2.0, 0.515), [0.0, 2.0, 2.0, 0.0])
    np.testing.assert_array_equal(rt.att_alpha(x, "none"), 0.0)


def test_coord_vec_and_attenuation_models():
    check_coord_vec_and_attenuation_models()


def check_luneburg_analytic():
    """The closed-form lens: grad n equals central differences of n (p = 2 and
    p = 3, n_sur = 1 and 1.3), n(0) = 2^(1/p) n_sur, n -> n_sur at r = R, the
    grid values equal the presampled profile, and parallel rays focus on the
    antipodal surface point."""
    cv = rt.coord_vec(32, 32, 32)
    rng = np.random.default_rng(3)
    for p, ns in ((2.0, 1.0), (3.0, 1.3)):
        med = rt.LuneburgMedium(cv, R=0.9, p=p, n_sur=ns)
        R = med.Rp
        x = rng.uniform(-0.6, 0.6, (50, 3)) * R
        f = med.sample(x)
        h = 1e-6 * R
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (med.sample(x + e)[:, 0] - med.sample(x - e)[:, 0]) / (2 * h)
            np.testing.assert_allclose(f[:, 1 + k], fd, rtol=1e-6, atol=1e-6 * ns / R)
        assert abs(med.sample(np.zeros((1, 3)))[0, 0] - 2 ** (1 / p) * ns) < 1e-14
        edge = med.sample(np.array([[R * (1 - 1e-9), 0.0, 0.0], [R * 1.01, 0.0, 0.0]]))
        assert abs(edge[0, 0] - ns) < 1e-6 and edge[1, 0] == ns and np.all(edge[1, 1:] == 0)
        X, Y, Z = np.meshgrid(*cv, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        np.testing.assert_allclose(med.sample(pts)[:, 0], med.n.ravel(), rtol=1e-14)
    med = rt.LuneburgMedium(rt.coord_vec(64, 64, 64))
    tr = rt.Tracer(med, param="physical_path_length", solver="rk4", step=med.voxel[0] / 4)
    R = med.Rp
    b = np.array([0.1, 0.2, 0.3]) * R
    x = np.stack([np.full(3, 1.2 * R), b, np.zeros(3)], axis=1)
    v = np.tile([-1.0, 0.0, 0.0], (3, 1))
    for _ in range(4000):
        xn, vn = tr.advance(x, v)
        if np.all(xn[:, 1] <= 0) or xn[0, 0] < -1.3 * R:
            break
        x, v = xn, vn
    assert np.all(np.abs(x[:, 1]) < 0.02 * R)
    assert np.all(np.abs(x[:, 0] + R) < 0.05 * R)


def test_luneburg_analytic():
    check_luneburg_analytic()
