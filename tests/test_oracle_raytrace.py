"""Pinning of the ray-tracer oracle (CPU): straight-ray line integrals, exact
transposition, integrator order and the Luneburg-lens focus."""

import numpy as np

import oracle.raytrace as rt


def _homog(n, nz, n0=1.0):
    cv = rt.coord_vec(n, n, nz)
    med = rt.GridMedium(cv, np.full((n, n, max(nz, 1)), n0))
    return cv, med


def test_straight_ray_line_integral():
    """Homogeneous n, no attenuation: sum of (h/voxel) * trilinear weights of f == 1
    along a ray equals its chord length (in voxels) through the grid, within one voxel."""
    n, nz = 24, 6
    cv, med = _homog(n, nz)
    x0, v0, shape, e0 = rt.setup_rays(cv, [0.0, 30.0, 90.0])
    tr = rt.Tracer(med, param="physical_path_length", solver="forward_symplectic_euler")
    g = tr.run(x0, v0, e0, f=np.ones((n, n, nz))).reshape(shape)
    # central ray at 0 degrees crosses n-1 voxel spacings plus the partial-weight margins
    c = g[shape[0] // 2, 0, nz // 2]
    assert abs(c - (n - 1)) < 1.0 + 1e-9
    # chord of the square at 30 degrees through the centre
    ch = (n - 1) / np.cos(np.deg2rad(30.0))
    assert abs(g[shape[0] // 2, 1, nz // 2] - ch) < 1.5


def test_deposit_is_transpose_of_integrate():
    rng = np.random.default_rng(0)
    n, nz = 12, 5
    cv = rt.coord_vec(n, n, nz)
    nf = 1.5 + 0.02 * rng.standard_normal((n, n, nz))
    med = rt.GridMedium(cv, nf)
    x0, v0, shape, e0 = rt.setup_rays(cv, np.arange(0.0, 180.0, 30.0), incl_deg=4.0)
    absg = rng.uniform(0.5, 1.0, (n, n, nz))
    for solver in ("forward_symplectic_euler", "rk4", "leapfrog", "forward_euler"):
        tr = rt.Tracer(med, param="canonical", solver=solver, att=("cylinder", 3.0, 0.8 * cv[0].max(), None),
                       abs_grid=absg)
        f = rng.random((n, n, nz))
        gg = rng.random(x0.shape[0])
        Wf = tr.run(x0, v0, e0, f=f)
        WTg = tr.run(x0, v0, e0, g=gg)
        assert abs(np.dot(Wf, gg) - np.sum(f * WTg)) <= 1e-10 * abs(np.dot(Wf, gg))


def test_rk4_fourth_order_and_euler_first_order():
    """Endpoint error vs step: RK4 ~ h^4, symplectic Euler ~ h.  The index is affine,
    n = 1.5 + a x + b y, which trilinear sampling reproduces exactly, so the ray
    equation is smooth and the integrators show their textbook orders."""
    n = 32
    cv = rt.coord_vec(n, n, n)
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    a, b = 20.0, -12.0
    nf = 1.5 + a * X + b * Y
    g = np.zeros(nf.shape + (3,))
    g[..., 0], g[..., 1] = a, b
    med = rt.GridMedium(cv, nf, grad=g)
    x0 = np.array([[0.025, 0.004, 0.001]])
    v0 = np.array([[-1.0, 0.2, 0.0]])
    L = 0.04

    def endpoint(solver, h):
        tr = rt.Tracer(med, param="canonical", solver=solver, step=h)
        x, v = x0.copy(), v0.copy()
        for _ in range(int(round(L / h))):
            x, v = tr.advance(x, v)
        return x[0]

    h = L / 8
    ref = endpoint("rk4", h / 32)
    e1 = np.linalg.norm(endpoint("rk4", h) - ref)
    e2 = np.linalg.norm(endpoint("rk4", h / 2) - ref)
    assert 12.0 < e1 / e2 < 20.0  # 16 for a 4th-order method
    s1 = np.linalg.norm(endpoint("forward_symplectic_euler", h) - ref)
    s2 = np.linalg.norm(endpoint("forward_symplectic_euler", h / 2) - ref)
    assert 1.6 < s1 / s2 < 2.6   # 2 for a 1st-order method


def test_luneburg_lens_focus():
    """Parallel rays entering a Luneburg lens (n = sqrt(2 - r^2/R^2), n_sur = 1)
    focus on the lens surface opposite to the entry (physical parametrisation,
    unit initial direction = n t in air)."""
    n = 64
    cv = rt.coord_vec(n, n, n)
    nf, g, R = rt.luneburg_n(cv)
    med = rt.GridMedium(cv, nf, grad=g)
    tr = rt.Tracer(med, param="physical_path_length", solver="rk4", step=med.voxel[0] / 4)
    b = np.array([0.1, 0.2, 0.3]) * R
    x = np.stack([np.full(3, 1.2 * R), b, np.zeros(3)], axis=1)
    v = np.tile([-1.0, 0.0, 0.0], (3, 1))
    cross = []
    for _ in range(4000):
        xn, vn = tr.advance(x, v)
        if np.all(xn[:, 1] <= 0) or xn[0, 0] < -1.3 * R:
            break
        x, v = xn, vn
    # the rays reach the x-axis (y = 0) near x = -R
    assert np.all(np.abs(x[:, 1]) < 0.03 * R)
    assert np.all(np.abs(x[:, 0] + R) < 0.08 * R)


def test_torch_cpu_port_matches_oracle():
    """The torch-CPU restatement of the reference's CPU path (the timed CPU
    baseline) traces the same rays as the float64 oracle: symplectic Euler,
    canonical parametrisation, cylinder attenuation, integrate and deposit."""
    import torch
    import oracle.raytrace_torch as rtt
    n, nz = 14, 6
    cv = rt.coord_vec(n, n, nz)
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    vox = cv[0][1] - cv[0][0]
    nf = 1.5 + 0.02 * np.exp(-((X - 2 * vox) ** 2 + Y ** 2 + Z ** 2) / (2 * (4 * vox) ** 2))
    alpha, R = 0.01 / vox, (n - 1) / 2 * vox
    ab = np.full((n, n, nz), 0.5)
    tr = rt.Tracer(rt.GridMedium(cv, nf), param="canonical", solver="forward_symplectic_euler",
                   att=("cylinder", alpha, R, None), abs_grid=ab)
    tt = rtt.TorchCPUTracer(cv, nf, att_alpha=alpha, att_radius=R, abs_grid=ab, param="canonical",
                            dtype=torch.float64)
    ang = np.array([0.0, 33.0, 90.0, 151.0])
    x0, v0, shape, e0 = rt.setup_rays(cv, ang)
    rng = np.random.default_rng(4)
    # off the voxel centres: the reference's floor / ceil taps count a tap
    # twice at exactly integral coordinates, the oracle's floor / floor + 1 do not
    x0 = x0 + rng.uniform(-0.05, 0.05, x0.shape) * vox
    f = rng.random((n, n, nz))
    g = rng.random(x0.shape[0])
    a = tt.run(x0, v0, e0, f=f)
    b = tr.run(x0, v0, e0, f=f)
    assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)
    a = tt.run(x0, v0, e0, g=g)
    b = tr.run(x0, v0, e0, g=g)
    assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)
