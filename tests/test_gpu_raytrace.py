"""Parity of the HIP ray tracer with the float64 oracle (relative L2 <= 1e-5),
exact transposition and determinism."""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _setup(n, nz, angles, solver, param, incl=0.0, att=True, absorb=True, bumps=True, seed=0):
    import vamtoolbox_amd as vam
    rng = np.random.default_rng(seed)
    tg = vam.geometry.TargetGeometry(target=np.ones((n, n, nz)))
    cv = tg.constructCoordVec()
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    nf = np.full(X.shape, 1.5)
    if bumps:
        for _ in range(4):
            c = rng.uniform(-0.4, 0.4, 3) * np.array([cv[0].max(), cv[1].max(), max(cv[2].max(), 1e-9)])
            s = rng.uniform(0.15, 0.3) * cv[0].max()
            nf += rng.choice([-0.02, 0.02]) * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    im = vam.medium.IndexModel(cv, type="interpolation", form="freeform", n_x=nf)
    vox = cv[0][1] - cv[0][0]
    am = vam.medium.AttenuationModel(cv, type="analytical", form="homogeneous_cylinder",
                                     alpha_internal=0.005 / vox) if att else None
    ab = vam.medium.AbsorptionModel(cv, type="analytical", form="homogeneous_cylinder",
                                    alpha_internal=0.005 / vox) if absorb else None
    pg = vam.geometry.ProjectionGeometry(angles, "ray_trace", CUDA=True, index_model=im, attenuation_model=am,
                                         absorption_model=ab, ray_trace_method="eikonal",
                                         eikonal_parametrization=param, ray_trace_ode_solver=solver,
                                         ray_trace_ray_config="parallel", inclination_angle=incl)
    return tg, pg, cv, nf, im, am, ab


def _oracle(tg, pg, cv, nf, am, ab, solver, param, incl):
    import oracle.raytrace as ort
    med = ort.GridMedium(cv, nf)
    att = ("none", 0.0, 0.0, None) if am is None else ("cylinder", am.params["alpha_internal"],
                                                       am.params["R_physical"], None)
    tr = ort.Tracer(med, param=param, solver=solver, att=att, abs_grid=None if ab is None else ab.alpha_at_grid())
    x0, v0, shape, e0 = ort.setup_rays(cv, pg.angles, incl_deg=incl)
    return tr, x0, v0, shape, e0


CASES = [
    ("forward_symplectic_euler", "physical_path_length", 0.0),
    ("forward_symplectic_euler", "canonical", 0.0),
    ("rk4", "canonical", 0.0),
    ("rk4", "optical_path_length", 6.0),
    ("leapfrog", "physical_path_length", 0.0),
    ("forward_euler", "canonical", -4.0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_integrate_deposit_parity(gpu, case):
    import torch
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    solver, param, incl = CASES[case]
    print("case", CASES[case], flush=True)
    n, nz = 20, 8
    ang = np.arange(0.0, 180.0, 15.0)
    tg, pg, cv, nf, im, am, ab = _setup(n, nz, ang, solver, param, incl)
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    tr, x0, v0, shape, e0 = _oracle(tg, pg, cv, nf, am, ab, solver, param, incl)
    rng = np.random.default_rng(1 + case)
    f = rng.random((n, n, nz))
    g = rng.random(int(np.prod(shape)))
    Wf = P.forward(torch.tensor(f, dtype=torch.float32, device=gpu)).cpu().numpy()
    steps, still = P.stats()
    WTg = P.backward(torch.tensor(g, dtype=torch.float32, device=gpu)).cpu().numpy()
    ref_f, ref_steps, ref_still = tr.run(x0, v0, e0, f=f, return_steps=True)
    ref_g = tr.run(x0, v0, e0, g=g)
    e1, e2 = rel_l2(Wf, ref_f), rel_l2(WTg, ref_g)
    print("rel L2 integrate %.2e deposit %.2e steps %d/%d still %d/%d" % (e1, e2, steps, ref_steps, still, ref_still))
    assert e1 <= TOL and e2 <= TOL
    assert steps == ref_steps and still == ref_still


def test_transpose_and_determinism(gpu):
    import torch
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    n, nz = 24, 12
    ang = np.arange(0.0, 360.0, 10.0)
    tg, pg, *_ = _setup(n, nz, ang, "rk4", "canonical", 3.0)
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    rng = np.random.default_rng(5)
    f = torch.tensor(rng.random((n, n, nz)), dtype=torch.float32, device=gpu)
    g = torch.tensor(rng.standard_normal(P.num_rays), dtype=torch.float32, device=gpu)
    Wf = P.forward(f)
    d1 = P.backward(g)
    d2 = P.backward(g)
    assert torch.equal(d1, d2), "deposit must be bitwise deterministic"
    assert torch.equal(P.forward(f), Wf)
    lhs = float(torch.dot(Wf.double(), g.double()))
    rhs = float(torch.sum(f.double() * d1.double()))
    assert abs(lhs - rhs) / (Wf.double().norm() * g.double().norm()) <= 1e-6


def test_two_dimensional(gpu):
    import torch
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    n = 32
    ang = np.arange(0.0, 180.0, 6.0)
    tg, pg, cv, nf, im, am, ab = _setup(n, 1, ang, "forward_symplectic_euler", "physical_path_length")
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    tr, x0, v0, shape, e0 = _oracle(tg, pg, cv, nf, am, ab, "forward_symplectic_euler", "physical_path_length", 0.0)
    f = np.random.default_rng(2).random((n, n, 1))
    Wf = P.forward(torch.tensor(f, dtype=torch.float32, device=gpu)).cpu().numpy()
    assert rel_l2(Wf, tr.run(x0, v0, e0, f=f)) <= TOL


MEDIA = ["ball_attenuation", "attenuation_grid", "ray_density", "luneburg_grid", "luneburg_analytic"]


@pytest.mark.parametrize("medium", MEDIA)
def test_media_variants(gpu, medium):
    """The media and ray options beyond the cylinder case against the oracle:
    homogeneous-ball attenuation (medium.py:905-915), an interpolated
    attenuation grid with zeros padding (medium.py:161-185), ray density 1.5
    (more rays, energy 1/density^2; pyTorchRayTrace.py:1679-1686), and the
    Luneburg lens presampled on the grid and in closed form
    (medium.py:486-549)."""
    import torch
    import oracle.raytrace as ort
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    print("medium", medium, flush=True)
    rng = np.random.default_rng(21)
    luneburg = medium.startswith("luneburg")
    n, nz = (24, 24) if luneburg else (20, 8)
    ang = np.arange(0.0, 180.0, 15.0)
    tg = vam.geometry.TargetGeometry(target=np.ones((n, n, nz)))
    cv = tg.constructCoordVec()
    vox = cv[0][1] - cv[0][0]
    am, att, rd = None, ("none", 0.0, 0.0, None), None
    param, solver = "canonical", "rk4"
    if luneburg:
        kind = "interpolation" if medium == "luneburg_grid" else "analytical"
        im = vam.medium.IndexModel(cv, type=kind, form="luneburg_lens")
        nf, gl, R = ort.luneburg_n(cv)
        med = ort.GridMedium(cv, nf, grad=gl) if kind == "interpolation" else ort.LuneburgMedium(cv)
        param = "physical_path_length"
    else:
        X, Y, Z = np.meshgrid(*cv, indexing="ij")
        nf = 1.5 + 0.02 * np.exp(-((X - 3 * vox) ** 2 + Y ** 2 + Z ** 2) / (2 * (6 * vox) ** 2))
        im = vam.medium.IndexModel(cv, type="interpolation", form="freeform", n_x=nf)
        med = ort.GridMedium(cv, nf)
        if medium == "ball_attenuation":
            am = vam.medium.AttenuationModel(cv, type="analytical", form="homogeneous_ball", R=0.8,
                                             alpha_internal=0.02 / vox)
            att = ("ball", am.params["alpha_internal"], am.params["R_physical"], None)
        elif medium == "attenuation_grid":
            a = rng.uniform(0.0, 0.03 / vox, (n, n, nz))
            am = vam.medium.AttenuationModel(cv, type="interpolation", form="freeform", alpha_x=a)
            att = ("grid", 0.0, 0.0, a)
        else:
            rd = 1.5
    pg = vam.geometry.ProjectionGeometry(ang, "ray_trace", CUDA=True, index_model=im, attenuation_model=am,
                                         ray_trace_method="eikonal", eikonal_parametrization=param,
                                         ray_trace_ode_solver=solver, ray_trace_ray_config="parallel",
                                         ray_density=rd)
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    tr = ort.Tracer(med, param=param, solver=solver, att=att)
    x0, v0, shape, e0 = ort.setup_rays(cv, ang, ray_density=1.0 if rd is None else rd)
    assert tuple(P.sino_shape) == tuple(shape)
    f = rng.random((n, n, nz))
    g = rng.random(int(np.prod(shape)))
    Wf = P.forward(torch.tensor(f, dtype=torch.float32, device=gpu)).cpu().numpy()
    steps, still = P.stats()
    WTg = P.backward(torch.tensor(g, dtype=torch.float32, device=gpu)).cpu().numpy()
    ref_f, ref_steps, ref_still = tr.run(x0, v0, e0, f=f, return_steps=True)
    ref_g = tr.run(x0, v0, e0, g=g)
    e1, e2 = rel_l2(Wf, ref_f), rel_l2(WTg, ref_g)
    print("rel L2 integrate %.2e deposit %.2e steps %d/%d still %d/%d" % (e1, e2, steps, ref_steps, still, ref_still))
    assert e1 <= TOL and e2 <= TOL
    assert steps == ref_steps and still == ref_still


def _rt_bclp_problem():
    """16 x 16 x 4 graded medium, 12 angles, RK4 canonical, attenuation and
    absorption in a cylinder of radius 7.5 voxels (the default radius, the
    smallest half-span, would be the 1.5-voxel half-height here), a smooth
    target in [0, 1] inside it."""
    import oracle.raytrace as ort
    import vamtoolbox_amd as vam
    n, nz = 16, 4
    ang = np.arange(0.0, 180.0, 15.0)
    tg0 = vam.geometry.TargetGeometry(target=np.ones((n, n, nz)))
    cv = tg0.constructCoordVec()
    vox = cv[0][1] - cv[0][0]
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    nf = 1.5 + 0.02 * np.exp(-((X - 2 * vox) ** 2 + Y ** 2 + Z ** 2) / (2 * (4 * vox) ** 2))
    T = np.clip(1.2 * np.exp(-(X ** 2 + (Y - vox) ** 2) / (2 * (2.5 * vox) ** 2)) - 0.1, 0, 1)
    alpha = 0.005 / vox
    cyl = dict(alpha_internal=alpha, R=7.5 * vox, length_unit_of_R="physical_coordinate")
    im = vam.medium.IndexModel(cv, type="interpolation", form="freeform", n_x=nf)
    am = vam.medium.AttenuationModel(cv, type="analytical", form="homogeneous_cylinder", **cyl)
    ab = vam.medium.AbsorptionModel(cv, type="analytical", form="homogeneous_cylinder", **cyl)
    pg = vam.geometry.ProjectionGeometry(ang, "ray_trace", CUDA=True, index_model=im, attenuation_model=am,
                                         absorption_model=ab, ray_trace_method="eikonal",
                                         eikonal_parametrization="canonical", ray_trace_ode_solver="rk4",
                                         ray_trace_ray_config="parallel")
    tg = vam.geometry.TargetGeometry(target=T, clip_to_circle=False)
    tr = ort.Tracer(ort.GridMedium(cv, nf), param="canonical", solver="rk4",
                    att=("cylinder", am.params["alpha_internal"], am.params["R_physical"], None),
                    abs_grid=ab.alpha_at_grid())
    x0, v0, shape, e0 = ort.setup_rays(cv, ang)
    return tg, pg, tr, (x0, v0, shape, e0), ab.alpha_at_grid(), ang


def test_inverse_backward_matches_oracle(gpu):
    """The initial sinogram of the ray-tracer BCLP (pyTorchRayTrace.py:135-166):
    f / alpha^2 (0 where alpha = 0), forward, Ram-Lak filter along t,
    times pi / (2 n_angles)."""
    import torch
    import oracle.optim as oo
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    tg, pg, tr, (x0, v0, shape, e0), ab, ang = _rt_bclp_problem()
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    f = np.asarray(tg.array, dtype=np.float64)
    g = P.inverseBackward(torch.tensor(f, dtype=torch.float32, device=gpu)).cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        fa = f / ab ** 2
    fa[~np.isfinite(fa)] = 0.0
    ref = oo.filter_sinogram(tr.run(x0, v0, e0, f=fa).reshape(shape), "ram-lak").ravel() * (np.pi / (2 * len(ang)))
    e = rel_l2(g, ref)
    print("inverseBackward rel L2 %.2e" % e)
    assert e <= TOL


def test_raytrace_bclp_fused_matches_oracle(gpu):
    """20 iterations of the fused ray-tracer BCLP (vamx_rt_bclp_step: deposit
    with the response / band / Lp epilogue, integrate fused with the update)
    against the oracle's BCLP loop on the float64 tracer, from the
    inverseBackward initial sinogram; and the graph-replayed loop equals the
    eager one."""
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, tr, (x0, v0, shape, e0), ab, ang = _rt_bclp_problem()
    T = np.asarray(tg.array, dtype=np.float64)
    fwd = lambda f: tr.run(x0, v0, e0, f=f)      # noqa: E731
    bwd = lambda g: tr.run(x0, v0, e0, g=g)      # noqa: E731
    ri = oo.Response("interpolation", B=20, M=0.6)
    f = ri.map_inv(T)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = f / ab ** 2
    f[~np.isfinite(f)] = 0.0
    # 0.7 x the inverseBackward sinogram: the dose falls short of the band and
    # the descent does real work (loss 7.27 -> 5.55, g moves by 19 %)
    g0 = 0.7 * np.clip(oo.filter_sinogram(fwd(f).reshape(shape), "ram-lak").ravel() * (np.pi / (2 * len(ang))), 0,
                       None)
    lr, n_iter = 1e-4, 20
    ro = oo.Response("analytical", B=20, M=0.6)
    g_ref, dose_ref, _, loss_ref = oo.bclp(T, fwd, bwd, ro, n_iter, g0, eps=0.05, lr=lr)
    assert np.all(np.diff(loss_ref) < 0), "test problem must descend"
    assert loss_ref[-1] < 0.85 * loss_ref[0] and rel_l2(g_ref, g0) > 0.1, "the update must move g"
    rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
    res = {}
    for graph in (False, True):
        opts = vam.optimize.Options(method="BCLP", n_iter=n_iter, response_model=rm, eps=0.05, p=2, q=1,
                                    learning_rate=lr, g0=g0, fused=True, hip_graph=graph)
        g, dose, mapped, logs = vam.optimize.optimize(tg, pg, opts, output="components")
        res[graph] = (np.asarray(g), np.asarray(dose), np.asarray(logs.loss))
    g, dose, loss = res[True]
    e_g, e_d = rel_l2(g.ravel(), g_ref.ravel()), rel_l2(dose.ravel(), dose_ref.ravel())
    e_l = np.max(np.abs(loss - loss_ref) / np.abs(loss_ref))
    print("ray-trace BCLP x%d: rel L2 g %.2e dose %.2e loss %.2e (%.4g -> %.4g)" % (n_iter, e_g, e_d, e_l,
                                                                                   loss_ref[0], loss_ref[-1]))
    assert e_g <= 1e-4 and e_d <= 1e-4 and e_l <= 1e-4
    np.testing.assert_array_equal(res[False][0], res[True][0])
    np.testing.assert_array_equal(res[False][2], res[True][2])


def test_autograd_ray_trace_and_deposit(gpu):
    """RayTrace / RayDeposit: each one's backward is the other (the exact
    transpose); gradients through autograd equal the oracle's deposit and
    integrate, and the dot-product identity <W f, w> = <f, W^T w> holds."""
    import torch
    from vamtoolbox_amd.autograd import RayDeposit, RayTrace
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    tg, pg, tr, (x0, v0, shape, e0), ab, ang = _rt_bclp_problem()
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    rng = np.random.default_rng(5)
    f = rng.uniform(0, 1, P.vol_shape)
    w = rng.uniform(-1, 1, P.num_rays)
    ft = torch.tensor(f, dtype=torch.float32, device=gpu, requires_grad=True)
    wt = torch.tensor(w, dtype=torch.float32, device=gpu)
    y = RayTrace.apply(ft, P)
    (y * wt).sum().backward()
    e_fwd = rel_l2(y.detach().cpu().numpy(), tr.run(x0, v0, e0, f=f))
    e_grad = rel_l2(ft.grad.cpu().numpy(), tr.run(x0, v0, e0, g=w))
    gt = torch.tensor(np.abs(w), dtype=torch.float32, device=gpu, requires_grad=True)
    u = rng.uniform(-1, 1, P.vol_shape)
    d = RayDeposit.apply(gt, P)
    (d * torch.tensor(u, dtype=torch.float32, device=gpu)).sum().backward()
    e_dep = rel_l2(d.detach().cpu().numpy(), tr.run(x0, v0, e0, g=np.abs(w)))
    e_grad2 = rel_l2(gt.grad.cpu().numpy(), tr.run(x0, v0, e0, f=u))
    lhs = float(np.dot(y.detach().cpu().numpy().astype(np.float64), w))
    rhs = float(np.sum(f * RayDeposit.apply(wt, P).cpu().numpy()))
    print("RayTrace fwd %.2e grad %.2e; RayDeposit fwd %.2e grad %.2e; dot %.6e vs %.6e" % (
        e_fwd, e_grad, e_dep, e_grad2, lhs, rhs))
    assert max(e_fwd, e_grad, e_dep, e_grad2) <= TOL
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


def test_autograd_raytrace_bclp_loss(gpu):
    """RayTraceBCLPLoss (vamx_rt_bclp_eval forward, integrate backward)
    against the oracle's BCLP loss and its gradient q L^((q-p)/q) W (alpha
    operand) on the float64 tracer (BCLP.py:277-386), for p = 2, q = 1 and
    p = 3, q = 2; and against a central difference of the GPU loss."""
    import torch
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.autograd import RayTraceBCLPLoss
    from vamtoolbox_amd.projector.pyTorchRayTrace import PyTorchRayTracingPropagator
    tg, pg, tr, (x0, v0, shape, e0), ab, ang = _rt_bclp_problem()
    P = PyTorchRayTracingPropagator(tg, pg, output_torch_tensor=True)
    T = np.asarray(tg.array, dtype=np.float64)
    fwd = lambda f: tr.run(x0, v0, e0, f=f)      # noqa: E731
    bwd = lambda g: tr.run(x0, v0, e0, g=g)      # noqa: E731
    ri = oo.Response("interpolation", B=20, M=0.6)
    f = ri.map_inv(T)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = f / ab ** 2
    f[~np.isfinite(f)] = 0.0
    g0 = 0.7 * np.clip(oo.filter_sinogram(fwd(f).reshape(shape), "ram-lak").ravel() * (np.pi / (2 * len(ang))), 0,
                       None)
    rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
    ro = oo.Response("analytical", B=20, M=0.6)
    Tt = torch.tensor(T, dtype=torch.float32, device=gpu)
    rng = np.random.default_rng(8)
    direction = rng.uniform(-1, 1, g0.shape) * g0
    for p, q in ((2.0, 1.0), (3.0, 2.0)):
        gt = torch.tensor(g0, dtype=torch.float32, device=gpu, requires_grad=True)
        L = RayTraceBCLPLoss.apply(gt, P, rm, Tt, 0.05, 1.0, p, q)
        L.backward()
        _, _, L_ref, operand = oo.bclp_state(g0, T, bwd, ro, 0.05, 1.0, p, q)
        G_ref = q * L_ref ** ((q - p) / q) * fwd(operand)
        e_L = abs(float(L.detach()) - L_ref) / L_ref
        e_G = rel_l2(gt.grad.cpu().numpy(), G_ref)
        h = 5e-2   # float64 oracle: the central difference is within 2.4e-3 of the gradient here
        with torch.no_grad():
            dt = torch.tensor(direction, dtype=torch.float32, device=gpu)
            Lp = float(RayTraceBCLPLoss.apply(gt + h * dt, P, rm, Tt, 0.05, 1.0, p, q))
            Lm = float(RayTraceBCLPLoss.apply(gt - h * dt, P, rm, Tt, 0.05, 1.0, p, q))
        fd = (Lp - Lm) / (2 * h)
        an = float(np.dot(gt.grad.cpu().numpy().astype(np.float64), direction))
        print("p=%g q=%g: loss %.6e vs %.6e (rel %.1e), grad rel L2 %.2e, central difference %.5e vs %.5e" % (
            p, q, float(L.detach()), L_ref, e_L, e_G, fd, an))
        assert e_L <= 1e-5 and e_G <= 1e-4
        assert abs(fd - an) <= 1e-2 * abs(an)
