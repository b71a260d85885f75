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
