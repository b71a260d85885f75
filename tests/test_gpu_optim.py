"""OSMO / CAL / BCLP on the GPU against the float64 oracle loops (iterates within
1e-4 relative L2), fused vs unfused vs graph-replayed BCLP, and the autograd
adjoints (dot-product and finite differences)."""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _ellipsoids(n, nz, seed, count=4, gray=False):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(np.arange(nz) - (nz - 1) / 2, np.arange(n) - (n - 1) / 2,
                             np.arange(n) - (n - 1) / 2, indexing="ij")
    vol = np.zeros((nz, n, n))
    for _ in range(count):
        a, b = rng.uniform(0.12, 0.3, 2) * n
        c = rng.uniform(0.3, 0.6) * nz
        cx, cy = rng.uniform(-0.15, 0.15, 2) * n
        cz = rng.uniform(-0.2, 0.2) * nz
        vol = np.maximum(vol, (((xx - cx) / a) ** 2 + ((yy - cy) / b) ** 2 + ((zz - cz) / c) ** 2 <= 1).astype(float))
    vol = np.transpose(vol, (1, 2, 0)).copy()
    if gray:
        from scipy.ndimage import gaussian_filter
        vol = np.clip(gaussian_filter(vol, 1.0), 0, 1)
    return vol


def _problem(n=32, nz=4, na=60, seed=0, gray=False, **pg_kw):
    import vamtoolbox_amd as vam
    t = _ellipsoids(n, nz, seed, gray=gray)
    tg = vam.geometry.TargetGeometry(target=t)
    ang = np.linspace(0, 360, na, endpoint=False)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True, **pg_kw)
    return tg, pg, ang


def _oracle_ops(ang, **kw):
    import oracle.d1 as d1
    return (lambda v: d1.forward(v, ang, **kw)), (lambda s: d1.backward(s, ang, **kw))


def test_osmo_matches_oracle(gpu):
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem()
    opts = vam.optimize.Options(method="OSMO", n_iter=8, d_h=0.85, d_l=0.6, filter="ram-lak")
    sino, recon, err = vam.optimize.optimize(tg, pg, opts)
    fwd, bwd = _oracle_ops(ang)
    b_ref, x_ref, err_ref = oo.osmo(tg.array, fwd, bwd, 8, d_h=0.85, d_l=0.6)
    e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
    print("OSMO rel L2 sinogram %.2e recon %.2e" % (e_b, e_x), err, err_ref)
    assert e_b <= 1e-4 and e_x <= 1e-4
    np.testing.assert_allclose(err, err_ref, atol=2e-3)


def test_cal_matches_oracle(gpu):
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem(seed=1)
    opts = vam.optimize.Options(method="CAL", n_iter=8, learning_rate=0.002, sigmoid=0.01)
    sino, recon, err = vam.optimize.optimize(tg, pg, opts)
    fwd, bwd = _oracle_ops(ang)
    b_ref, x_ref, err_ref = oo.cal(tg.array, fwd, bwd, 8, lr=0.002, sigmoid=0.01)
    e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
    print("CAL rel L2 sinogram %.2e recon %.2e" % (e_b, e_x))
    assert e_b <= 1e-4 and e_x <= 1e-4


def test_cal_threshold_search_matches_oracle(gpu):
    """CAL without d_h (CAL.py:39-110): each iteration picks the threshold of
    100 that maximises gel-above minus void-above fractions; the device
    counts (sorted values) give the oracle's threshold and iterates."""
    import torch
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.optimizer.CAL import _threshold_reconstruction
    tg, pg, ang = _problem(seed=3)
    gel, void = np.asarray(tg.gel_inds), np.asarray(tg.void_inds)
    rng = np.random.default_rng(8)
    x = rng.random(tg.array.shape).astype(np.float32)
    xt, th = _threshold_reconstruction(torch.tensor(x, device=gpu), tg, 0.5)
    xt_ref, th_ref = oo.cal_threshold(x.astype(np.float64), gel, void, 0.5)
    assert th == th_ref
    np.testing.assert_allclose(xt.cpu().numpy(), xt_ref, rtol=1e-12, atol=0)
    opts = vam.optimize.Options(method="CAL", n_iter=6, learning_rate=0.002, sigmoid=20.0, d_h=None)
    sino, recon, err = vam.optimize.optimize(tg, pg, opts)
    fwd, bwd = _oracle_ops(ang)
    b_ref, x_ref, err_ref = oo.cal(tg.array, fwd, bwd, 6, d_h=None, lr=0.002, sigmoid=20.0)
    e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
    print("CAL (threshold search) rel L2 sinogram %.2e recon %.2e" % (e_b, e_x))
    assert e_b <= 1e-4 and e_x <= 1e-4


def _bclp_setup(gpu, kind="analytical", n_iter=20, **pg_kw):
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem(n=32, nz=8, na=45, seed=2, gray=True, **pg_kw)
    if kind == "analytical":
        rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
        ro = oo.Response("analytical", B=20, M=0.6)
    else:
        rm = vam.response.ResponseModel(B=25)
        ro = oo.Response("interpolation", B=25)
    # initial sinogram from the interpolation inverse (finite at zero response)
    fwd, bwd = _oracle_ops(ang, **{k: v for k, v in (("mu", pg.absorption_per_voxel),
                                                     ("det_offset", pg.det_offset)) if v})
    # 0.8 x the filtered initial guess: the band is violated and the descent is monotone
    g0 = 0.8 * oo.bclp_init(tg.array, fwd, oo.Response("interpolation", B=20, M=0.6), len(ang))
    lr = 1e-4 if kind == "analytical" else 1e-5
    return tg, pg, ang, rm, ro, g0, fwd, bwd, lr


@pytest.mark.parametrize("kind", ["analytical", "interpolation"])
def test_bclp_fused_matches_oracle(gpu, kind):
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, kind)
    n_iter = 20
    opts = vam.optimize.Options(method="BCLP", n_iter=n_iter, response_model=rm, eps=0.05, p=2, q=1,
                                learning_rate=lr, g0=g0, fused=True)
    g, dose, mapped, logs = vam.optimize.optimize(tg, pg, opts, output="components")
    g_ref, dose_ref, m_ref, loss_ref = oo.bclp(tg.array, fwd, bwd, ro, n_iter, g0, eps=0.05, lr=lr)
    e_g, e_d = rel_l2(g, g_ref), rel_l2(dose, dose_ref)
    assert np.all(loss_ref > 0) and np.all(np.diff(loss_ref) < 0), "test problem must descend"
    e_l = np.max(np.abs(logs.loss - loss_ref) / np.abs(loss_ref))
    print("BCLP(%s) rel L2 g %.2e dose %.2e loss %.2e" % (kind, e_g, e_d, e_l))
    assert e_g <= 1e-4 and e_d <= 1e-4 and e_l <= 1e-4


@pytest.mark.parametrize("method", ["OSMO", "BCLP"])
def test_fifty_iterations_match_oracle(gpu, method):
    """50 iterations (the length the iterate tolerance is quoted at): OSMO
    and the fused BCLP iterates within 1e-4 relative L2 of the float64
    oracle loops."""
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    n_iter = 50
    if method == "OSMO":
        tg, pg, ang = _problem(seed=4)
        opts = vam.optimize.Options(method="OSMO", n_iter=n_iter, d_h=0.85, d_l=0.6, filter="ram-lak")
        sino, recon, err = vam.optimize.optimize(tg, pg, opts)
        fwd, bwd = _oracle_ops(ang)
        b_ref, x_ref, err_ref = oo.osmo(tg.array, fwd, bwd, n_iter, d_h=0.85, d_l=0.6)
        e1, e2 = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
        print("OSMO x50 rel L2 sinogram %.2e recon %.2e; VER %.4f vs %.4f" % (e1, e2, err[-1], err_ref[-1]))
    else:
        tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, "analytical")
        opts = vam.optimize.Options(method="BCLP", n_iter=n_iter, response_model=rm, eps=0.05, p=2, q=1,
                                    learning_rate=lr, g0=g0, fused=True)
        g, dose, mapped, logs = vam.optimize.optimize(tg, pg, opts, output="components")
        g_ref, dose_ref, m_ref, loss_ref = oo.bclp(tg.array, fwd, bwd, ro, n_iter, g0, eps=0.05, lr=lr)
        e1, e2 = rel_l2(g, g_ref), rel_l2(dose, dose_ref)
        # the band is met before the end (loss 0 exactly): the loss error is
        # measured against the initial loss
        e_l = np.max(np.abs(logs.loss - loss_ref)) / loss_ref[0]
        print("BCLP x50 rel L2 g %.2e dose %.2e; loss %.2e (%.4g -> %.4g)" % (e1, e2, e_l, loss_ref[0],
                                                                              loss_ref[-1]))
        assert e_l <= 1e-4
    assert e1 <= 1e-4 and e2 <= 1e-4


def test_bclp_fused_unfused_graph_agree(gpu):
    import vamtoolbox_amd as vam
    tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, "analytical", mu=0.01, det_offset=2.0)
    res = {}
    for name, kw in (("graph", dict(fused=True, hip_graph=True)), ("eager", dict(fused=True, hip_graph=False)),
                     ("unfused", dict(fused=False))):
        opts = vam.optimize.Options(method="BCLP", n_iter=12, response_model=rm, eps=0.05, learning_rate=lr,
                                    g0=g0, **kw)
        res[name] = vam.optimize.optimize(tg, pg, opts, output="components")
    assert np.array_equal(res["graph"][0], res["eager"][0]), "graph replay must equal eager bitwise"
    assert np.array_equal(res["graph"][3].loss, res["eager"][3].loss)
    e = rel_l2(res["unfused"][0], res["eager"][0])
    print("fused vs unfused rel L2 g %.2e" % e)
    assert e <= 1e-5


def test_bclp_exact_gradient_matches_oracle(gpu):
    import oracle.d1 as d1
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, "analytical")
    opts = vam.optimize.Options(method="BCLP", n_iter=10, response_model=rm, eps=0.05, learning_rate=lr,
                                g0=g0, gradient="exact")
    g, dose, mapped, logs = vam.optimize.optimize(tg, pg, opts, output="components")
    g_ref, dose_ref, _, loss_ref = oo.bclp(tg.array, fwd, bwd, ro, 10, g0, eps=0.05, lr=lr,
                                           grad_op=lambda v: d1.backward_adjoint(v, ang))
    e_g = rel_l2(g, g_ref)
    print("BCLP exact-gradient rel L2 g %.2e" % e_g)
    assert e_g <= 1e-4


def test_bclp_autograd_gradient(gpu):
    """BCLPLoss backward == oracle exact gradient, and == central finite differences
    (eps = 1e-3) at 64 random sinogram pixels."""
    import torch
    import oracle.d1 as d1
    import oracle.optim as oo
    from vamtoolbox_amd.autograd import BCLPLoss
    from vamtoolbox_amd.projectorconstructor import projectorconstructor
    tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, "analytical")
    P = projectorconstructor(tg, pg)
    T = torch.tensor(tg.array, dtype=torch.float32, device=gpu)
    g = torch.tensor(g0, dtype=torch.float32, device=gpu, requires_grad=True)
    L = BCLPLoss.apply(g, P, rm, T, 0.05, 1.0, 2.0, 1.0)
    L.backward()
    grad = g.grad.double().cpu().numpy()
    _, _, Lr, operand = oo.bclp_state(g0.astype(np.float32).astype(np.float64), tg.array, bwd, ro, 0.05, 1.0, 2.0, 1.0)
    grad_ref = (1.0 * Lr ** ((1.0 - 2.0) / 1.0)) * d1.backward_adjoint(operand, ang)
    Lg = float(L.detach())
    print("loss %.6e vs %.6e; grad rel L2 %.2e" % (Lg, Lr, rel_l2(grad, grad_ref)))
    assert abs(Lg - Lr) / Lr <= 1e-5
    assert rel_l2(grad, grad_ref) <= 1e-4
    rng = np.random.default_rng(0)
    h = 1e-3
    g64 = g0.astype(np.float64)
    worst = 0.0
    for _ in range(64):
        ii = tuple(rng.integers(0, s) for s in g0.shape)
        gp, gm = g64.copy(), g64.copy()
        gp[ii] += h
        gm[ii] -= h
        fd = (oo.bclp_state(gp, tg.array, bwd, ro, 0.05, 1.0, 2.0, 1.0)[2] -
              oo.bclp_state(gm, tg.array, bwd, ro, 0.05, 1.0, 2.0, 1.0)[2]) / (2 * h)
        worst = max(worst, abs(fd - grad[ii]) / max(np.abs(grad).max(), 1e-12))
    print("finite-difference worst error relative to max|grad|: %.2e" % worst)
    assert worst <= 1e-3


def test_autograd_dot_products(gpu):
    import torch
    from vamtoolbox_amd.autograd import ParallelBackproject, ParallelProject
    from vamtoolbox_amd.projectorconstructor import projectorconstructor
    tg, pg, ang = _problem(n=24, nz=4, na=30, seed=3)
    P = projectorconstructor(tg, pg)
    rng = np.random.default_rng(1)
    # mixed signs: the warp output clip (active only for slices without zeros
    # or of one sign) is inactive, so the forward is linear
    x = torch.tensor(rng.uniform(-1, 1, P.vol_shape), dtype=torch.float32, device=gpu, requires_grad=True)
    y = torch.tensor(rng.random(P.sino_shape), dtype=torch.float32, device=gpu)
    (ParallelProject.apply(x, P) * y).sum().backward()
    lhs = float((P.plan.forward(x.detach()).double() * y.double()).sum())
    rhs = float((x.detach().double() * x.grad.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-6
    s = torch.tensor(rng.random(P.sino_shape), dtype=torch.float32, device=gpu, requires_grad=True)
    v = torch.tensor(rng.random(P.vol_shape), dtype=torch.float32, device=gpu)
    (ParallelBackproject.apply(s, P) * v).sum().backward()
    lhs = float((P.plan.backward(s.detach()).double() * v.double()).sum())
    rhs = float((s.detach().double() * s.grad.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-6


@pytest.mark.parametrize("nz", [64, 40, 128, 256])
def test_bclp_fused_wave_path_matches_oracle(gpu, nz):
    """nz = 64 and 40 (a partial z chunk; one z per lane), 128 and 256 (two
    and four z per lane, the benchmark's kernels): the fused BCLP kernels take
    the wavefront-uniform path."""
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem(n=24, nz=nz, na=30, seed=4, gray=True, absorption_per_voxel=0.01, det_offset=1.0)
    fwd, bwd = _oracle_ops(ang, mu=0.01, det_offset=1.0)
    g0 = 0.8 * oo.bclp_init(tg.array, fwd, oo.Response("interpolation", B=20, M=0.6), len(ang))
    rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
    ro = oo.Response("analytical", B=20, M=0.6)
    lr = 1e-4
    opts = vam.optimize.Options(method="BCLP", n_iter=10, response_model=rm, eps=0.05, learning_rate=lr, g0=g0)
    g, dose, _, logs = vam.optimize.optimize(tg, pg, opts, output="components")
    g_ref, dose_ref, _, loss_ref = oo.bclp(tg.array, fwd, bwd, ro, 10, g0, eps=0.05, lr=lr)
    e_g, e_d = rel_l2(g, g_ref), rel_l2(dose, dose_ref)
    e_l = np.max(np.abs(logs.loss - loss_ref) / np.abs(loss_ref))
    print("BCLP wave path rel L2 g %.2e dose %.2e loss %.2e" % (e_g, e_d, e_l), loss_ref[[0, -1]])
    assert loss_ref[-1] < loss_ref[0]
    assert e_g <= 1e-4 and e_d <= 1e-4 and e_l <= 1e-4


def test_osmo_with_insert_matches_oracle(gpu):
    """OSMO through projectorconstructor with an insert (occlusion shadowing)."""
    import oracle.d1 as d1
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    n, nz, na = 32, 4, 60
    t = _ellipsoids(n, nz, 6)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    insert = np.repeat(((xx + 9) ** 2 + yy ** 2 <= 9)[:, :, None], nz, axis=2).astype(np.uint8)
    t = t * (1 - insert)
    tg = vam.geometry.TargetGeometry(target=t, insert=insert)
    ang = np.linspace(0, 360, na, endpoint=False)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True)
    opts = vam.optimize.Options(method="OSMO", n_iter=6, d_h=0.85, d_l=0.6, filter="ram-lak")
    sino, recon, err = vam.optimize.optimize(tg, pg, opts)
    occ = d1.occlusion_sinogram(insert, ang)
    b_ref, x_ref, err_ref = oo.osmo(tg.array, lambda v: d1.forward(v, ang, occ=occ),
                                    lambda s: d1.backward(s, ang, occ=occ), 6, d_h=0.85, d_l=0.6)
    e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
    print("OSMO with insert: rel L2 sinogram %.2e recon %.2e" % (e_b, e_x))
    assert e_b <= 1e-4 and e_x <= 1e-4


def test_fbp_matches_oracle(gpu):
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem(seed=3)
    for offset in (False, True):
        sino, recon, err = vam.optimize.optimize(tg, pg, vam.optimize.Options(method="FBP", offset=offset))
        fwd, bwd = _oracle_ops(ang)
        b_ref, x_ref, ver_ref = oo.fbp(tg.array, fwd, bwd, offset=offset)
        e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
        print("FBP offset=%s rel L2 sinogram %.2e recon %.2e VER %.4f / %.4f" % (offset, e_b, e_x, err[0], ver_ref))
        assert e_b <= 1e-5 and e_x <= 1e-5
        assert abs(err[0] - ver_ref) <= 1e-3


def test_pm_matches_oracle(gpu):
    """PM (L-BFGS-B on the host, projections on the GPU) against the same
    scipy iteration on the float64 oracle operators."""
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    tg, pg, ang = _problem(n=24, nz=2, na=40, seed=5)
    opts = vam.optimize.Options(method="PM", n_iter=4, d_h=0.8, d_l=0.7, p=2)
    sino, recon, err = vam.optimize.optimize(tg, pg, opts)
    fwd, bwd = _oracle_ops(ang)
    b_ref, x_ref = oo.pm(tg.array, fwd, bwd, 4, p=2.0, d_h=0.8, d_l=0.7)
    e_b, e_x = rel_l2(sino.array, b_ref), rel_l2(recon.array, x_ref)
    L0, _, _ = oo.pm_loss_grad(np.clip(oo.filter_sinogram(fwd(tg.array)), 0, None) /
                               np.max(np.clip(oo.filter_sinogram(fwd(tg.array)), 0, None)), tg.array, bwd, fwd, p=2.0)
    L1, _, _ = oo.pm_loss_grad(sino.array, tg.array, bwd, fwd, p=2.0)
    print("PM rel L2 sinogram %.2e recon %.2e; loss %.4g -> %.4g" % (e_b, e_x, L0, L1))
    assert L1 < L0
    assert e_b <= 1e-3 and e_x <= 1e-3


def test_bclp_eval_then_update_equals_step(gpu):
    """vamx_bclp_eval (operand left in the workspace) followed by
    vamx_bclp_update is vamx_bclp_step bit for bit, iteration after iteration."""
    import ctypes
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.parallel import ParallelPlan
    from vamtoolbox_amd.response import ResponseModel
    rng = np.random.default_rng(8)
    n, nz, na = 48, 64, 90
    plan = ParallelPlan(n, nz, np.linspace(0, 360, na, endpoint=False), mu=0.01, det_offset=3.0, device=gpu)
    T = torch.tensor(np.clip(rng.random((n, n, nz)) * 1.2 - 0.1, 0, 1), dtype=torch.float32, device=gpu)
    g0 = torch.tensor(rng.random((n, na, nz)) * 0.05, dtype=torch.float32, device=gpu)
    lib = _lib.load()
    resp, _ = ResponseModel(type="analytical", B=20, M=0.6).vamx_struct(gpu)
    prm = _lib.BclpParams(p=2.0, q=1.0, eps=0.05, weight=1.0, learning_rate=1e-3, glb=0.0, gub=np.inf, dvol=1.0,
                          grad_op=0)
    nbytes = ctypes.c_int64()
    _lib.check(lib.vamx_bclp_workspace_bytes(plan.handle, ctypes.byref(nbytes)), "workspace")
    st = _lib.stream_ptr(gpu)
    ws1 = torch.empty(int(nbytes.value), device=gpu, dtype=torch.uint8)
    ws2 = torch.empty_like(ws1)
    g1, g2 = g0.clone(), g0.clone()
    d1, d2 = torch.empty(plan.vol_shape, device=gpu), torch.empty(plan.vol_shape, device=gpu)
    log = torch.zeros(4, device=gpu, dtype=torch.float64)
    cnt = torch.zeros(1, device=gpu, dtype=torch.int32)
    l2 = torch.zeros(1, device=gpu, dtype=torch.float64)
    for it in range(3):
        _lib.check(lib.vamx_bclp_step(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g1.data_ptr(), T.data_ptr(),
                                      0, 0, d1.data_ptr(), log.data_ptr(), cnt.data_ptr(), ws1.data_ptr(), st), "step")
        _lib.check(lib.vamx_bclp_eval(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g2.data_ptr(), T.data_ptr(),
                                      0, 0, d2.data_ptr(), 0, l2.data_ptr(), ws2.data_ptr(), st), "eval")
        _lib.check(lib.vamx_bclp_update(plan.handle, ctypes.byref(prm), g2.data_ptr(), ws2.data_ptr(), st), "update")
        torch.cuda.synchronize()
        assert torch.equal(d1, d2), it
        assert float(log[it].item()) == float(l2.item()), it
        assert torch.equal(g1, g2), it
    assert not torch.equal(g1, g0)


@pytest.mark.parametrize("nz", [64, 40, 256])
def test_bclp_exact_gradient_step_kernels_agree(gpu, nz):
    """vamx_bclp_step with grad_op = 1 (the exact transpose of the back
    projection fused with the update): the wavefront kernel pairing opposite
    bins equals the unpaired wavefront kernel bit for bit, and both agree with
    the per-lane kernels (whose loss partials are summed in another order) to
    float32 rounding, over three iterations; an occlusion plan is rejected."""
    import ctypes
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.parallel import ParallelPlan
    from vamtoolbox_amd.response import ResponseModel
    rng = np.random.default_rng(9)
    n, na = 40, 90
    ang = np.linspace(0, 360, na, endpoint=False)
    T = torch.tensor(np.clip(rng.random((n, n, nz)) * 1.2 - 0.1, 0, 1), dtype=torch.float32, device=gpu)
    g0 = torch.tensor(rng.random((n, na, nz)) * 0.05, dtype=torch.float32, device=gpu)
    lib = _lib.load()
    resp, _ = ResponseModel(type="analytical", B=20, M=0.6).vamx_struct(gpu)
    prm = _lib.BclpParams(p=2.0, q=1.0, eps=0.05, weight=1.0, learning_rate=1e-3, glb=0.0, gub=np.inf, dvol=1.0,
                          grad_op=1)
    st = _lib.stream_ptr(gpu)

    def run(plan):
        nbytes = ctypes.c_int64()
        _lib.check(lib.vamx_bclp_workspace_bytes(plan.handle, ctypes.byref(nbytes)), "workspace")
        ws = torch.empty(int(nbytes.value), device=gpu, dtype=torch.uint8)
        g = g0.clone()
        d = torch.empty(plan.vol_shape, device=gpu)
        log = torch.zeros(4, device=gpu, dtype=torch.float64)
        cnt = torch.zeros(1, device=gpu, dtype=torch.int32)
        for _ in range(3):
            _lib.check(lib.vamx_bclp_step(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(),
                                          T.data_ptr(), 0, 0, d.data_ptr(), log.data_ptr(), cnt.data_ptr(),
                                          ws.data_ptr(), st), "step")
        torch.cuda.synchronize()
        return g, log[:3].cpu().numpy()

    kw = dict(mu=0.01, det_offset=3.0, device=gpu)
    paired = ParallelPlan(n, nz, ang, **kw)
    unpaired = ParallelPlan(n, nz, ang, **kw)
    unpaired.set_tuning(fp_pair=0)
    lane = ParallelPlan(n, nz, ang, per_lane_geometry=True, **kw)
    gp, lp = run(paired)
    gu, lu = run(unpaired)
    gl, ll = run(lane)
    assert not torch.equal(gp, g0)
    assert torch.equal(gp, gu) and np.array_equal(lp, lu)
    e = rel_l2(gp.cpu().numpy(), gl.cpu().numpy())
    print("exact-gradient step: wave vs per-lane rel L2 %.2e, losses" % e, lp)
    assert e <= 1e-6 and np.allclose(lp, ll, rtol=1e-9, atol=0)
    insert = np.zeros((n, n, nz))
    insert[n // 2 - 2:n // 2 + 2, n // 2 - 2:n // 2 + 2, :] = 1.0
    occ = ParallelPlan(n, nz, ang, occlusion=insert, device=gpu)
    assert occ.occluded
    with pytest.raises(RuntimeError, match="occlusion"):
        run(occ)


def _on_levels_close(a, ref, bits, top, max_frac, max_levels=1):
    """Quantised iterates: float32 rounding moves a value across a level
    boundary now and then, so at most max_frac of the entries may differ, and
    each by at most max_levels level steps; all others agree to 1e-5 of the
    top."""
    a, ref = np.asarray(a, np.float64).ravel(), np.asarray(ref, np.float64).ravel()
    step = top / (2 ** bits - 1)
    d = np.abs(a - ref)
    off = d > 1e-5 * top
    print("  %d of %d entries one level apart (max diff %.3g, level %.3g)" % (off.sum(), a.size, d.max(), step))
    assert off.mean() <= max_frac
    assert d.max() <= (max_levels + 0.01) * step + 1e-5 * top


def test_discretize_on_device_matches_oracle(gpu):
    """util/data.py:495-523 on a device tensor: bucketise against the float64
    levels, index -1 (below the range) wrapping to the top level as NumPy's
    indexing does; bit for bit equal to the oracle's levels."""
    import torch
    import oracle.optim as oo
    from vamtoolbox_amd.util import data
    rng = np.random.default_rng(3)
    lo, hi = 0.0, 1.7
    lev = np.linspace(lo, hi, 2 ** 8)
    x = np.concatenate([rng.uniform(-0.2, 1.9, 100000), lev, [lo, hi, -1e-30, hi * 2]]).astype(np.float32)
    for bits in (1, 4, 8, 12):
        got = data.discretize(torch.as_tensor(x, device="cuda"), bits, [lo, hi]).cpu().numpy()
        ref = oo.discretize(x.astype(np.float64), bits, lo, hi).astype(np.float32)
        np.testing.assert_array_equal(got, ref)
    # range given as device scalars, as the loops pass torch.amin / torch.amax
    t = torch.as_tensor(x, device="cuda")
    got = data.discretize(t, 8, [torch.amin(t), torch.amax(t)]).cpu().numpy()
    ref = oo.discretize(x.astype(np.float64), 8, float(x.min()), float(x.max())).astype(np.float32)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("method", ["OSMO", "CAL", "FBP", "BCLP"])
def test_bit_depth_matches_oracle(gpu, method):
    """The optimisers with bit_depth = 8 against the oracle loops with the
    same quantisation (OSMO.py:61-64,99-102,164-165, CAL.py:171-173,
    FBP.py:56-57, BCLP.py:528-533).  Quantisation inside a loop makes it
    discontinuous: a value that a rounding moves across a level boundary
    changes the next projection, which moves more values across boundaries.
    On this problem the float64 oracle itself, with relative noise of 3e-6 on
    its projections, ends 4 OSMO iterations with 1.5 % of the sinogram
    entries on another level.  The loops are therefore compared over their
    first quantised iterations (OSMO: one void and one gel step)."""
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    bits = 8
    if method == "BCLP":
        tg, pg, ang, rm, ro, g0, fwd, bwd, lr = _bclp_setup(gpu, "analytical")
        n_iter = 3
        opts = vam.optimize.Options(method="BCLP", n_iter=n_iter, response_model=rm, eps=0.05, p=2, q=1,
                                    learning_rate=lr, g0=g0, bit_depth=bits)
        g, dose, mapped, logs = vam.optimize.optimize(tg, pg, opts, output="components")
        g_ref, dose_ref, _, loss_ref = oo.bclp(tg.array, fwd, bwd, ro, n_iter, g0, eps=0.05, lr=lr, bit_depth=bits)
        print("BCLP bit depth 8: rel L2 g %.2e dose %.2e" % (rel_l2(g, g_ref), rel_l2(dose, dose_ref)))
        _on_levels_close(g, g_ref, bits, g_ref.max(), 0.005)
        assert rel_l2(dose, dose_ref) <= 1e-3
        # the fused step followed by the constraints is the unfused loop: the
        # same iterates bit for bit (the loss sums in another order)
        opts_u = vam.optimize.Options(method="BCLP", n_iter=n_iter, response_model=rm, eps=0.05, p=2, q=1,
                                      learning_rate=lr, g0=g0, bit_depth=bits, fused=False)
        g_u, dose_u, _, logs_u = vam.optimize.optimize(tg, pg, opts_u, output="components")
        assert np.array_equal(np.asarray(g), np.asarray(g_u))
        np.testing.assert_allclose(logs.loss[:n_iter], logs_u.loss[:n_iter], rtol=1e-12, atol=0)
        return
    tg, pg, ang = _problem(seed=4)
    fwd, bwd = _oracle_ops(ang)
    if method == "OSMO":
        opts = vam.optimize.Options(method="OSMO", n_iter=2, d_h=0.85, d_l=0.6, bit_depth=bits)
        b_ref, x_ref, _ = oo.osmo(tg.array, fwd, bwd, 2, d_h=0.85, d_l=0.6, bit_depth=bits)
    elif method == "CAL":
        opts = vam.optimize.Options(method="CAL", n_iter=3, learning_rate=0.002, sigmoid=0.01, bit_depth=bits)
        b_ref, x_ref, _ = oo.cal(tg.array, fwd, bwd, 3, lr=0.002, sigmoid=0.01, bit_depth=bits)
    else:
        opts = vam.optimize.Options(method="FBP", bit_depth=bits)
        b_ref, x_ref, _ = oo.fbp(tg.array, fwd, bwd, bit_depth=bits)
    sino, recon, _ = vam.optimize.optimize(tg, pg, opts)
    print("%s bit depth 8: rel L2 sinogram %.2e recon %.2e" % (method, rel_l2(sino.array, b_ref),
                                                              rel_l2(recon.array, x_ref)))
    top = b_ref.max()
    # OSMO re-quantises the [min, max]-level projection on [0, max] levels at
    # the end: one boundary crossing inside the loop can show as up to 3 levels
    # (so does the float64 oracle with 1e-6 noise on its projections)
    _on_levels_close(sino.array, b_ref, bits, top, 0.005, 3 if method == "OSMO" else 1)
    assert rel_l2(recon.array, x_ref) <= 1e-3


@pytest.mark.parametrize("kind", ["glf", "glf_nu", "linear", "identity", "interpolation"])
def test_response_map_autograd(gpu, kind):
    """ResponseMap (vamx_response_map): map and map' of response.py:191-291 on
    the device against the oracle's float64 response, and its autograd
    backward grad * map'(x)."""
    import torch
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.autograd import ResponseMap
    if kind == "glf":
        rm, ro = vam.response.ResponseModel(type="analytical", B=20, M=0.6), oo.Response("analytical", B=20, M=0.6)
    elif kind == "glf_nu":
        kw = dict(A=0.1, K=0.9, B=12, M=0.4, nu=2.5)
        rm, ro = vam.response.ResponseModel(type="analytical", **kw), oo.Response("analytical", **kw)
    elif kind == "linear":
        rm = vam.response.ResponseModel(type="analytical", form="linear", M=1.7, C=-0.2)
        ro = None
    elif kind == "identity":
        rm = vam.response.ResponseModel(type="analytical", form="identity")
        ro = None
    else:
        rm, ro = vam.response.ResponseModel(B=25), oo.Response("interpolation", B=25)
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(-0.5, 1.5, 50000), np.linspace(0, 1, 512), [0.0, 1.0, -3.0, 4.0]])
    x = x.astype(np.float32).astype(np.float64)
    if ro is not None:
        m_ref, d_ref = ro.map(x), ro.dmap(x)
    elif kind == "linear":
        m_ref, d_ref = 1.7 * x - 0.2, np.full_like(x, 1.7)
    else:
        m_ref, d_ref = x.copy(), np.ones_like(x)
    xt = torch.tensor(x, dtype=torch.float32, device=gpu, requires_grad=True)
    m = ResponseMap.apply(xt, rm)
    w = rng.uniform(-1, 1, x.size)
    (m * torch.tensor(w, dtype=torch.float32, device=gpu)).sum().backward()
    m_np, g_np = m.detach().cpu().numpy(), xt.grad.cpu().numpy()
    e_m = np.max(np.abs(m_np - m_ref) / np.maximum(np.abs(m_ref), 1e-3))
    e_g = np.max(np.abs(g_np - w * d_ref) / np.maximum(np.abs(w * d_ref), 1e-3))
    print("%s: map max rel %.2e, grad max rel %.2e" % (kind, e_m, e_g))
    # float64 evaluation, float32 storage: one rounding of the result
    assert e_m <= 2e-7 and e_g <= 5e-7


@pytest.mark.parametrize("method", ["OSMO", "CAL", "FBP", "BCLP"])
def test_zero_dose_loops_match_oracle(gpu, method):
    """A zero-dose body (projectorconstructor.py:162-163): the reference's
    back projection zeroes the sinogram's rays through it in place, so its
    loops iterate on, and return, the zeroed sinogram.  The optimisers here
    against the oracle loops that zero those rays before each back
    projection; the returned sinogram is exactly zero on them.  BCLP runs
    its fused step (the iterate zeroed on those rays after every step) and
    agrees with its unfused loop."""
    import oracle.d1 as d1
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    n, nz, na = 32, 4, 60
    ang = np.linspace(0, 360, na, endpoint=False)
    t = _ellipsoids(n, nz, 7, gray=(method == "BCLP"))
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    zd = np.repeat(((xx - 3) ** 2 + (yy + 2) ** 2 <= 9)[:, :, None], nz, axis=2).astype(float)
    t = np.where(zd > 0, 0.0, t)
    tg = vam.geometry.TargetGeometry(target=t, zero_dose=zd)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True)
    fwd, bwd = _oracle_ops(ang)
    mask = d1.forward(zd, ang) != 0
    if method == "BCLP":
        rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
        ro = oo.Response("analytical", B=20, M=0.6)
        g0 = 0.8 * oo.bclp_init(tg.array, fwd, oo.Response("interpolation", B=20, M=0.6), na)
        opts = vam.optimize.Options(method="BCLP", n_iter=10, response_model=rm, eps=0.05, p=2, q=1,
                                    learning_rate=1e-4, g0=g0)
        g, dose, _, logs = vam.optimize.optimize(tg, pg, opts, output="components")
        g_ref, dose_ref, _, _ = oo.bclp(tg.array, fwd, bwd, ro, 10, g0, eps=0.05, lr=1e-4, zero_sino=mask)
        opts_u = vam.optimize.Options(method="BCLP", n_iter=10, response_model=rm, eps=0.05, p=2, q=1,
                                      learning_rate=1e-4, g0=g0, fused=False)
        g_u, dose_u, _, logs_u = vam.optimize.optimize(tg, pg, opts_u, output="components")
        e_fu = rel_l2(np.asarray(g), np.asarray(g_u))
        print("BCLP zero-dose fused vs unfused rel L2 g %.2e" % e_fu)
        assert e_fu <= 1e-6 and np.allclose(logs.loss, logs_u.loss, rtol=1e-6, atol=0)
        s_gpu, s_ref, x_gpu, x_ref = np.asarray(g).reshape(g_ref.shape), g_ref, dose, dose_ref
    else:
        if method == "OSMO":
            opts = vam.optimize.Options(method="OSMO", n_iter=6, d_h=0.85, d_l=0.6)
            s_ref, x_ref, _ = oo.osmo(tg.array, fwd, bwd, 6, d_h=0.85, d_l=0.6, zero_sino=mask)
        elif method == "CAL":
            opts = vam.optimize.Options(method="CAL", n_iter=6, learning_rate=0.002, sigmoid=0.01)
            s_ref, x_ref, _ = oo.cal(tg.array, fwd, bwd, 6, lr=0.002, sigmoid=0.01, zero_sino=mask)
        else:
            opts = vam.optimize.Options(method="FBP")
            s_ref, x_ref, _ = oo.fbp(tg.array, fwd, bwd, zero_sino=mask)
        sino, recon, _ = vam.optimize.optimize(tg, pg, opts)
        s_gpu, x_gpu = sino.array, recon.array
    e_s, e_x = rel_l2(s_gpu, s_ref), rel_l2(x_gpu, x_ref)
    print("%s with a zero-dose body: rel L2 sinogram %.2e dose %.2e; %d of %d rays zeroed" % (
        method, e_s, e_x, mask.sum(), mask.size))
    assert np.array_equal(np.asarray(pg.zero_dose_sino), mask)
    assert mask.sum() > 0 and np.all(np.asarray(s_gpu)[mask] == 0)
    assert e_s <= 1e-4 and e_x <= 1e-4


def test_osmo_elementwise_kernels_equal_torch(gpu):
    """vamx_osmo_model_update and vamx_scale_clamp equal the torch expressions
    of OSMO.py:54-56, 92-94 and 60-69 bit for bit, NaN included."""
    import torch
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    st = _lib.stream_ptr(gpu)
    g = torch.Generator(device=gpu).manual_seed(4)
    n = 100003
    x = torch.rand(n, device=gpu, generator=g) * 1.4 - 0.2
    x[17] = float("nan")
    xm0 = torch.rand(n, device=gpu, generator=g) - 0.5
    mask = torch.rand(n, device=gpu, generator=g) > 0.4
    for d, gel in ((0.7, 0), (0.8, 1)):
        xm = xm0.clone()
        _lib.check(lib.vamx_osmo_model_update(x.data_ptr(), xm.data_ptr(), mask.data_ptr(), n, d, gel, st), "update")
        ref = (torch.where(mask, xm0 + torch.clamp(d - x, min=0.0), xm0) if gel else
               torch.where(mask, xm0 - torch.clamp(x - d, min=0.0), xm0))
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(xm, nan=7.0), torch.nan_to_num(ref, nan=7.0))
        assert bool(torch.isnan(xm[17])) == bool(mask[17])
    for lo in (0.0, 0.05):
        v = x.clone()
        m = torch.amax(torch.nan_to_num(x, nan=0.0))
        _lib.check(lib.vamx_scale_clamp(v.data_ptr(), n, m.data_ptr(), lo, st), "scale_clamp")
        ref = torch.clamp(x / m, min=lo)
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(v, nan=7.0), torch.nan_to_num(ref, nan=7.0)) and bool(torch.isnan(v[17]))


def test_segmented_scale_clamp_equals_per_segment_torch(gpu):
    """vamx_scale_clamp_segments equals torch's amax / div / clamp on every
    segment alone, bit for bit: uneven and odd segment sizes, an empty
    segment, a segment holding a NaN (all NaN after), one of negatives only
    (its maximum negative), and a second call with fewer, smaller segments
    that reuses the workspace; 16-byte aligned and misaligned buffers."""
    import torch
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    st = _lib.stream_ptr(gpu)
    g = torch.Generator(device=gpu).manual_seed(6)
    work = torch.empty(16, dtype=torch.int32, device=gpu)
    for sizes in ((70001, 3, 0, 131072, 1, 5000, 257, 9999), (1000, 17, 4)):
        n = sum(sizes)
        x = torch.rand(n, device=gpu, generator=g) * 3.0 - 0.5
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        planted = len(sizes) == 8
        if planted:
            x[off[3] + 77] = float("nan")             # segment 3
            x[off[6]:off[7]] = -x[off[6]:off[7]].abs() - 0.1   # segment 6: negatives only
        seg = torch.as_tensor(off, device=gpu)
        for lo, shift in ((float("-inf"), 0), (0.0, 0), (0.05, 0), (0.05, 1)):
            # shift 1: a buffer 4 bytes past a 16-byte boundary
            v = torch.empty(n + shift, device=gpu)[shift:]
            v.copy_(x)
            _lib.check(lib.vamx_scale_clamp_segments(v.data_ptr(), n, seg.data_ptr(), len(sizes), lo, work.data_ptr(),
                                                     st), "scale_clamp_segments")
            ref = x.clone()
            for s in range(len(sizes)):
                r = ref[off[s]:off[s + 1]]
                if r.numel():
                    r.copy_(torch.clamp(r / torch.amax(r), min=lo))
            torch.cuda.synchronize()
            assert torch.equal(torch.nan_to_num(v, nan=7.0), torch.nan_to_num(ref, nan=7.0)), (sizes, lo)
            assert torch.equal(torch.isnan(v), torch.isnan(ref))
            if planted:
                assert bool(torch.isnan(v[off[3]:off[4]]).all())


def test_ver_kernel_equals_numpy(gpu):
    """calcVER on a device dose (vamx_ver: the least gel dose, then the count
    of void voxels at or above it) equals the NumPy path (metrics.py:47-87),
    with ties at the least gel dose, signed zeros and a NaN gel dose (VER 0,
    as np.min and >= give)."""
    import torch
    import vamtoolbox_amd as vam
    from tests.test_gpu_optim import _ellipsoids
    rng = np.random.default_rng(3)
    # 40 x 40 x 12 voxels: four per step; 33 x 33 x 7 (odd count): one per step
    for case, (n, nz) in enumerate([(40, 12)] * 4 + [(33, 7)] * 4):
        t = _ellipsoids(n, nz, seed=9)
        tg = vam.geometry.TargetGeometry(target=t)
        case %= 4
        d = rng.random(t.shape).astype(np.float32)
        gel, void = np.asarray(tg.gel_inds), np.asarray(tg.void_inds)
        d[gel] = np.float32(0.5) + np.float32(0.5) * d[gel]   # about half of the void voxels count
        if case == 1:   # ties: void voxels equal to the least gel dose, and -0 / +0
            vi = np.argwhere(void)
            d[tuple(vi[:50].T)] = d[gel].min()
            d[tuple(vi[50:60].T)] = -0.0
        if case == 2:
            d[tuple(np.argwhere(gel)[3])] = np.nan
        if case == 3:
            d[gel] = -0.0
        ref = vam.metrics.calcVER(tg, d)
        got = vam.metrics.calcVER(tg, torch.tensor(d, device=gpu))
        print("case %d: VER %.6f vs %.6f" % (case, got, ref))
        assert got == ref
