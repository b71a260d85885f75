"""One persistent launch over uneven problems equals the per-problem launches bitwise, and
batched OSMO equals OSMO on each problem."""

import ctypes

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("span", [180, 360])
def test_batched_equals_individual(gpu, span):
    """Over 360 deg the wavefront forward pairs opposite angles, in the batch
    as in the single-problem kernel; over 180 deg nothing pairs."""
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(11)
    # thin z (per-lane kernels) and z ranges the wavefront kernels take (full and partial chunks,
    # odd nz with lanes holding z and z + 64)
    # (the last problem fills the GPU alone: its forward runs as its own launch)
    specs = [(32, 8, 30, 0.0), (48, 5, 45, 0.01), (24, 12, 90, 0.0), (40, 3, 20, 0.02), (16, 1, 60, 0.0),
             (32, 64, 40, 0.01), (24, 100, 30, 0.0), (21, 40, 25, 0.02), (18, 256, 12, 0.0), (26, 77, 36, 0.01),
             (19, 201, 20, 0.0), (128, 64, 360, 0.01)]
    plans, vols, sinos, ref_s, bps, ref_b = [], [], [], [], [], []
    for n, nz, na, mu in specs:
        ang = np.linspace(0, span, na, endpoint=False)
        p = ParallelPlan(n, nz, ang, mu=mu, device=gpu)
        v = torch.tensor(rng.random((n, n, nz)), dtype=torch.float32, device=gpu)
        s = torch.tensor(rng.random((n, na, nz)), dtype=torch.float32, device=gpu)
        plans.append(p)
        vols.append(v)
        sinos.append(torch.empty((n, na, nz), device=gpu))
        ref_s.append(p.forward(v))
        bps.append((s, torch.empty((n, n, nz), device=gpu)))
        ref_b.append(p.backward(s))
    lib = _lib.load()
    k = len(plans)
    H = (ctypes.c_void_p * k)(*[p.handle.value for p in plans])
    IN = (ctypes.c_void_p * k)(*[v.data_ptr() for v in vols])
    OUT = (ctypes.c_void_p * k)(*[s.data_ptr() for s in sinos])
    _lib.check(lib.vamx_par_forward_batched(H, k, IN, OUT, _lib.stream_ptr(gpu)), "forward_batched")
    IN2 = (ctypes.c_void_p * k)(*[s.data_ptr() for s, _ in bps])
    OUT2 = (ctypes.c_void_p * k)(*[o.data_ptr() for _, o in bps])
    _lib.check(lib.vamx_par_backward_batched(H, k, IN2, OUT2, _lib.stream_ptr(gpu)), "backward_batched")
    torch.cuda.synchronize()
    for i in range(k):
        assert torch.equal(sinos[i], ref_s[i]), i
        assert torch.equal(bps[i][1], ref_b[i]), i


def _osmo_problems():
    import vamtoolbox_amd as vam
    from tests.test_gpu_optim import _ellipsoids
    specs = [(24, 3, 30, 0.0), (40, 5, 45, 0.01), (32, 8, 60, 0.0), (16, 2, 20, 0.02)]
    tgs, pgs = [], []
    for s, (n, nz, na, mu) in enumerate(specs):
        tgs.append(vam.geometry.TargetGeometry(target=_ellipsoids(n, nz, seed=20 + s)))
        pgs.append(vam.geometry.ProjectionGeometry(np.linspace(0, 360, na, endpoint=False), "parallel", CUDA=True,
                                                   absorption_per_voxel=mu))
    return tgs, pgs


def test_batched_osmo_matches_single_problems(gpu):
    """Batched OSMO (graph-replayed) = OSMO on each problem alone."""
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.batch import minimizeOSMOBatched
    tgs, pgs = _osmo_problems()
    opts = vam.optimize.Options(method="OSMO", n_iter=6, d_h=0.85, d_l=0.6, filter="ram-lak")
    res = minimizeOSMOBatched(tgs, pgs, opts, hip_graph=True)
    for i, (tg, pg) in enumerate(zip(tgs, pgs)):
        sino, recon, err = vam.optimize.optimize(tg, pg, vam.optimize.Options(method="OSMO", n_iter=6, d_h=0.85,
                                                                             d_l=0.6, filter="ram-lak"))
        b, x, e = res[i]
        eb, ex = rel_l2(b.array, sino.array), rel_l2(x.array, recon.array)
        print("problem %d: rel L2 sinogram %.2e recon %.2e" % (i, eb, ex))
        assert eb <= 1e-6 and ex <= 1e-6
        np.testing.assert_allclose(e, err, atol=1e-6)


def test_batched_osmo_autotuned_plans(gpu):
    """With autotune every problem's plan takes its own fastest work order;
    the batched projections agree with the default work order to float32
    rounding (tile shapes are bitwise, pairing sums in another order)."""
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.batch import OSMOBatch
    from tests.test_gpu_optim import _ellipsoids
    specs = [(40, 64, 60, 0.01), (36, 77, 45, 0.0), (28, 130, 30, 0.02)]
    tgs = [vam.geometry.TargetGeometry(target=_ellipsoids(n, nz, seed=30 + i)) for i, (n, nz, _, _) in enumerate(specs)]
    pgs = [vam.geometry.ProjectionGeometry(np.linspace(0, 360, na, endpoint=False), "parallel", CUDA=True,
                                           absorption_per_voxel=mu) for _, _, na, mu in specs]
    opts = vam.optimize.Options(method="OSMO", n_iter=3, d_h=0.85, d_l=0.6, filter="ram-lak")
    tuned = OSMOBatch(tgs, pgs, opts, device=gpu, autotune=True)
    plain = OSMOBatch(tgs, pgs, opts, device=gpu)
    for p in tuned.batch.plans:
        print(p.get_tuning())
    eb = rel_l2(tuned.b.cpu().numpy(), plain.b.cpu().numpy())
    ex = rel_l2(tuned.x.cpu().numpy(), plain.x.cpu().numpy())
    print("autotuned vs default: sinogram %.2e recon %.2e" % (eb, ex))
    assert eb <= 1e-6 and ex <= 1e-6


def test_batched_osmo_graph_replays_equal_eager(gpu):
    """Every replay of the captured iteration equals the eager iteration, with
    eager work and host synchronisation between the replays."""
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.batch import OSMOBatch
    tgs, pgs = _osmo_problems()
    opts = vam.optimize.Options(method="OSMO", n_iter=6, d_h=0.85, d_l=0.6, filter="ram-lak")
    eager = OSMOBatch(tgs, pgs, opts, device=gpu)
    graph = OSMOBatch(tgs, pgs, opts, device=gpu)
    eager.step()
    graph.step()
    graph.capture()
    for it in range(4):
        eager.step()
        torch.cuda.synchronize()
        graph.graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(eager.x, graph.x), it
        assert torch.equal(eager.b, graph.b), it
        assert torch.equal(eager.x_model, graph.x_model), it
