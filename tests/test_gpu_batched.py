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


def _filling_set(gpu, seed):
    """three problems that each fill the GPU alone (their forwards run as
    single-problem launches on the caller's stream and two side streams) and
    three small ones (persistent launch)"""
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(seed)
    specs = [(128, 64, 360, 0.01), (128, 64, 360, 0.0), (120, 77, 360, 0.02), (32, 8, 30, 0.0), (24, 40, 45, 0.01),
             (20, 3, 20, 0.0)]
    plans, vols, sinos = [], [], []
    for n, nz, na, mu in specs:
        p = ParallelPlan(n, nz, np.linspace(0, 360, na, endpoint=False), mu=mu, device=gpu)
        plans.append(p)
        vols.append(torch.tensor(rng.random((n, n, nz)), dtype=torch.float32, device=gpu))
        sinos.append(torch.tensor(rng.random((n, na, nz)), dtype=torch.float32, device=gpu))
    return plans, vols, sinos


def test_batched_side_streams_equal_individual(gpu):
    """The event fork / join path (three GPU-filling forwards) equals the
    per-problem forwards bitwise, through the plan-list entry points and
    through an explicit batch."""
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.batch import ParallelBatch
    plans, vols, sinos = _filling_set(gpu, 41)
    lib = _lib.load()
    k = len(plans)
    ref_s = [p.forward(v) for p, v in zip(plans, vols)]
    ref_b = [p.backward(s) for p, s in zip(plans, sinos)]
    outs = [torch.full_like(r, float("nan")) for r in ref_s]
    H = (ctypes.c_void_p * k)(*[p.handle.value for p in plans])
    IN = (ctypes.c_void_p * k)(*[v.data_ptr() for v in vols])
    OUT = (ctypes.c_void_p * k)(*[o.data_ptr() for o in outs])
    _lib.check(lib.vamx_par_forward_batched(H, k, IN, OUT, _lib.stream_ptr(gpu)), "forward_batched")
    torch.cuda.synchronize()
    for i in range(k):
        assert torch.equal(outs[i], ref_s[i]), i
    B = ParallelBatch(plans)
    s = B.forward(B.pack_vols(vols))
    x = B.backward(torch.cat([t.reshape(-1) for t in sinos]))
    torch.cuda.synchronize()
    for i, (sv, xv) in enumerate(zip(B.sino_views(s), B.vol_views(x))):
        assert torch.equal(sv, ref_s[i]), i
        assert torch.equal(xv, ref_b[i]), i


def test_batched_repeated_plan_rejected(gpu):
    """A plan may appear once per batch: its slice min/max scratch serves one problem."""
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.batch import ParallelBatch
    from vamtoolbox_amd.parallel import ParallelPlan
    p = ParallelPlan(24, 8, np.linspace(0, 360, 30, endpoint=False), device=gpu)
    with pytest.raises(ValueError):
        ParallelBatch([p, p])
    lib = _lib.load()
    a = torch.rand(p.vol_shape, device=gpu)
    b = torch.rand(p.vol_shape, device=gpu)
    oa, ob = torch.empty(p.sino_shape, device=gpu), torch.empty(p.sino_shape, device=gpu)
    H = (ctypes.c_void_p * 2)(p.handle.value, p.handle.value)
    IN = (ctypes.c_void_p * 2)(a.data_ptr(), b.data_ptr())
    OUT = (ctypes.c_void_p * 2)(oa.data_ptr(), ob.data_ptr())
    assert lib.vamx_par_forward_batched(H, 2, IN, OUT, _lib.stream_ptr(gpu)) == -1
    assert b"repeats" in lib.vamx_last_error()
    h = ctypes.c_void_p()
    assert lib.vamx_batch_create(H, 2, ctypes.byref(h)) == -1


def test_batched_concurrent_batches_on_two_streams(gpu):
    """Two batches on two streams at once: each has its own tile counter, so
    neither drains the other's queue."""
    import torch
    from vamtoolbox_amd.batch import ParallelBatch
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(43)
    sets = []
    for base in (0, 1):
        plans = [ParallelPlan(n, nz, np.linspace(0, 360, na, endpoint=False), device=gpu)
                 for n, nz, na in [(48 + 8 * base, 32, 90), (32, 16, 60), (40, 8, 45)]]
        vols = [torch.tensor(rng.random(p.vol_shape), dtype=torch.float32, device=gpu) for p in plans]
        B = ParallelBatch(plans)
        sets.append((plans, vols, B, B.pack_vols(vols), [p.forward(v) for p, v in zip(plans, vols)]))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=gpu) for _ in sets]
    outs = [B.empty_sino() for _, _, B, _, _ in sets]
    for rep in range(3):
        for st, (_, _, B, flat, _), o in zip(streams, sets, outs):
            st.wait_stream(torch.cuda.current_stream(gpu))
            with torch.cuda.stream(st):
                B.forward(flat, out=o)
        torch.cuda.synchronize()
        for (_, _, B, _, ref), o in zip(sets, outs):
            for i, v in enumerate(B.sino_views(o)):
                assert torch.equal(v, ref[i]), (rep, i)


def test_batched_graph_survives_other_batches(gpu):
    """A graph that captured a batched forward keeps replaying correctly after
    many other batches ran on the thread (each batch owns its tables)."""
    import torch
    from vamtoolbox_amd.batch import ParallelBatch
    from vamtoolbox_amd.parallel import ParallelPlan
    plans, vols, _ = _filling_set(gpu, 47)
    B = ParallelBatch(plans)
    flat = B.pack_vols(vols)
    out = B.empty_sino()
    B.forward(flat, out=out)           # uploads the tables eagerly
    ref = out.clone()
    s = torch.cuda.Stream(device=gpu)
    s.wait_stream(torch.cuda.current_stream(gpu))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            B.forward(flat, out=out)
    torch.cuda.current_stream(gpu).wait_stream(s)
    for rep in range(3):
        # twenty other batches, each with its own tables
        for j in range(20):
            ps = [ParallelPlan(16 + j % 5, 4 + j % 3, np.linspace(0, 180, 10 + j, endpoint=False), device=gpu)]
            Bo = ParallelBatch(ps)
            Bo.forward(Bo.pack_vols([np.ones(ps[0].vol_shape)]))
            Bo.close()
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep


def test_batched_osmo_graph_with_side_streams(gpu):
    """OSMOBatch capture and replay over a set whose forwards use the side
    streams equals the eager iteration bitwise."""
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.batch import OSMOBatch
    from tests.test_gpu_optim import _ellipsoids
    specs = [(128, 64, 360, 0.01), (128, 64, 360, 0.0), (120, 77, 360, 0.02), (32, 8, 30, 0.0), (24, 40, 45, 0.01)]
    tgs = [vam.geometry.TargetGeometry(target=_ellipsoids(n, nz, seed=50 + i)) for i, (n, nz, _, _) in enumerate(specs)]
    pgs = [vam.geometry.ProjectionGeometry(np.linspace(0, 360, na, endpoint=False), "parallel", CUDA=True,
                                           absorption_per_voxel=mu) for _, _, na, mu in specs]
    opts = vam.optimize.Options(method="OSMO", n_iter=4, d_h=0.85, d_l=0.6, filter="ram-lak")
    eager = OSMOBatch(tgs, pgs, opts, device=gpu)
    graph = OSMOBatch(tgs, pgs, opts, device=gpu)
    eager.step()
    graph.step()
    graph.capture()
    for it in range(3):
        eager.step()
        torch.cuda.synchronize()
        graph.graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(eager.x, graph.x), it
        assert torch.equal(eager.b, graph.b), it
