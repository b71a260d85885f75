"""Benchmark: fused BCLP iterations per second at 256^3 voxels x 360 angles on one MI355X.

Workload (the BCLP configuration): 256^3 float32 grayscale target (union of 8
seeded random ellipsoids, semi-axes 15-70 voxels, Gaussian-smoothed with
sigma = 2 voxels, clipped to [0, 1] and to the inscribed cylinder), 360
angles (0..359 deg), 256 x 256 detector, exponential absorption 0.01 per voxel
inside the cylinder, 10-voxel detector offset, zero inclination; analytical
sigmoid response (threshold 0.6, steepness 20), band +-0.05, p = 2, q = 1,
sinogram >= 0, learning rate 0.01.  The initial sinogram is the ramp-filtered
forward projection of the target, clipped at 0 and scaled so that the mean
dose inside the target is 0.7.  A step is one BCLP iteration (back projection of the
sinogram, response / band / Lp epilogue and loss, forward projection of the
operand fused with the sinogram update), replayed from a captured HIP graph.

Besides the JSON line's required fields it reports: the same iteration as a
plain unfused GPU loop (HIP projectors + torch element-wise ops + the
reference's host-side loss branch) and as a plain PyTorch loop without custom
kernels, the relative sinogram / dose error of the fused loop against the
unfused one, the forward / back projection kernel times, and the float64
NumPy port of the reference CPU projector timed on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

N, NZ, NA = 256, 256, 360
MU, OFFSET = 0.01, 10.0
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# synthetic target
# ---------------------------------------------------------------------------
def make_target(device, seed=3):
    import torch
    rng = np.random.default_rng(seed)
    c = (N - 1) / 2
    ax = torch.arange(N, device=device, dtype=torch.float32) - c
    zz = (torch.arange(NZ, device=device, dtype=torch.float32) - (NZ - 1) / 2)
    Y, X, Z = torch.meshgrid(ax, ax, zz, indexing="ij")
    vol = torch.zeros((N, N, NZ), device=device)
    for _ in range(8):
        a, b, cc = rng.uniform(15, 70, 3)
        ctr = rng.uniform(-50, 50, 3)
        R = _rotation(rng)
        d = torch.stack([X - ctr[0], Y - ctr[1], Z - ctr[2]], dim=-1)
        Rt = torch.tensor(R, device=device, dtype=torch.float32)
        u = d @ Rt.T
        inside = (u[..., 0] / a) ** 2 + (u[..., 1] / b) ** 2 + (u[..., 2] / cc) ** 2 <= 1
        vol = torch.maximum(vol, inside.float())
    # Gaussian smoothing, sigma = 2 voxels (separable, truncated at 4 sigma)
    r = torch.arange(-8, 9, device=device, dtype=torch.float32)
    k = torch.exp(-r ** 2 / (2 * 2.0 ** 2))
    k = k / k.sum()
    v = vol[None, None]
    for dim in range(3):
        shape = [1, 1, 1, 1, 1]
        shape[2 + dim] = 17
        pad = [0, 0, 0, 0, 0, 0]
        pad[2 * (2 - dim)] = 8
        pad[2 * (2 - dim) + 1] = 8
        v = torch.nn.functional.conv3d(torch.nn.functional.pad(v, pad), k.reshape(shape))
    vol = v[0, 0].clamp(0, 1)
    from vamtoolbox_amd.util.data import clipToCircle
    return clipToCircle(vol).contiguous()


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------------------
# plain PyTorch projector (no custom kernels): the baseline GPU port
# ---------------------------------------------------------------------------
class TorchD1:
    """Rotate-and-sum forward / np.interp back projection with torch ops only."""

    def __init__(self, n, nz, angles_deg, device, mu=0.0, offset=0.0, chunk=6):
        import torch
        self.n, self.nz, self.dev, self.chunk = n, nz, device, chunk
        self.c = n // 2
        th = torch.tensor(np.deg2rad(angles_deg), device=device, dtype=torch.float64)
        self.co, self.sn = torch.cos(th), torch.sin(th)
        self.mu, self.off = mu, offset
        self.R = (n - 1) / 2.0
        kk = torch.arange(n, device=device, dtype=torch.float64)
        self.t = (kk - self.c) - offset
        self.s = kk - self.c
        yy, xx = torch.meshgrid(kk - self.c, kk - self.c, indexing="ij")
        self.x, self.y = xx, yy
        from vamtoolbox_amd.util.data import outside_mask
        self.out_mask = torch.as_tensor(outside_mask(n, n), device=device)

    def _abs(self, t, s):
        import torch
        if self.mu == 0:
            return torch.ones_like(t + s)
        inside = t.abs() < self.R
        half = torch.sqrt(torch.clamp(self.R ** 2 - t * t, min=0))
        return torch.where(inside, torch.exp(-self.mu * torch.clamp(half - s, min=0)), torch.zeros_like(t + s))

    def forward(self, vol):
        import torch
        n, nz = self.n, self.nz
        flat = vol.reshape(n * n, nz)
        out = torch.empty((n, self.co.numel(), nz), device=self.dev)
        T, S = self.t[None, :], self.s[:, None]          # (r, k)
        W = self._abs(T, S).float()
        for a0 in range(0, self.co.numel(), self.chunk):
            co, sn = self.co[a0:a0 + self.chunk, None, None], self.sn[a0:a0 + self.chunk, None, None]
            row = self.c - sn * T + co * S
            col = self.c + co * T + sn * S
            r0, c0 = torch.floor(row), torch.floor(col)
            dr, dc = (row - r0).float(), (col - c0).float()
            acc = 0
            for rr, wr in ((r0, 1 - dr), (r0 + 1, dr)):
                for cc, wc in ((c0, 1 - dc), (c0 + 1, dc)):
                    ok = (rr >= 0) & (rr < n) & (cc >= 0) & (cc < n)
                    idx = (rr.clamp(0, n - 1) * n + cc.clamp(0, n - 1)).long()
                    vals = flat.index_select(0, idx.reshape(-1)).reshape(idx.shape + (nz,))
                    acc = acc + (wr * wc * ok)[..., None] * vals
            out[:, a0:a0 + self.chunk, :] = (W[None, ..., None] * acc).sum(dim=1).permute(1, 0, 2)
        return out

    def backward(self, sino):
        import torch
        n, nz = self.n, self.nz
        out = torch.zeros((n, n, nz), device=self.dev)
        for i in range(self.co.numel()):
            t = self.x * self.co[i] - self.y * self.sn[i]
            s = self.x * self.sn[i] + self.y * self.co[i]
            u = t - (-self.c - self.off)
            ok = (u >= 0) & (u <= n - 1)
            j = torch.clamp(torch.floor(u), 0, n - 1).long()
            j1 = torch.clamp(j + 1, max=n - 1)
            w = torch.where(j == n - 1, torch.zeros_like(u), u - j).float()
            b = sino[:, i, :]
            v = b[j] + w[..., None] * (b[j1] - b[j])
            v = v * (ok * self._abs(t, s)).float()[..., None]
            out += v
        return torch.where(self.out_mask[..., None], torch.zeros((), device=self.dev), out)


# ---------------------------------------------------------------------------
# CPU baseline: float64 NumPy port of the reference's per-(z, angle) loops
# ---------------------------------------------------------------------------
def cpu_baseline(target_np, g_np, angles, k_slices=2):
    import oracle.d1 as d1
    import oracle.optim as oo
    resp = oo.Response("analytical", B=20, M=0.6)
    t0 = time.perf_counter()
    sl = slice(0, k_slices)
    dose = d1.backward(g_np[:, :, sl].astype(np.float64), angles, det_offset=OFFSET, mu=MU)
    m = resp.map(dose)
    eT = m - target_np[:, :, sl]
    eb = np.abs(eT) - 0.05
    v = eb > 0
    operand = v * np.abs(eb) * np.sign(eT) * resp.dmap(dose)
    d1.forward(operand, angles, det_offset=OFFSET, mu=MU)
    dt = time.perf_counter() - t0
    per_iter = dt * (NZ / k_slices)
    return {"value": 1.0 / per_iter, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": "float64 NumPy port of Projector3DParallelPython back+forward projection and the BCLP "
                      "epilogue on %d of %d z-slices (all %d angles), %.1f s measured, scaled x%d to one "
                      "iteration" % (k_slices, NZ, NA, dt, NZ // k_slices)}


def _cpu_slice_iteration(job):
    """One z-slice of the BCLP iteration on the CPU port (a worker of the
    all-core baseline): back projection, epilogue, forward projection."""
    import oracle.d1 as d1
    import oracle.optim as oo
    target_sl, g_sl, angles = job
    resp = oo.Response("analytical", B=20, M=0.6)
    dose = d1.backward(g_sl.astype(np.float64), angles, det_offset=OFFSET, mu=MU)
    eT = resp.map(dose) - target_sl
    eb = np.abs(eT) - 0.05
    operand = (eb > 0) * np.abs(eb) * np.sign(eT) * resp.dmap(dose)
    return dose, d1.forward(operand, angles, det_offset=OFFSET, mu=MU)


def cpu_baseline_all_cores(pool, workers, target_np, g_np, angles):
    """The z-slab variant: the iteration is independent per z-slice at zero
    inclination, so `workers` processes each run one slice of the port at
    once; the wall time of those slices is scaled to all NZ slices."""
    zs = [k * NZ // workers + NZ // (2 * workers) for k in range(workers)]   # spread over the volume
    jobs = [(target_np[:, :, z:z + 1], g_np[:, :, z:z + 1], angles) for z in zs]
    pool.map(_cpu_slice_iteration, jobs[:1])          # warm the workers' imports
    t0 = time.perf_counter()
    res = pool.map(_cpu_slice_iteration, jobs, chunksize=1)
    dt = time.perf_counter() - t0
    per_iter = dt * (NZ / workers)
    return {"value": 1.0 / per_iter, "unit": "iters/s", "cores": workers, "kind": "port", "slices": zs,
            "results": res,
            "sample": "float64 NumPy port of Projector3DParallelPython back+forward projection and the BCLP "
                      "epilogue, %d worker processes (one z-slice each, OMP_NUM_THREADS=1) over %d of %d z-slices "
                      "(all %d angles), %.1f s wall, scaled x%.0f to one iteration" % (
                          workers, workers, NZ, NA, dt, NZ / workers)}


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-torch-baseline", action="store_true")
    ap.add_argument("--lr", type=float, default=0.01, help="BCLP learning rate (the reference's default)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the candidate work orders first (bclp: the plan's; batched: every problem's plan); "
                         "the library defaults are the fastest measured at 256^3 x 360")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the forward / back projection kernels of the workload (5 launches each) "
                         "and a 1 GiB device copy, for hardware-counter profiling")
    ap.add_argument("--workload", choices=("bclp", "osmo", "batched", "raytrace", "imageseq"), default="bclp",
                    help="bclp (default): fused BCLP at 256^3 x 360; osmo: plain forward/back projection and "
                         "OSMO at 256^3 x 360; batched: OSMO on 64 uneven problems, one batched call per projection; "
                         "raytrace: curved-ray tracer at 192^3 x 360 angles (RK4) and its BCLP iteration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the light-propagation path runs on a single MI355X, "
              "and --gpus N with N > 1 is not supported",
              file=sys.stderr)
        sys.exit(2)
    args.cpu_pool = None
    if args.workload == "bclp" and not args.no_cpu and not args.kernels_only:
        # the all-core CPU baseline's worker processes start before this
        # process touches the GPU (they run NumPy only)
        import multiprocessing as mp
        args.cpu_workers = max(1, min(16, os.cpu_count() or 1))
        omp = os.environ.get("OMP_NUM_THREADS")
        os.environ["OMP_NUM_THREADS"] = "1"
        try:
            args.cpu_pool = mp.get_context("spawn").Pool(args.cpu_workers)
        finally:
            if omp is None:
                del os.environ["OMP_NUM_THREADS"]
            else:
                os.environ["OMP_NUM_THREADS"] = omp
    try:
        {"bclp": run_bclp, "osmo": run_osmo, "batched": run_batched, "raytrace": run_raytrace,
         "imageseq": run_imageseq}[args.workload](args)
    finally:
        if args.cpu_pool is not None:
            args.cpu_pool.terminate()
            args.cpu_pool.join()


def run_bclp(args):
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.optimizer.BCLP import BCLPNorm
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    angles = np.arange(NA, dtype=np.float64)

    log("building target")
    T = make_target(dev)
    tg = vam.geometry.TargetGeometry(target=T.cpu().numpy())
    pg = vam.geometry.ProjectionGeometry(angles, "parallel", CUDA=True, absorption_per_voxel=MU, det_offset=OFFSET)
    rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
    inv = vam.response.ResponseModel(type="interpolation", B=20, M=0.6)
    P = vam.projectorconstructor.projectorconstructor(tg, pg)
    log("initial sinogram")
    g0 = P.forward(inv.map_inv(T.double()).float())
    g0 = vam.util.data.filterSinogram(g0, "ram-lak") * (np.pi / (2 * NA))
    g0 = torch.clamp(g0, min=0).float().contiguous()
    # absorption dims the far side of the vial: rescale so that the mean dose
    # inside the target sits just above the response threshold (0.6)
    d0 = P.backward(g0)
    g0 = (g0 * (0.7 / float(d0[T > 0.5].mean()))).contiguous()
    lr = args.lr

    if args.autotune:
        log("autotuning the work order")
        rep = P.plan.autotune(T, torch.rand(P.plan.sino_shape, device=dev), reps=3)
        for what in ("forward", "backward"):
            log("  %s: %s" % (what, ", ".join("%s %.3f ms" % (k, v) for k, v in sorted(rep[what].items(),
                                                                                   key=lambda kv: kv[1]))))
        log("tuning choice %s" % rep["choice"])
    tuning = P.plan.get_tuning()

    def options(**kw):
        return vam.optimize.Options(method="BCLP", n_iter=kw.pop("n_iter"), response_model=rm, eps=0.05, p=2, q=1,
                                    learning_rate=lr, g0=g0, **kw)

    # ---- fused, graph-replayed iteration ------------------------------------
    log("fused BCLP: warmup %d, timed %d" % (args.warmup, args.steps))
    bc = BCLPNorm(tg, pg, options(n_iter=args.steps, fused=True, hip_graph=True))
    lib = _lib.load()
    plan = bc.P.plan
    plan.set_tuning(**tuning)
    g = bc.g0.reshape(plan.sino_shape).contiguous().clone()
    dose = torch.empty(plan.vol_shape, device=dev)
    loss_log = torch.zeros(args.steps + args.warmup + 2, device=dev, dtype=torch.float64)
    counter = torch.zeros(1, device=dev, dtype=torch.int32)
    nb = ctypes.c_int64()
    _lib.check(lib.vamx_bclp_workspace_bytes(plan.handle, ctypes.byref(nb)), "workspace")
    ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
    resp, prm = bc._vamx_args()

    def step():
        _lib.check(lib.vamx_bclp_step(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(),
                                      bc.T.data_ptr(), 0, 0, dose.data_ptr(), loss_log.data_ptr(),
                                      counter.data_ptr(), ws.data_ptr(), _lib.stream_ptr(dev)), "vamx_bclp_step")

    if args.kernels_only:
        # each projection 5 times, the fused iteration 5 times (its two
        # projection kernels), then a 1 GiB device copy 3 times as the
        # reference for the traffic counters
        sino = torch.empty(P.plan.sino_shape, device=dev)
        vol = torch.empty(P.plan.vol_shape, device=dev)
        for _ in range(5):
            P.plan.forward(T, out=sino)
        for _ in range(5):
            P.plan.backward(g0, out=vol)
        for _ in range(5):
            step()
        a = torch.empty(1 << 28, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)  # 1 GiB read + 1 GiB write: counter calibration
        torch.cuda.synchronize()
        log("kernels-only run done")
        return

    step()
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            step()
    torch.cuda.current_stream(dev).wait_stream(s)
    for _ in range(max(args.warmup - 1, 0)):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    iters_s = args.steps / dt
    log("fused: %.3f ms/iter, %.1f iters/s" % (ms_per_step, iters_s))
    n_fused = args.warmup + args.steps
    losses = loss_log[:n_fused].cpu().numpy()
    log("loss first %.6g last %.6g" % (losses[0], losses[-1]))
    g_fused = g.clone()
    dose_fused = dose.clone()

    # ---- unfused GPU loop (same projector kernels + torch element-wise) -----
    log("unfused GPU loop")
    bu = BCLPNorm(tg, pg, options(n_iter=n_fused, fused=False, log_metrics=False))
    bu.P.plan.set_tuning(**tuning)
    gu = bu.g0.reshape(-1).clone()
    t_unf = []
    for k in range(n_fused):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        d_, m_, eT, eb, v, loss, operand = bu._eval_state(gu)
        L = float(loss.item())
        G = bu._grad(operand, L)
        gu = bu.imposeSinogramConstraints(bu._cyl(gu - lr * G.reshape(-1))).reshape(-1)
        torch.cuda.synchronize()
        t_unf.append(time.perf_counter() - ta)
    unf_ms = 1e3 * float(np.mean(t_unf[args.warmup:])) if n_fused > args.warmup else 1e3 * float(np.mean(t_unf))
    dose_u = bu.P.backward(bu._cyl(gu)).reshape(plan.vol_shape)
    # the fused dose buffer holds B g of the last step's input; compare like with like
    rel_g = float(torch.linalg.norm((g_fused.reshape(-1) - gu).double()) / torch.linalg.norm(gu.double()))
    dose_f_last = bu.P.backward(g_fused.reshape(plan.sino_shape))
    rel_d = float(torch.linalg.norm((dose_f_last - dose_u).double()) / torch.linalg.norm(dose_u.double()))
    log("unfused: %.3f ms/iter; fused vs unfused rel L2 g %.2e dose %.2e" % (unf_ms, rel_g, rel_d))

    # ---- kernel times with HIP events on the launch stream -------------------
    def time_kernel(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    op = torch.clamp(bc.T, 0, 1)  # a volume of the workload's shape and sign structure
    sino_out = torch.empty(plan.sino_shape, device=dev)
    vol_out = torch.empty(plan.vol_shape, device=dev)
    fp_ms = time_kernel(lambda: plan.forward(op, out=sino_out))
    bp_ms = time_kernel(lambda: plan.backward(g, out=vol_out))
    V, S = N * N * NZ, N * NA * NZ
    fp_bytes = 4.0 * (NA * V + S)
    bp_bytes = 4.0 * (NA * V + V)
    log("FP %.3f ms (%.0f GB/s algorithmic)  BP %.3f ms (%.0f GB/s algorithmic)" % (
        fp_ms, fp_bytes / fp_ms / 1e6, bp_ms, bp_bytes / bp_ms / 1e6))
    # the two kernels the timed iteration replays, each timed alone on the
    # stream it runs on: kernel A (back projection + epilogue, and the
    # one-block loss reduction) through vamx_bclp_eval, kernel C (forward
    # projection of the operand fused with the update, and the 16 us slice
    # min/max pre-pass of the warp clip) through vamx_bclp_update
    l_out = torch.zeros(1, device=dev, dtype=torch.float64)
    g_t = g.clone()
    st = _lib.stream_ptr(dev)
    ws_t = torch.empty_like(ws)
    eval_fn = lambda: _lib.check(lib.vamx_bclp_eval(plan.handle, ctypes.byref(resp), ctypes.byref(prm),  # noqa
                                                    g_t.data_ptr(), bc.T.data_ptr(), 0, 0, vol_out.data_ptr(), 0,
                                                    l_out.data_ptr(), ws_t.data_ptr(), st), "vamx_bclp_eval")
    upd_fn = lambda: _lib.check(lib.vamx_bclp_update(plan.handle, ctypes.byref(prm), g_t.data_ptr(),  # noqa
                                                     ws_t.data_ptr(), st), "vamx_bclp_update")
    epi_ms = time_kernel(eval_fn)
    eval_fn()
    upd_ms = time_kernel(upd_fn)
    epi_alg = 4.0 * (NA * V + 4 * V)        # sinogram samples + target read, dose and operand written
    upd_alg = 4.0 * (NA * V + 3 * S)        # operand samples + g read and written + sinogram output
    prof = _profile_table()

    def roof(name, ms, alg):
        tr = prof.get(name, {})
        traffic = tr.get("bytes")
        achieved = traffic / (ms * 1e-3) / 1e9 if traffic else None
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic, "kernel": name,
                "kernel_ms": ms, "algorithmic_bytes": alg, "algorithmic_gbs": alg / (ms * 1e-3) / 1e9,
                "td_busy": tr.get("td_busy"), "ta_busy": tr.get("ta_busy"),
                "limiter": "texture data path (TD busy %.0f %% of the kernel)" % (100 * tr["td_busy"])
                if tr.get("td_busy") else None}

    # ---- the same iteration through autograd: BCLPLoss forward (the fused
    # epilogue kernel) and backward (the exact transpose of the back
    # projection, wavefront-uniform kernel), then the clipped update ---------
    from vamtoolbox_amd.autograd import BCLPLoss
    g_ag = g.clone().requires_grad_(True)

    def autograd_step():
        g_ag.grad = None
        L = BCLPLoss.apply(g_ag, bc.P, rm, bc.T, 0.05, 1.0, 2.0, 1.0)
        L.backward()
        with torch.no_grad():
            g_ag.sub_(lr * g_ag.grad).clamp_(min=0.0)

    autograd_ms = time_kernel(autograd_step)
    log("autograd BCLP iteration (BCLPLoss forward + exact-adjoint backward + update): %.3f ms" % autograd_ms)
    # the fused iteration with the exact gradient (grad_op 1: the transpose of
    # the back projection fused with the update instead of the forward
    # projection), eager launches
    prm_x = _lib.BclpParams.from_buffer_copy(prm)
    prm_x.grad_op = 1
    g_x = g.clone()
    log_x = torch.zeros(64, device=dev, dtype=torch.float64)
    cnt_x = torch.zeros(1, device=dev, dtype=torch.int32)
    exact_ms = time_kernel(lambda: (cnt_x.zero_(), _lib.check(lib.vamx_bclp_step(  # noqa
        plan.handle, ctypes.byref(resp), ctypes.byref(prm_x), g_x.data_ptr(), bc.T.data_ptr(), 0, 0,
        vol_out.data_ptr(), log_x.data_ptr(), cnt_x.data_ptr(), ws_t.data_ptr(), st), "vamx_bclp_step")))
    log("fused BCLP iteration with the exact gradient (grad_op 1): %.3f ms" % exact_ms)

    roof_c = roof("bclp_fp_update_pair", upd_ms, upd_alg)
    roof_a = roof("bclp_bp_epilogue_wave", epi_ms, epi_alg)
    log("kernel C (FP + update) %.3f ms: %s GB/s measured traffic; kernel A (BP + epilogue) %.3f ms" % (
        upd_ms, "%.0f" % roof_c["achieved"] if roof_c["achieved"] else "n/a", epi_ms))

    # ---- plain PyTorch loop (no custom kernels) -------------------------------
    torch_ms = None
    if not args.no_torch_baseline:
        log("plain PyTorch BCLP iteration")
        tp = TorchD1(N, NZ, angles, dev, mu=MU, offset=OFFSET)
        gt = g0.clone()
        Td = bc.T
        times = []
        for k in range(2):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            dd = tp.backward(gt)
            mm = rm.map(dd)
            eT = mm - Td
            eb = torch.abs(eT) - 0.05
            vv = eb > 0
            loss = torch.sum(vv * eb ** 2) ** 0.5
            Lh = float(loss.item())
            opd = (vv * torch.abs(eb) * torch.sign(eT) * rm.dmapdf(dd)).float()
            G = tp.forward(opd)
            if Lh != 0:
                G = G / Lh
            gt = torch.clamp(gt - lr * G, min=0)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ta)
        torch_ms = 1e3 * times[-1]
        log("plain torch: %.1f ms/iter" % torch_ms)

    # ---- CPU baseline ----------------------------------------------------------
    cpu = None
    oracle_check = None
    if not args.no_cpu:
        log("CPU baseline (NumPy port, bounded sample)")
        tnp, gnp = bc.T.cpu().numpy().astype(np.float64), g.cpu().numpy()
        one = cpu_baseline(tnp, gnp, angles)
        log("CPU, one core: %.3e iters/s" % one["value"])
        if args.cpu_pool is not None:
            cpu = cpu_baseline_all_cores(args.cpu_pool, args.cpu_workers, tnp, gnp, angles)
            cpu["one_core"] = {"value": one["value"], "sample": one["sample"]}
            # the port's dose and unscaled gradient F(operand) on its slices
            # against the GPU's at the same g (the oracle check of the timed
            # iteration's two projections and epilogue at this geometry)
            zs, res = cpu.pop("slices"), cpu.pop("results")
            op_t = torch.empty(plan.vol_shape, device=dev)
            dose_t = torch.empty(plan.vol_shape, device=dev)
            _lib.check(lib.vamx_bclp_eval(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(),
                                          bc.T.data_ptr(), 0, 0, dose_t.data_ptr(), op_t.data_ptr(),
                                          l_out.data_ptr(), ws_t.data_ptr(), st), "vamx_bclp_eval")
            fwd_t = plan.forward(op_t)
            dg, fg = dose_t.cpu().numpy(), fwd_t.cpu().numpy()
            dref = np.concatenate([r[0] for r in res], axis=2)
            fref = np.concatenate([r[1] for r in res], axis=2)
            e_d = float(np.linalg.norm(dg[:, :, zs] - dref) / np.linalg.norm(dref))
            e_f = float(np.linalg.norm(fg[:, :, zs] - fref) / np.linalg.norm(fref))
            oracle_check = {"rel_l2_dose": e_d, "rel_l2_gradient": e_f, "z_slices": zs}
            log("GPU vs float64 port on %d slices: rel L2 dose %.2e, gradient %.2e" % (len(zs), e_d, e_f))
        else:
            cpu = one
        cpu["host"] = "%d host cores visible (%s)" % (os.cpu_count() or 0, _cpu_model())
        log("CPU: %.3e iters/s on %d cores" % (cpu["value"], cpu["cores"]))

    out = {
        "metric": "BCLP iters/s at 256³×360 angles (rel err ≤1e-4); fwd/back HBM GB/s vs peak",
        "value": iters_s,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "BCLP 256^3 voxels x 360 angles, 256x256 detector, smoothed ellipsoid target, "
                               "sigmoid response (M=0.6, B=20), eps=0.05, p=2, q=1, absorption 0.01/voxel, "
                               "detector offset 10 voxels, fused HIP-graph iteration",
                   "n": N, "nz": NZ, "n_angles": NA},
        # the dominant kernel of the timed iteration (kernel C); achieved =
        # its counter-measured HBM/fabric bytes per launch (profiles/) over
        # its event-timed duration; the algorithmic rate (one 4-byte read per
        # interpolated sample) is a separate field: L2 and the opposite-ray
        # pairing serve the repeated tap reads, so it exceeds the HBM peak
        "roofline": roof_c,
        "roofline_bp": roof_a,
        "cpu_baseline": cpu,
        "forward_ms": fp_ms,
        "backward_ms": bp_ms,
        "bp_epilogue_ms": epi_ms,
        "fp_update_ms": upd_ms,
        "autograd_ms_per_iter": autograd_ms,
        "exact_gradient_ms_per_iter": exact_ms,
        "rel_err_vs_oracle": oracle_check,
        "forward_gbs": fp_bytes / fp_ms / 1e6,
        "backward_gbs": bp_bytes / bp_ms / 1e6,
        "unfused_ms_per_iter": unf_ms,
        "speedup_vs_unfused_hip": unf_ms / ms_per_step,
        "plain_torch_ms_per_iter": torch_ms,
        "speedup_vs_plain_torch": (torch_ms / ms_per_step) if torch_ms else None,
        "rel_err_sino_vs_unfused": rel_g,
        "rel_err_dose_vs_unfused": rel_d,
        "loss_first": float(losses[0]),
        "loss_last": float(losses[-1]),
        "tuning": tuning,
    }
    print(json.dumps(out), flush=True)


def run_imageseq(args):
    """Frames for a DLP projector: 360 frames of 1920 x 1080 from a 256 x 360 x
    256 float64 sinogram, two arrayed copies, 99.5th-percentile normalisation
    (ImageSeq, imagesequence.py:85-182).  The compose kernel writes every frame
    byte once and reads each image sample once per copy."""
    import torch
    from vamtoolbox_amd.imagesequence import ImageConfig, ImageSeq
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(0)
    s = torch.as_tensor(rng.uniform(0, 1, (N, NA, NZ)), device=dev)
    cfg = ImageConfig((1920, 1080), array_num=2, array_offset=300, normalization_percentile=99.5)
    ImageSeq(cfg, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seq = ImageSeq(cfg, s)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    # the frames alone (clear + compose), through the C ABI, on torch's stream
    import ctypes
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.imagesequence import _windows
    lib = _lib.load()
    prm = _lib.ImgSeqParams()
    prm.n_u, prm.n_v = 1920, 1080
    wins = _windows(cfg, N, NZ)
    prm.u1, prm.n_copies = wins[0][1], len(wins)
    for k, (v1, _) in enumerate(wins):
        prm.v1[k] = v1
    prm.bit_depth, prm.normalize, prm.norm_value, prm.intensity_scale = 8, 1, 0.995, 1.0
    frames = torch.empty((NA, 1080, 1920), device=dev, dtype=torch.uint8)
    comp_ms = _time_events(lambda: _lib.check(lib.vamx_imgseq_compose(
        s.data_ptr(), N, NA, NZ, ctypes.byref(prm), frames.data_ptr(), _lib.stream_ptr(dev)), "compose"), 10)
    alg = NA * 1080 * 1920 + 8.0 * N * NA * NZ * len(wins)
    log("ImageSeq: %.2f ms per sequence (percentile + frames); frames (clear + compose) %.3f ms, %.0f GB/s "
        "algorithmic" % (
        ms, comp_ms, alg / comp_ms / 1e6))
    roof = {"bound": "hbm", "achieved": alg / comp_ms / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": alg / comp_ms / 1e6 / HBM_PEAK_GBS, "traffic": None, "kernel": "imgseq_clear + imgseq_compose",
            "kernel_ms": comp_ms, "algorithmic_bytes": alg}
    _line("image sequences/s (360 frames of 1920x1080 from a 256x360x256 sinogram)", 1e3 / ms, "sequences/s", ms,
          args, {"workload": "ImageSeq: 360 frames 1920x1080 uint8, 2 arrayed copies, 99.5th-percentile "
                             "normalisation, float64 sinogram 256x360x256 resident on the GPU"}, roof, None,
          compose_ms=comp_ms, frames_bytes=int(seq.frames.numel()))


def _profile_table():
    """profiles/traffic.txt: kernel, HBM/fabric bytes per launch (calibrated
    FETCH_SIZE + WRITE_SIZE), TA and TD busy fractions, from the committed
    counter passes."""
    out = {}
    tfile = os.path.join(HERE, "profiles", "traffic.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            parts = line.split()
            if len(parts) >= 2 and not line.startswith("#"):
                try:
                    e = {"bytes": float(parts[1])}
                    if len(parts) >= 4:
                        e["ta_busy"], e["td_busy"] = float(parts[2]), float(parts[3])
                    out[parts[0]] = e
                except ValueError:
                    pass
    return out


def _line(metric, value, unit, ms, args, config, roofline, cpu, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _time_events(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _ellipsoid_union(shape, count, semi, centre_radius, rng, device):
    """Binary union of `count` random ellipsoids (semi-axes U[semi], centres in a
    ball of `centre_radius` voxels, random rotations), built on the GPU."""
    import torch
    n, _, nz = shape
    ax = torch.arange(n, device=device, dtype=torch.float32) - (n - 1) / 2
    zz = torch.arange(nz, device=device, dtype=torch.float32) - (nz - 1) / 2
    Y, X, Z = torch.meshgrid(ax, ax, zz, indexing="ij")
    vol = torch.zeros(shape, device=device)
    for _ in range(count):
        a, b, c = rng.uniform(semi[0], semi[1], 3)
        c = min(c, nz / 2) if nz < 2 * semi[1] else c
        ctr = rng.uniform(-centre_radius, centre_radius, 3)
        ctr[2] = rng.uniform(-nz / 4, nz / 4)
        Rt = torch.tensor(_rotation(rng), device=device, dtype=torch.float32)
        u = torch.stack([X - ctr[0], Y - ctr[1], Z - ctr[2]], dim=-1) @ Rt.T
        inside = (u[..., 0] / a) ** 2 + (u[..., 1] / b) ** 2 + (u[..., 2] / c) ** 2 <= 1
        vol = torch.maximum(vol, inside.float())
    return vol


# ---------------------------------------------------------------------------
# OSMO at 256^3 x 360 (plain forward / back projection + 50 OSMO iterations)
# ---------------------------------------------------------------------------
def run_osmo(args):
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.util.data import clipToCircle
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(2)
    T = clipToCircle(_ellipsoid_union((N, N, NZ), 8, (15, 70), 50, rng, dev))
    angles = np.arange(NA, dtype=np.float64)
    tg = vam.geometry.TargetGeometry(target=T.cpu().numpy())
    pg = vam.geometry.ProjectionGeometry(angles, "parallel", CUDA=True, absorption_per_voxel=MU, det_offset=OFFSET)
    P = vam.projectorconstructor.projectorconstructor(tg, pg)
    plan = P.plan
    sino = torch.empty(plan.sino_shape, device=dev)
    vol = torch.empty(plan.vol_shape, device=dev)
    fp_ms = _time_events(lambda: plan.forward(T, out=sino), 10)
    bp_ms = _time_events(lambda: plan.backward(sino, out=vol), 10)
    V, S = N * N * NZ, N * NA * NZ
    fp_bytes, bp_bytes = 4.0 * (NA * V + S), 4.0 * (NA * V + V)
    log("FP %.3f ms (%.0f GB/s)  BP %.3f ms (%.0f GB/s)" % (fp_ms, fp_bytes / fp_ms / 1e6, bp_ms,
                                                         bp_bytes / bp_ms / 1e6))
    n_iter = 50
    opts = vam.optimize.Options(method="OSMO", n_iter=n_iter, d_h=0.8, d_l=0.7, filter="ram-lak")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sino_o, recon, err = vam.optimize.optimize(tg, pg, opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    it_s = (n_iter - 1) / max(float(opts.iter_times[-1] - opts.iter_times[0]), 1e-9)
    log("OSMO %d iterations: %.2f s total (with the VER of each iteration), %.1f iters/s; VER %.4f -> %.4f" % (
        n_iter, dt, it_s, err[0], err[-1]))
    # the forward's measured HBM/fabric bytes per launch (profiles/traffic.txt)
    # over its event-timed duration; the algorithmic rate is a separate field
    tr = _profile_table().get("par_fp_pair_kernel", {})
    traffic = tr.get("bytes")
    achieved = traffic / (fp_ms * 1e-3) / 1e9 if traffic else None
    _line("OSMO iters/s at 256^3 x 360 angles", it_s, "iters/s", 1e3 / it_s, args,
          {"workload": "OSMO 50 iterations, 256^3 binary target of 8 ellipsoids, 360 angles, absorption "
                       "0.01/voxel, detector offset 10 voxels (2 forward + 2 back projections per iteration)",
           "n": N, "nz": NZ, "n_angles": NA},
          {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
           "kernel": "par_fp_pair_kernel", "kernel_ms": fp_ms, "algorithmic_bytes": fp_bytes,
           "algorithmic_gbs": fp_bytes / (fp_ms * 1e-3) / 1e9, "td_busy": tr.get("td_busy")},
          None, forward_ms=fp_ms, backward_ms=bp_ms, forward_gbs=fp_bytes / fp_ms / 1e6,
          backward_gbs=bp_bytes / bp_ms / 1e6, ver_first=float(err[0]), ver_last=float(err[-1]),
          total_s=dt)


# ---------------------------------------------------------------------------
# 64 uneven OSMO problems, one batched call (vamx_par_{forward,backward}_batched) per projection
# ---------------------------------------------------------------------------
def run_batched(args):
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.batch import OSMOBatch
    from vamtoolbox_amd.util.data import clipToCircle
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(4)
    tgs, pgs, specs = [], [], []
    log("building 64 problems")
    for _ in range(64):
        n = int(rng.choice([128, 192, 256, 384]))
        nz = int(rng.integers(32, 129))
        na = int(rng.choice([90, 180, 360, 720]))
        mu = float(rng.uniform(0.0, 0.02))
        T = clipToCircle(_ellipsoid_union((n, n, nz), 5, (0.08 * n, 0.27 * n), 0.2 * n, rng, dev))
        tgs.append(vam.geometry.TargetGeometry(target=T.cpu().numpy()))
        pgs.append(vam.geometry.ProjectionGeometry(np.linspace(0, 360, na, endpoint=False), "parallel", CUDA=True,
                                                   absorption_per_voxel=mu))
        specs.append((n, nz, na, mu))
    opts = vam.optimize.Options(method="OSMO", n_iter=args.steps + args.warmup + 1, d_h=0.8, d_l=0.7,
                                filter="ram-lak")
    log("setting up the batch")
    st = OSMOBatch(tgs, pgs, opts, device=dev, autotune=args.autotune)
    B = st.batch
    samples = sum(na * n * n * nz for n, nz, na, _ in specs)
    sino_el = sum(n * na * nz for n, nz, na, _ in specs)
    vol_el = sum(n * n * nz for n, nz, _, _ in specs)
    fp_bytes = 4.0 * (samples + sino_el)
    bp_bytes = 4.0 * (samples + vol_el)
    fp_ms = _time_events(lambda: B.forward(st.x_model, out=st.b), 5)
    bp_ms = _time_events(lambda: B.backward(st.b, out=st.x), 5)
    log("batched FP %.3f ms (%.0f GB/s algorithmic)  BP %.3f ms (%.0f GB/s)" % (
        fp_ms, fp_bytes / fp_ms / 1e6, bp_ms, bp_bytes / bp_ms / 1e6))
    # restart from the initial state for the timed iterations
    st = OSMOBatch(tgs, pgs, opts, device=dev, autotune=args.autotune)
    st.step()
    st.capture()
    for _ in range(max(args.warmup - 1, 0)):
        st.graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.graph.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    ver = st.ver()
    log("batched OSMO: %.2f ms per iteration of all 64 problems (%.1f iters/s); mean VER %.4f" % (
        ms, 1e3 / ms, float(np.mean(ver))))
    # no counter pass of the batched launches: the HBM fraction is not
    # reported; the algorithmic rate is (one 4-byte read per interpolated sample)
    _line("batched OSMO iters/s, 64 uneven problems", 1e3 / ms, "iters/s", ms, args,
          {"workload": "64 OSMO problems, n in {128,192,256,384}, nz in U{32..128}, angles in {90,180,360,720}, "
                       "absorption U[0,0.02]/voxel, ellipsoid-union targets; one batched forward and one "
                       "batched back projection launch per half step, HIP-graph-replayed iteration",
           "problems": 64, "voxels": vol_el, "sinogram_elements": sino_el, "samples_per_projection": samples},
          {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": None, "traffic": None, "kernel": "vamx_par_forward_batched", "kernel_ms": fp_ms,
           "algorithmic_bytes": fp_bytes, "algorithmic_gbs": fp_bytes / (fp_ms * 1e-3) / 1e9},
          None, forward_ms=fp_ms, backward_ms=bp_ms, forward_gbs=fp_bytes / fp_ms / 1e6,
          backward_gbs=bp_bytes / bp_ms / 1e6, mean_ver=float(np.mean(ver)))


# ---------------------------------------------------------------------------
# curved-ray tracer at 192^3, 360 angles x 192^2 rays, RK4, and its BCLP iteration
# ---------------------------------------------------------------------------
def run_raytrace(args):
    import torch
    import vamtoolbox_amd as vam
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.optimizer.BCLP import BCLPNorm
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n, na = 192, 360
    rng = np.random.default_rng(5)
    T = _ellipsoid_union((n, n, n), 8, (12, 50), 40, rng, dev)
    r = torch.arange(-8, 9, device=dev, dtype=torch.float32)
    k = torch.exp(-r ** 2 / 8.0)
    k = k / k.sum()
    v = T[None, None]
    for dim in range(3):
        shape = [1, 1, 1, 1, 1]
        shape[2 + dim] = 17
        pad = [0, 0, 0, 0, 0, 0]
        pad[2 * (2 - dim)] = 8
        pad[2 * (2 - dim) + 1] = 8
        v = torch.nn.functional.conv3d(torch.nn.functional.pad(v, pad), k.reshape(shape))
    T = v[0, 0].clamp(0, 1)
    tg = vam.geometry.TargetGeometry(target=T.cpu().numpy())
    cv = tg.constructCoordVec()
    vox = float(cv[0][1] - cv[0][0])
    X, Y, Z = np.meshgrid(*cv, indexing="ij")
    nf = np.full(X.shape, 1.5)
    for _ in range(6):
        c = rng.uniform(-0.5, 0.5, 3) * np.array([cv[0].max(), cv[1].max(), cv[2].max()])
        s = rng.uniform(10, 30) * vox
        nf += rng.choice([-0.02, 0.02]) * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    im = vam.medium.IndexModel(cv, type="interpolation", form="freeform", n_x=nf)
    am = vam.medium.AttenuationModel(cv, type="analytical", form="homogeneous_cylinder", alpha_internal=0.005 / vox)
    ab = vam.medium.AbsorptionModel(cv, type="analytical", form="homogeneous_cylinder", alpha_internal=0.005 / vox)
    angles = np.arange(na, dtype=np.float64)
    pg = vam.geometry.ProjectionGeometry(angles, "ray_trace", CUDA=True, index_model=im, attenuation_model=am,
                                         absorption_model=ab, ray_trace_method="eikonal",
                                         eikonal_parametrization="canonical", ray_trace_ode_solver="rk4",
                                         ray_trace_ray_config="parallel")
    P = vam.projectorconstructor.projectorconstructor(tg, pg)
    P.output_torch_tensor = True
    f = T.reshape(-1).contiguous()
    g = P.forward(f)
    if args.kernels_only:
        # two more integrate and two deposit traces, then a 1 GiB device copy
        # 3 times as the reference for the traffic counters
        for _ in range(2):
            P.forward(f)
        for _ in range(2):
            P.backward(g)
        a = torch.empty(1 << 28, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("kernels-only run done")
        return
    fwd_ms = _time_events(lambda: P.forward(f), 3)
    steps_f = P.stats()
    bwd_ms = _time_events(lambda: P.backward(g), 3)
    steps_b = P.stats()
    in_box = float(steps_f[0])
    log("ray trace: integrate %.1f ms, deposit %.1f ms, %.3e ray-steps per trace (%d unfinished)" % (
        fwd_ms, bwd_ms, in_box, steps_f[1]))
    # BCLP on the ray tracer: the fused iteration (vamx_rt_bclp_step: deposit
    # with the response / band / Lp epilogue, integrate fused with the
    # update) replayed from a HIP graph, and the unfused loop (trace-deposit,
    # torch element-wise response, trace-integrate) for comparison
    rm = vam.response.ResponseModel(type="analytical", B=20, M=0.6)
    g0 = torch.clamp(P.forward(f), min=0)
    d0 = P.backward(g0).reshape(T.shape)
    g0 = (g0 * (0.7 / float(d0[T > 0.5].mean()))).reshape(-1).contiguous()
    opts = vam.optimize.Options(method="BCLP", n_iter=20, response_model=rm, eps=0.05, p=2, q=1,
                                learning_rate=0.01, g0=g0, log_metrics=False, fused=False)
    bc = BCLPNorm(tg, pg, opts)
    gi = bc.g0.reshape(-1).clone()
    # step size: the first step moves g by 5 % of its mean magnitude
    _, _, _, _, _, loss0, op0 = bc._eval_state(gi)
    G0 = bc._grad(op0, float(loss0.item()))
    lr = 0.05 * float(gi.abs().mean()) / max(float(G0.abs().mean()), 1e-30)
    times, losses = [], []
    for it in range(6):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        d_, m_, eT, eb, vmask, loss, operand = bc._eval_state(gi)
        L = float(loss.item())
        G = bc._grad(operand, L)
        gi = bc.imposeSinogramConstraints(gi - lr * G.reshape(-1)).reshape(-1)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - ta)
        losses.append(L)
    unf_ms = 1e3 * float(np.mean(times[2:]))
    log("ray-trace BCLP, unfused: %.1f ms per iteration" % unf_ms)
    lib = _lib.load()
    resp, prm = bc._vamx_args()
    prm.learning_rate = lr
    gf = bc.g0.reshape(-1).contiguous().clone()
    Tt = bc.T.contiguous()
    dose_f = torch.empty(P.vol_shape, device=dev)
    n_it = args.steps + args.warmup + 1
    loss_log = torch.zeros(n_it + 1, device=dev, dtype=torch.float64)
    counter = torch.zeros(1, device=dev, dtype=torch.int32)
    nb = ctypes.c_int64()
    _lib.check(lib.vamx_rt_bclp_workspace_bytes(P.handle, ctypes.byref(nb)), "workspace")
    ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)

    def rt_step():
        _lib.check(lib.vamx_rt_bclp_step(P.handle, ctypes.byref(resp), ctypes.byref(prm), gf.data_ptr(),
                                         Tt.data_ptr(), 0, 0, dose_f.data_ptr(), loss_log.data_ptr(),
                                         counter.data_ptr(), ws.data_ptr(), _lib.stream_ptr(dev)),
                   "vamx_rt_bclp_step")

    rt_step()
    graph = torch.cuda.CUDAGraph()
    sg = torch.cuda.Stream(device=dev)
    sg.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(sg):
        with torch.cuda.graph(graph, stream=sg):
            rt_step()
    torch.cuda.current_stream(dev).wait_stream(sg)
    for _ in range(max(args.warmup - 1, 0)):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps_t = max(args.steps, 1)
    for _ in range(steps_t):
        graph.replay()
    torch.cuda.synchronize()
    it_ms = (time.perf_counter() - t0) / steps_t * 1e3
    fl = loss_log[:steps_t + args.warmup].cpu().numpy()
    log("ray-trace BCLP, fused (graph): %.1f ms per iteration; loss %.4g -> %.4g" % (it_ms, fl[0], fl[-1]))
    losses = [float(fl[0]), float(fl[-1])]
    bytes_per_step = 72.0
    achieved = in_box * bytes_per_step / (fwd_ms * 1e-3) / 1e9
    rt_tr = _profile_table().get("rt_trace_integrate", {})
    traffic = rt_tr.get("bytes")
    cpu = None
    if not args.no_cpu:
        cpu = _raytrace_cpu_baseline(tg, pg, cv, nf, am, ab, na)
    _line("ray-trace BCLP iters/s at 192^3 x 360 angles (RK4)", 1e3 / it_ms, "iters/s", it_ms, args,
          {"workload": "curved-ray tracer, 192^3 medium n = 1.5 + 6 Gaussian bumps (+-0.02, sigma 10-30 voxels), "
                       "attenuation and absorption 0.005/voxel in the cylinder, 360 angles x 192 x 192 rays, RK4, "
                       "step 0.5 voxel; BCLP iteration = deposit fused with the response / band / Lp epilogue, "
                       "integrate fused with the update, replayed from a HIP graph",
           "n": n, "n_angles": na, "rays": int(192 * 192 * na)},
          {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "rt_trace_kernel<0, RK4, canonical>",
           "kernel_ms": fwd_ms, "algorithmic_bytes": in_box * bytes_per_step,
           "traffic_gbs": traffic / (fwd_ms * 1e-3) / 1e9 if traffic else None,
           "td_busy": rt_tr.get("td_busy"), "ta_busy": rt_tr.get("ta_busy")},
          cpu, integrate_ms=fwd_ms, deposit_ms=bwd_ms, ray_steps=in_box, ray_steps_per_s=in_box / (fwd_ms * 1e-3),
          unfinished_rays=int(steps_f[1]), loss_first=losses[0], loss_last=losses[-1], learning_rate=lr,
          unfused_ms_per_iter=unf_ms, speedup_vs_unfused=unf_ms / it_ms)


def _raytrace_cpu_baseline(tg, pg, cv, nf, am, ab, na, k_angles=2):
    """The reference's CPU path (pyTorchRayTrace.py:378-571 on CPU torch:
    symplectic Euler, grid_sample, sparse-COO deposit) restated in
    oracle/raytrace_torch.py, on all host cores of the GPU's share, for
    k_angles of the angles, scaled to one integrate pass."""
    import torch
    import oracle.raytrace as ort
    import oracle.raytrace_torch as rtt
    threads = max(1, min(16, os.cpu_count() or 1))
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        tr = rtt.TorchCPUTracer(cv, nf, att_alpha=am.params["alpha_internal"], att_radius=am.params["R_physical"],
                                abs_grid=ab.alpha_at_grid(), param="canonical")
        x0, v0, shape, e0 = ort.setup_rays(cv, pg.angles[:k_angles])
        f = np.asarray(tg.array, dtype=np.float64)
        t0 = time.perf_counter()
        tr.run(x0, v0, e0, f=f)
        dt = time.perf_counter() - t0
    finally:
        torch.set_num_threads(prev)
    per = dt * (na / k_angles)
    return {"value": 1.0 / per, "unit": "integrate passes/s", "cores": threads, "kind": "port",
            "sample": "the reference's CPU-torch tracer (symplectic Euler, grid_sample, sparse-COO deposit; "
                      "oracle/raytrace_torch.py) with torch.set_num_threads(%d), %d of %d angles (%d rays), "
                      "%.1f s measured, scaled x%d" % (threads, k_angles, na, x0.shape[0], dt, na // k_angles),
            "host": "%d host cores visible (%s)" % (os.cpu_count() or 0, _cpu_model())}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
