"""vamx_bclp_step_noalloc (no workspace argument, no allocation) equals
vamx_bclp_step bitwise, call after call, through the C ABI."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(gpu, n, nz, na, mu, offset, seed):
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(seed)
    ang = np.linspace(0.0, 360.0, na, endpoint=False)
    plan = ParallelPlan(n, nz, ang, mu=mu, det_offset=offset, device=gpu)
    target = torch.tensor(np.clip(rng.random((n, n, nz)) * 1.2 - 0.1, 0, 1), dtype=torch.float32, device=gpu)
    g0 = torch.tensor(rng.random((n, na, nz)) * 0.05, dtype=torch.float32, device=gpu)
    return plan, target, g0


def _structs(lr):
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.response import ResponseModel
    resp, _ = ResponseModel(type="analytical", B=20, M=0.6).vamx_struct(None)
    prm = _lib.BclpParams(p=2.0, q=1.0, eps=0.05, weight=1.0, learning_rate=lr, glb=0.0, gub=np.inf, dvol=1.0,
                          grad_op=0)
    return resp, prm


def _run_pair(gpu, plan, target, g0, iters, lr):
    """iters iterations with the workspace entry point and with the allocation-free one"""
    import torch
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    resp, prm = _structs(lr)
    st = _lib.stream_ptr(gpu)
    nbytes = ctypes.c_int64()
    _lib.check(lib.vamx_bclp_workspace_bytes(plan.handle, ctypes.byref(nbytes)), "workspace_bytes")
    ws = torch.empty(int(nbytes.value), device=gpu, dtype=torch.uint8)
    out = {}
    for mode in ("workspace", "noalloc"):
        g = g0.clone()
        dose = torch.empty(plan.vol_shape, device=gpu)
        operand = torch.empty(plan.vol_shape, device=gpu)
        loss = torch.full((iters + 1,), float("nan"), device=gpu, dtype=torch.float64)
        cnt = torch.zeros(1, device=gpu, dtype=torch.int32)
        for _ in range(iters):
            if mode == "workspace":
                rc = lib.vamx_bclp_step(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(),
                                        target.data_ptr(), 0, 0, dose.data_ptr(), loss.data_ptr(), cnt.data_ptr(),
                                        ws.data_ptr(), st)
            else:
                rc = lib.vamx_bclp_step_noalloc(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(),
                                                target.data_ptr(), 0, 0, dose.data_ptr(), operand.data_ptr(),
                                                loss.data_ptr(), cnt.data_ptr(), st)
            _lib.check(rc, mode)
        torch.cuda.synchronize()
        out[mode] = (g, dose, loss, cnt)
    return out


def test_noalloc_step_equals_workspace_step(gpu):
    import torch
    # first call: 64^3 x 90 angles (wavefront kernels); second: a smaller
    # 40 x 40 x 24 problem, on the same scratch array
    for n, nz, na, mu, off, seed, iters in [(64, 64, 90, 0.01, 2.0, 1, 3), (40, 24, 60, 0.0, 0.0, 2, 2)]:
        plan, target, g0 = _problem(gpu, n, nz, na, mu, off, seed)
        print("noalloc vs workspace: n %d nz %d angles %d" % (n, nz, na), flush=True)
        out = _run_pair(gpu, plan, target, g0, iters, lr=1e-3)
        gw, dw, lw, cw = out["workspace"]
        gn, dn, ln, cn = out["noalloc"]
        assert int(cw.item()) == iters and int(cn.item()) == iters
        assert torch.equal(gw, gn)
        assert torch.equal(dw, dn)
        assert torch.equal(lw[:iters], ln[:iters])
        assert bool(torch.isfinite(ln[:iters]).all())
        assert not torch.equal(gn, g0)   # the iterations did update g


def test_noalloc_step_rejects_oversized_problem(gpu):
    """More loss partials than the scratch holds: VAMX_ERR_ARG before any launch."""
    import torch
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    resp, prm = _structs(1e-3)
    # per-lane kernels, one z per lane (odd nz): 256 * 256 * 513 / 256 = 131328 blocks > 131040
    from vamtoolbox_amd.parallel import ParallelPlan
    plan = ParallelPlan(256, 513, [0.0, 90.0], per_lane_geometry=True, device=gpu)
    vol = torch.zeros(plan.vol_shape, device=gpu)
    dose = torch.empty_like(vol)
    operand = torch.empty_like(vol)
    g = torch.zeros(plan.sino_shape, device=gpu)
    loss = torch.zeros(2, device=gpu, dtype=torch.float64)
    cnt = torch.zeros(1, device=gpu, dtype=torch.int32)
    rc = lib.vamx_bclp_step_noalloc(plan.handle, ctypes.byref(resp), ctypes.byref(prm), g.data_ptr(), vol.data_ptr(),
                                    0, 0, dose.data_ptr(), operand.data_ptr(), loss.data_ptr(), cnt.data_ptr(),
                                    _lib.stream_ptr(gpu))
    assert rc == -1
    assert b"scratch" in lib.vamx_last_error()
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
    # a problem within the limit still runs afterwards
    small, target, g0 = _problem(gpu, 32, 16, 30, 0.0, 0.0, 4)
    out = _run_pair(gpu, small, target, g0, 1, lr=1e-3)
    assert torch.equal(out["workspace"][0], out["noalloc"][0])
