"""One persistent launch over uneven problems equals the per-problem launches bitwise."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_equals_individual(gpu):
    import torch
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(11)
    specs = [(32, 8, 30, 0.0), (48, 5, 45, 0.01), (24, 12, 90, 0.0), (40, 3, 20, 0.02), (16, 1, 60, 0.0)]
    plans, vols, sinos, ref_s, bps, ref_b = [], [], [], [], [], []
    for n, nz, na, mu in specs:
        ang = np.linspace(0, 180, na, endpoint=False)
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
