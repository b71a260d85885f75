"""Parity of the HIP parallel-beam kernels with the float64 oracle.

Tolerance: relative L2 <= 1e-5 against the float64 restatement of the
reference discretisation (float32 data arithmetic, float64 geometry); the
exact-adjoint pairs pass the dot-product test to 1e-6.
"""

import itertools

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _ellipses(n, nz, rng, count=5):
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    vol = np.zeros((n, n, nz))
    for z in range(nz):
        for _ in range(count):
            a, b = rng.uniform(0.1, 0.3, 2) * n
            cx, cy = rng.uniform(-0.2, 0.2, 2) * n
            th = rng.uniform(0, np.pi)
            u = (xx - cx) * np.cos(th) + (yy - cy) * np.sin(th)
            v = -(xx - cx) * np.sin(th) + (yy - cy) * np.cos(th)
            vol[:, :, z] = np.maximum(vol[:, :, z], ((u / a) ** 2 + (v / b) ** 2 <= 1).astype(float))
    return vol


CASES = [
    # n, nz, angles, kwargs
    (32, 8, np.arange(0, 360, 12.0), {}),
    (32, 8, np.arange(0, 180, 7.5), dict(det_offset=3.0)),
    (32, 8, np.arange(0, 360, 12.0), dict(mu=0.05)),
    (32, 8, np.arange(0, 360, 12.0), dict(mu=0.02, det_offset=2.0)),
    (33, 5, np.arange(0, 360, 10.0), {}),                 # odd n, nz % 4 != 0 (scalar path)
    (32, 1, np.arange(0, 360, 5.0), {}),                  # 2-D
    (24, 6, np.arange(0, 360, 15.0), dict(incl_deg=8.0)),
    (24, 6, np.arange(0, 360, 15.0), dict(incl_deg=-5.0, mu=0.03, det_offset=1.5)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_forward_backward_parity(gpu, case):
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan

    n, nz, ang, kw = CASES[case]
    print("case", case, n, nz, len(ang), kw, flush=True)
    rng = np.random.default_rng(100 + case)
    vol = _ellipses(n, nz, rng) * rng.uniform(0.5, 1.0, (1, 1, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    plan = ParallelPlan(n, nz, ang, device=gpu, **kw)
    fp = plan.forward(torch.tensor(vol, dtype=torch.float32, device=gpu)).cpu().numpy()
    bp = plan.backward(torch.tensor(sino, dtype=torch.float32, device=gpu)).cpu().numpy()
    fp_ref = d1.forward(vol, ang, **kw)
    bp_ref = d1.backward(sino, ang, **kw)
    e_fp, e_bp = rel_l2(fp, fp_ref), rel_l2(bp, bp_ref)
    print("rel L2 forward %.3e backward %.3e" % (e_fp, e_bp))
    assert e_fp <= TOL
    assert e_bp <= TOL


@pytest.mark.parametrize("case", range(len(CASES)))
def test_adjoint_parity(gpu, case):
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan

    n, nz, ang, kw = CASES[case]
    print("case", case, n, nz, len(ang), kw, flush=True)
    rng = np.random.default_rng(200 + case)
    vol = rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(-1, 1, (n, len(ang), nz))
    plan = ParallelPlan(n, nz, ang, device=gpu, warp_clip=False, **kw)
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    fT = plan.forward_adjoint(ts).cpu().numpy()
    bT = plan.backward_adjoint(tv).cpu().numpy()
    e1 = rel_l2(fT, d1.forward_adjoint(sino, ang, **kw))
    e2 = rel_l2(bT, d1.backward_adjoint(vol, ang, **kw))
    print("rel L2 F^T %.3e B^T %.3e" % (e1, e2))
    assert e1 <= TOL
    assert e2 <= TOL
    # dot-product tests <A x, y> = <x, A^T y> on the GPU results
    F = plan.forward(tv).double().cpu().numpy()
    B = plan.backward(ts).double().cpu().numpy()
    lhs1, rhs1 = np.sum(F * sino), np.sum(vol * fT)
    lhs2, rhs2 = np.sum(B * vol), np.sum(sino * bT)
    d1e = abs(lhs1 - rhs1) / (np.linalg.norm(F) * np.linalg.norm(sino))
    d2e = abs(lhs2 - rhs2) / (np.linalg.norm(B) * np.linalg.norm(vol))
    print("dot-product rel err %.3e %.3e" % (d1e, d2e))
    assert d1e <= 1e-6
    assert d2e <= 1e-6


def test_warp_clip_positive_slices(gpu):
    """Slices without zeros exercise skimage's output clip (golden vectors)."""
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "d1_vectors.npz"))
    keys = sorted(set(k.rsplit("_", 1)[0] for k in z.files))
    for key in keys:
        vol, ang = z[key + "_vol"], z[key + "_ang"]
        n, _, nz = vol.shape
        plan = ParallelPlan(n, nz, ang, device=gpu)
        fp = plan.forward(torch.tensor(vol, dtype=torch.float32, device=gpu)).cpu().numpy()
        bp = plan.backward(torch.tensor(z[key + "_sino"], dtype=torch.float32, device=gpu)).cpu().numpy()
        e1, e2 = rel_l2(fp, z[key + "_fwd"]), rel_l2(bp, z[key + "_bwd"])
        print(key, "%.2e %.2e" % (e1, e2))
        assert e1 <= TOL and e2 <= TOL, key


def test_determinism(gpu):
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(7)
    n, nz = 48, 16
    ang = np.arange(0, 360, 3.0)
    plan = ParallelPlan(n, nz, ang, device=gpu)
    v = torch.tensor(rng.random((n, n, nz)), dtype=torch.float32, device=gpu)
    s = torch.tensor(rng.random((n, len(ang), nz)), dtype=torch.float32, device=gpu)
    a1, a2 = plan.forward(v), plan.forward(v)
    b1, b2 = plan.backward(s), plan.backward(s)
    assert torch.equal(a1, a2) and torch.equal(b1, b2)


# shapes taken by the wavefront-uniform kernels: a wavefront covers one ray /
# voxel quad and 64 x ZV consecutive z; the last z chunk may be partial
WAVE_CASES = [
    (32, 64, np.arange(0, 360, 9.0), {}),
    (24, 128, np.arange(0, 180, 6.0), dict(mu=0.03, det_offset=2.0)),
    (20, 256, np.arange(0, 360, 20.0), dict(det_offset=-1.5)),
    (27, 40, np.arange(0, 360, 11.0), dict(mu=0.02)),                 # one partial chunk, ZV = 1
    (22, 100, np.arange(0, 180, 7.5), dict(det_offset=0.5)),          # partial chunk, ZV = 2
    (17, 198, np.arange(3, 360, 13.0), {}),                           # odd n, 2 chunks of 128, ZV = 2
]


@pytest.mark.parametrize("case", range(len(WAVE_CASES)))
@pytest.mark.parametrize("positive", [False, True])
def test_wave_kernels_parity(gpu, case, positive):
    """Wavefront-uniform kernels against the oracle, with the warp clip inactive
    (signed data) and active (strictly positive slices), and against the
    per-lane-geometry kernels."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz, ang, kw = WAVE_CASES[case]
    print("case", case, n, nz, len(ang), kw, "positive" if positive else "signed", flush=True)
    rng = np.random.default_rng(300 + case)
    vol = rng.uniform(0.2, 1.0, (n, n, nz)) if positive else rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    wave = ParallelPlan(n, nz, ang, device=gpu, **kw)
    lane = ParallelPlan(n, nz, ang, device=gpu, per_lane_geometry=True, **kw)
    fw, bw = wave.forward(tv).cpu().numpy(), wave.backward(ts).cpu().numpy()
    fl, bl = lane.forward(tv).cpu().numpy(), lane.backward(ts).cpu().numpy()
    e1, e2 = rel_l2(fw, d1.forward(vol, ang, **kw)), rel_l2(bw, d1.backward(sino, ang, **kw))
    print("rel L2 forward %.2e backward %.2e; wave vs lane %.2e %.2e" % (e1, e2, rel_l2(fw, fl), rel_l2(bw, bl)))
    assert e1 <= TOL and e2 <= TOL
    assert rel_l2(fw, fl) <= 1e-6 and rel_l2(bw, bl) <= 1e-6


@pytest.mark.parametrize("nz, kw", [(6, dict(mu=0.02)), (34, dict(det_offset=0.5)), (18, {})])
@pytest.mark.parametrize("positive", [False, True])
def test_per_lane_two_z_per_thread(gpu, nz, kw, positive):
    """The per-lane forward and back projection with two z per thread (even
    nz, not a multiple of 4, 8-byte aligned buffers) against the oracle, and
    bitwise equal to the one-z-per-thread kernels that buffers 4 bytes off an
    8-byte boundary take; warp clip inactive (signed data) and active
    (positive slices)."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, ang = 23, np.arange(0, 360, 9.0)
    print("nz", nz, kw, "positive" if positive else "signed", flush=True)
    rng = np.random.default_rng(700 + nz)
    vol = rng.uniform(0.2, 1.0, (n, n, nz)) if positive else rng.uniform(-1, 1, (n, n, nz))
    lane = ParallelPlan(n, nz, ang, device=gpu, per_lane_geometry=True, **kw)
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    f2 = lane.forward(tv)
    vb = torch.empty(tv.numel() + 1, device=gpu)[1:].view(lane.vol_shape)
    vb.copy_(tv)
    sb = torch.empty(f2.numel() + 1, device=gpu)[1:].view(lane.sino_shape)
    lane.forward(vb, out=sb)
    assert f2.data_ptr() % 8 == 0 and sb.data_ptr() % 8 == 4
    assert torch.equal(f2, sb)
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    b2 = lane.backward(ts)
    tb = torch.empty(ts.numel() + 1, device=gpu)[1:].view(lane.sino_shape)
    tb.copy_(ts)
    bb = torch.empty(b2.numel() + 1, device=gpu)[1:].view(lane.vol_shape)
    lane.backward(tb, out=bb)
    assert b2.data_ptr() % 8 == 0 and bb.data_ptr() % 8 == 4
    assert torch.equal(b2, bb)
    e1 = rel_l2(f2.cpu().numpy(), d1.forward(vol, ang, **kw))
    e2 = rel_l2(b2.cpu().numpy(), d1.backward(sino, ang, **kw))
    print("rel L2 forward %.2e backward %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL


@pytest.mark.parametrize("na", [1, 15, 16, 17, 47])
@pytest.mark.parametrize("nz", [40, 64, 128, 256])
def test_angle_batch_boundaries(gpu, na, nz):
    """The wavefront kernels evaluate geometry for 16 angles (back
    projection) or 64 samples (forward) at a time: angle counts below, at and
    just past a batch, with every z width (1, 2, 4 per lane), match the oracle
    and the per-lane-geometry kernels."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n = 21
    ang = np.linspace(3.0, 363.0, na, endpoint=False)
    rng = np.random.default_rng(na * 1000 + nz)
    vol = rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(0, 1, (n, na, nz))
    kw = dict(mu=0.015, det_offset=1.0)
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    wave = ParallelPlan(n, nz, ang, device=gpu, **kw)
    lane = ParallelPlan(n, nz, ang, device=gpu, per_lane_geometry=True, **kw)
    fw, bw = wave.forward(tv).cpu().numpy(), wave.backward(ts).cpu().numpy()
    fl, bl = lane.forward(tv).cpu().numpy(), lane.backward(ts).cpu().numpy()
    e1, e2 = rel_l2(fw, d1.forward(vol, ang, **kw)), rel_l2(bw, d1.backward(sino, ang, **kw))
    print("na %d nz %d: rel L2 forward %.2e backward %.2e" % (na, nz, e1, e2))
    assert e1 <= TOL and e2 <= TOL
    assert rel_l2(fw, fl) <= 1e-6 and rel_l2(bw, bl) <= 1e-6


def test_tuning_settings_agree(gpu):
    """Every work order (tile shape, z values per lane) gives the oracle's result."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz = 40, 128
    ang = np.arange(0.0, 360.0, 6.0)      # opposite angles: the paired candidates pair them
    rng = np.random.default_rng(9)
    vol = rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    plan = ParallelPlan(n, nz, ang, device=gpu, mu=0.02)
    ref_f, ref_b = d1.forward(vol, ang, mu=0.02), d1.backward(sino, ang, mu=0.02)
    for (ta, tk, zv, ru, pr), (tile, bzv) in zip(ParallelPlan.FP_CANDIDATES,
                                                 itertools.cycle(ParallelPlan.BP_CANDIDATES)):
        plan.set_tuning(fp_tile_angles=ta, fp_tile_rays=tk, fp_zv=zv, fp_reuse=ru, fp_pair=pr, bp_tile=tile,
                        bp_zv=bzv)
        e1 = rel_l2(plan.forward(tv).cpu().numpy(), ref_f)
        e2 = rel_l2(plan.backward(ts).cpu().numpy(), ref_b)
        assert e1 <= TOL and e2 <= TOL, (ta, tk, zv, ru, pr, tile, bzv, e1, e2)
    # the tile shape changes the order waves run in and tap reuse which
    # instruction loads a tap, not any value
    for pr in (0, 1):
        plan.set_tuning(fp_tile_angles=32, fp_tile_rays=32, fp_zv=2, fp_reuse=0, fp_pair=pr, bp_tile=16, bp_zv=2)
        a, b = plan.forward(tv), plan.backward(ts)
        plan.set_tuning(fp_tile_angles=8, fp_tile_rays=64, fp_reuse=1, bp_tile=7)
        assert torch.equal(a, plan.forward(tv)) and torch.equal(b, plan.backward(ts))
    rep = plan.autotune(tv, ts, reps=1)
    assert set(rep["choice"]) == {"fp_tile_angles", "fp_tile_rays", "fp_zv", "fp_reuse", "fp_pair", "bp_tile",
                                  "bp_zv"}
    tuned = plan.get_tuning()
    assert {k: tuned[k] for k in rep["choice"]} == rep["choice"] and tuned["z_split"] == 1


PAIR_CASES = [
    # n, nz, angles, det_offset, mu, data sign
    (40, 128, np.arange(0.0, 360.0, 3.0), 0.0, 0.02, "signed"),
    (41, 64, np.arange(0.0, 360.0, 4.0), 2.5, 0.0, "positive"),     # odd n, half-pixel offset; clipped
    (41, 64, np.arange(0.0, 360.0, 4.0), 2.5, 0.0, "signed"),       # waves project the rays separately
    (64, 40, np.arange(0.0, 360.0, 3.0), -3.0, 0.01, "signed"),     # partial z chunk
    (33, 256, np.concatenate([np.arange(0.0, 360.0, 10.0), [5.0, 95.0]]), 4.0, 0.01, "signed"),  # + single angles
    (32, 64, np.array([0.0, 0.0, 180.0, 45.0, 225.0, 225.0, 90.0, -90.0]), 1.5, 0.01, "signed"),  # duplicates, unsorted
]


@pytest.mark.parametrize("case", range(len(PAIR_CASES)))
def test_opposite_angle_pairing(gpu, case):
    """Projecting rays k at angle a and 2c + 2 det_offset - k at a + 180 deg
    from one pass over the taps (unclipped waves): the result matches the
    oracle and the unpaired kernel to float32 rounding; clipped waves project
    the two rays separately, bitwise as unpaired."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz, ang, off, mu, sign = PAIR_CASES[case]
    rng = np.random.default_rng(40 + case)
    vol = rng.uniform(-1, 1, (n, n, nz)) if sign == "signed" else rng.uniform(0.2, 1, (n, n, nz))
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    plan = ParallelPlan(n, nz, ang, det_offset=off, mu=mu, device=gpu)
    assert plan.get_tuning()["fp_pair"] == 1
    paired = plan.forward(tv).cpu().numpy()
    plan.set_tuning(fp_pair=0)
    single = plan.forward(tv).cpu().numpy()
    ref = d1.forward(vol, ang, det_offset=off, mu=mu)
    e_ref, e_pair = rel_l2(paired, ref), rel_l2(paired, single)
    print("case %d: rel L2 vs oracle %.2e, paired vs unpaired %.2e" % (case, e_ref, e_pair))
    assert e_ref <= TOL and e_pair <= 1e-6
    if sign == "positive":
        assert np.array_equal(paired, single)


@pytest.mark.parametrize("n,nz,na,mu", [(384, 77, 90, 0.013), (192, 128, 720, 0.004), (256, 33, 180, 0.0),
                                         (128, 100, 360, 0.02)])
def test_pairing_matches_unpaired_at_batch_shapes(gpu, n, nz, na, mu):
    """Paired against unpaired forward at the uneven-batch shapes (odd and
    partial z chunks, 90..720 angles, n up to 384): every sinogram value agrees
    to float32 rounding, not only the norm."""
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    rng = np.random.default_rng(n + nz)
    vol = torch.tensor(rng.uniform(-1, 1, (n, n, nz)), dtype=torch.float32, device=gpu)
    plan = ParallelPlan(n, nz, np.linspace(0, 360, na, endpoint=False), mu=mu, device=gpu)
    a = plan.forward(vol).cpu().numpy()
    plan.set_tuning(fp_pair=0)
    b = plan.forward(vol).cpu().numpy()
    scale = np.abs(b).max()
    err = np.abs(a - b).max() / scale
    print("n %d nz %d na %d: max |paired - unpaired| / max %.2e, rel L2 %.2e" % (n, nz, na, err, rel_l2(a, b)))
    assert err <= 2e-6 and rel_l2(a, b) <= 1e-6


# shapes where a lane holds one z (odd nz, or buffers 4 bytes off 8-byte
# alignment) and the 64-z chunk count is even: lanes then hold z and z + 64
SPLIT_CASES = [
    # n, nz, angles, det_offset, mu, buffers offset by one float
    (30, 77, np.arange(0.0, 360.0, 8.0), 0.0, 0.02, False),
    (23, 127, np.arange(0.0, 180.0, 6.0), 1.5, 0.0, False),
    (26, 65, np.arange(2.0, 360.0, 12.0), -2.0, 0.01, False),
    (20, 201, np.arange(0.0, 360.0, 15.0), 0.5, 0.0, False),         # 4 chunks of 64 -> 2 of 128
    (24, 100, np.arange(0.0, 360.0, 10.0), 1.0, 0.03, True),         # even nz, misaligned buffers
]


@pytest.mark.parametrize("case", range(len(SPLIT_CASES)))
@pytest.mark.parametrize("positive", [False, True])
def test_split_z_layout(gpu, case, positive):
    """Lanes holding z and z + 64 give the oracle's result and, z by z, the
    same operations as one z per lane: paired and unpaired forward and the
    back projection are bitwise equal with the split layout on and off."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz, ang, off, mu, shifted = SPLIT_CASES[case]
    print("case", case, n, nz, len(ang), off, mu, shifted, "positive" if positive else "signed", flush=True)
    rng = np.random.default_rng(500 + case)
    vol = rng.uniform(0.2, 1.0, (n, n, nz)) if positive else rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))

    def place(a):
        t = torch.tensor(a, dtype=torch.float32, device=gpu)
        if not shifted:
            return t
        buf = torch.zeros(t.numel() + 1, dtype=torch.float32, device=gpu)
        buf[1:] = t.reshape(-1)
        return buf[1:].view(t.shape)

    def empty(shape):
        buf = torch.zeros(int(np.prod(shape)) + 1, dtype=torch.float32, device=gpu)
        return (buf[1:] if shifted else buf[:-1]).view(shape)

    tv, ts = place(vol), place(sino)
    assert (tv.data_ptr() % 8 != 0) == shifted
    plan = ParallelPlan(n, nz, ang, det_offset=off, mu=mu, device=gpu)
    assert plan.get_tuning()["z_split"] == 1
    res = {}
    for split in (1, 0):
        for pair in (1, 0):
            plan.set_tuning(z_split=split, fp_pair=pair)
            res[("f", split, pair)] = plan.forward(tv, out=empty(plan.sino_shape)).cpu().numpy()
        res[("b", split)] = plan.backward(ts, out=empty(plan.vol_shape)).cpu().numpy()
    ref_f, ref_b = d1.forward(vol, ang, det_offset=off, mu=mu), d1.backward(sino, ang, det_offset=off, mu=mu)
    e1, e2 = rel_l2(res[("f", 1, 1)], ref_f), rel_l2(res[("b", 1)], ref_b)
    print("rel L2 forward %.2e backward %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL
    assert rel_l2(res[("f", 1, 0)], ref_f) <= TOL
    assert np.array_equal(res[("f", 1, 1)], res[("f", 0, 1)])
    assert np.array_equal(res[("f", 1, 0)], res[("f", 0, 0)])
    assert np.array_equal(res[("b", 1)], res[("b", 0)])


# --------------------------------------------------------------------------
# occlusion shadowing (Projector3DParallel.py:72-111, 147-153, 192-216)
# --------------------------------------------------------------------------
def test_occlusion_golden_vectors(gpu):
    """The reference's occlusion path, from scikit-image / numpy golden vectors."""
    import os
    import torch
    from vamtoolbox_amd.parallel import ParallelPlan
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "d1_occlusion_vectors.npz"))
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files})
    for key in keys:
        ins, ang = g[key + "_insert"], g[key + "_ang"]
        n, _, nz = ins.shape
        print("case", key, flush=True)
        plan = ParallelPlan(n, nz, ang, occlusion=ins, device=gpu)
        f = plan.forward(torch.tensor(g[key + "_vol"], dtype=torch.float32, device=gpu)).cpu().numpy()
        b = plan.backward(torch.tensor(g[key + "_sino"], dtype=torch.float32, device=gpu)).cpu().numpy()
        e1, e2 = rel_l2(f, g[key + "_fwd"]), rel_l2(b, g[key + "_bwd"])
        print("  rel L2 forward %.2e backward %.2e" % (e1, e2))
        assert e1 <= TOL and e2 <= TOL, key


OCC_CASES = [
    (40, 5, np.arange(0, 360, 9.0), dict(mu=0.02, det_offset=1.5)),
    (33, 64, np.arange(0, 180, 7.0), {}),                              # wave-eligible shape, occlusion -> per lane
    (24, 1, np.arange(0, 360, 5.0), dict(det_offset=-2.0)),
]


@pytest.mark.parametrize("case", range(len(OCC_CASES)))
def test_occlusion_parity(gpu, case):
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz, ang, kw = OCC_CASES[case]
    print("case", case, n, nz, len(ang), kw, flush=True)
    rng = np.random.default_rng(400 + case)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    rod = (xx - 0.2 * n) ** 2 + (yy - 0.05 * n) ** 2 <= (0.1 * n) ** 2
    ins = np.repeat(rod[:, :, None], nz, axis=2) & (rng.uniform(size=(1, 1, nz)) > 0.2)
    vol = rng.uniform(0.1, 1.0, (n, n, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    plan = ParallelPlan(n, nz, ang, occlusion=ins, device=gpu, **kw)
    occ = d1.occlusion_sinogram(ins, ang, det_offset=kw.get("det_offset", 0.0))
    f = plan.forward(torch.tensor(vol, dtype=torch.float32, device=gpu)).cpu().numpy()
    b = plan.backward(torch.tensor(sino, dtype=torch.float32, device=gpu)).cpu().numpy()
    e1 = rel_l2(f, d1.forward(vol, ang, occ=occ, **kw))
    e2 = rel_l2(b, d1.backward(sino, ang, occ=occ, **kw))
    print("rel L2 forward %.2e backward %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL
    # the shadow removes something: differs from the unoccluded projection
    free = ParallelPlan(n, nz, ang, device=gpu, **kw)
    assert rel_l2(f, free.forward(torch.tensor(vol, dtype=torch.float32, device=gpu)).cpu().numpy()) > 1e-3


def test_zero_dose_constraint(gpu):
    """Zero-dose body: the sinogram entries whose rays cross it are zeroed
    before every back projection (geometry.py:157-159,
    Projector3DParallel.py:43-44 via projectorconstructor.py:162-163)."""
    import oracle.d1 as d1
    import vamtoolbox_amd as vam
    n, nz = 28, 3
    ang = np.arange(0.0, 180.0, 6.0)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    zd = np.repeat(((xx - 7) ** 2 + (yy + 3) ** 2 <= 6)[:, :, None], nz, axis=2).astype(np.float64)
    rng = np.random.default_rng(21)
    tg = vam.geometry.TargetGeometry(target=rng.uniform(0, 1, (n, n, nz)), zero_dose=zd)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True)
    A = vam.projectorconstructor.projectorconstructor(tg, pg)
    zs = d1.forward(zd, ang) != 0
    assert np.array_equal(np.asarray(pg.zero_dose_sino), zs)
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    got = A.backward(sino)
    ref = d1.backward(np.where(zs, 0.0, sino), ang)
    e = rel_l2(got, ref)
    print("zero-dose backward rel L2 %.2e, %d of %d sinogram entries zeroed" % (e, zs.sum(), zs.size))
    assert e <= TOL and zs.sum() > 0



def test_radial_absorption_mask(gpu):
    """The radial absorption mask of geometry.py:161-193 multiplies the
    volume before the forward projection and the reconstruction after the
    back projection (Projector3DParallelCUDA.py:43-44, 66-67); optimize()
    builds it whenever absorption_coeff is set (optimize.py:199-200)."""
    import oracle.d1 as d1
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    n, nz = 36, 5
    ang = np.arange(0.0, 360.0, 8.0)
    rng = np.random.default_rng(22)
    kw = dict(absorption_coeff=0.08, container_radius=n / 2 * 0.1 + 0.9, projector_pixel_size=0.1)
    tg = vam.geometry.TargetGeometry(target=rng.uniform(0, 1, (n, n, nz)))
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True, **kw)
    pg.calcAbsorptionMask(tg)
    mask = np.asarray(pg.absorption_mask)
    assert mask.shape == (n, n, nz) and mask.max() < 1.0 and np.count_nonzero(mask == 0) > 0
    A = vam.projectorconstructor.projectorconstructor(tg, pg)
    x = rng.uniform(0, 1, (n, n, nz))
    sino = rng.uniform(0, 1, (n, len(ang), nz))
    e_f = rel_l2(A.forward(x), d1.forward(mask * x, ang))
    e_b = rel_l2(A.backward(sino), mask * d1.backward(sino, ang))
    print("radial mask: forward rel L2 %.2e backward %.2e" % (e_f, e_b))
    assert e_f <= TOL and e_b <= TOL
    # through optimize(): OSMO with the mask against the oracle loop on masked operators
    t = np.zeros((n, n, nz))
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    t[(xx / 9) ** 2 + (yy / 6) ** 2 <= 1] = 1.0
    tg = vam.geometry.TargetGeometry(target=t)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True, **kw)
    opts = vam.optimize.Options(method="OSMO", n_iter=4, d_h=0.85, d_l=0.6)
    s_gpu, r_gpu, _ = vam.optimize.optimize(tg, pg, opts)
    assert pg.absorption_mask is not None
    b_ref, x_ref, _ = oo.osmo(t, lambda v: d1.forward(mask * v, ang), lambda s: mask * d1.backward(s, ang), 4,
                              d_h=0.85, d_l=0.6)
    e_s, e_r = rel_l2(s_gpu.array, b_ref), rel_l2(r_gpu.array, x_ref)
    print("OSMO with the radial mask: rel L2 sinogram %.2e recon %.2e" % (e_s, e_r))
    assert e_s <= 1e-4 and e_r <= 1e-4


ADJ_WAVE_CASES = [
    # n, nz, angles, kwargs: z widths 4 / 2 / 1, partial 64-z chunks, odd n,
    # absorption and offset, angle counts around the 64-angle lane batches
    (32, 256, np.arange(0, 360, 5.0), dict(mu=0.01, det_offset=2.0)),
    (33, 130, np.arange(0, 360, 12.0), {}),
    (24, 77, np.arange(0, 180, 2.5), dict(det_offset=-1.5)),
    (40, 64, np.linspace(0, 360, 65, endpoint=False), dict(mu=0.03)),
    (16, 40, np.arange(0, 360, 3.0), {}),
    # odd detector shift 2 det_offset = 1 (opposite bins k and 2c + 1 - k) and
    # 45-degree multiples, where cos and sin tie
    (24, 128, np.arange(0, 360, 15.0), dict(mu=0.02, det_offset=0.5)),
]


@pytest.mark.parametrize("case", range(len(ADJ_WAVE_CASES)))
def test_adjoint_wave_kernels_bitwise(gpu, case):
    """The wavefront-uniform exact transposes sum the same terms in the same
    order as the per-lane kernels: bitwise equal results, and within 1e-5 of
    the oracle's transposes."""
    import torch
    import oracle.d1 as d1
    from vamtoolbox_amd.parallel import ParallelPlan
    n, nz, ang, kw = ADJ_WAVE_CASES[case]
    print("case", case, n, nz, len(ang), kw, flush=True)
    rng = np.random.default_rng(300 + case)
    vol = rng.uniform(-1, 1, (n, n, nz))
    sino = rng.uniform(-1, 1, (n, len(ang), nz))
    wave = ParallelPlan(n, nz, ang, device=gpu, warp_clip=False, **kw)
    lane = ParallelPlan(n, nz, ang, device=gpu, warp_clip=False, per_lane_geometry=True, **kw)
    tv = torch.tensor(vol, dtype=torch.float32, device=gpu)
    ts = torch.tensor(sino, dtype=torch.float32, device=gpu)
    fw, fl = wave.forward_adjoint(ts), lane.forward_adjoint(ts)
    bw, bl = wave.backward_adjoint(tv), lane.backward_adjoint(tv)
    assert torch.equal(fw, fl)
    assert torch.equal(bw, bl)
    # B^T pairs opposite bins like the forward (one pass over the shared
    # strip); unpaired it runs one bin per wavefront: bitwise the same
    wave.set_tuning(fp_pair=0)
    assert torch.equal(wave.backward_adjoint(tv), bl)
    z = [0, nz // 2, nz - 1]
    e1 = rel_l2(fw[:, :, z].cpu().numpy(), d1.forward_adjoint(sino[:, :, z], ang, **kw))
    e2 = rel_l2(bw[:, :, z].cpu().numpy(), d1.backward_adjoint(vol[:, :, z], ang, **kw))
    print("rel L2 F^T %.3e B^T %.3e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL
