"""Pinning of the image-sequence oracle (oracle/imageseq.py) on the CPU, and
the host-side logic of vamtoolbox_amd.imagesequence (no GPU)."""

import numpy as np
import pytest

import oracle.imageseq as oi


def check_place_known_answer():
    """imagesequence.py:300-361 by hand on a 64 x 48 screen, 20 x 10 image:
    u1 = int(32 - 2 - 10) = 20, v1 = int(24 - 3 - extra - 5) = 16 - extra.
    Odd array_num = 3, pitch 12: extra = int(12 d + 3) for d = 0, 1, -1, so
    v_offset counts twice for arrayed copies.  Even array_num = 2, pitch 10:
    extra = int(10 d / 2 + 3) for d = 1, -1."""
    assert oi.place(64, 48, 20, 10, u_offset=2, v_offset=3) == [(16, 26, 20, 40)]
    assert oi.place(64, 48, 20, 10, 2, 3, array_num=3, array_offset=12) == [(13, 23, 20, 40), (1, 11, 20, 40),
                                                                              (25, 35, 20, 40)]
    assert oi.place(64, 48, 20, 10, 2, 3, array_num=2, array_offset=10) == [(8, 18, 20, 40), (18, 28, 20, 40)]
    with pytest.raises(Exception):
        oi.place(64, 48, 20, 10, u_offset=30)
    with pytest.raises(Exception):
        oi.place(64, 48, 20, 10, 0, 0, array_num=3, array_offset=40)


def check_resize_closed_forms():
    """INTER_LINEAR: the same size is a copy; a constant stays constant; a
    2x upscale of a ramp samples it at (d + 0.5) / 2 - 0.5 (clamped at the
    edges); a 2x downscale averages pixel pairs."""
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 1, (7, 9))
    np.testing.assert_array_equal(oi.resize_linear(a, 9, 7), a)
    np.testing.assert_allclose(oi.resize_linear(np.full((5, 6), 2.5), 11, 3), 2.5, rtol=0, atol=1e-15)
    ramp = np.tile(np.arange(8.0), (3, 1))
    up = oi.resize_linear(ramp, 16, 3)
    expect = np.clip((np.arange(16) + 0.5) / 2 - 0.5, 0, 7)
    np.testing.assert_allclose(up, np.tile(expect, (3, 1)), rtol=0, atol=1e-14)
    down = oi.resize_linear(a[:, :8], 4, 7)
    np.testing.assert_allclose(down, 0.5 * (a[:, 0:8:2] + a[:, 1:8:2]), rtol=0, atol=1e-15)
    # rows and columns are independent: a ramp along rows
    col = oi.resize_linear(np.tile(np.arange(4.0)[:, None], (1, 5)), 5, 8)
    np.testing.assert_allclose(col[:, 0], np.clip((np.arange(8) + 0.5) / 2 - 0.5, 0, 3), atol=1e-14)


def check_frames_known_answer():
    """The value path by hand: bit depth 8 truncates toward zero and clips
    at 255; bit depth 10 clips at 1023 and keeps the low byte in the uint8
    frame; the 100th percentile normalisation maps the maximum to 2^b - 1;
    negative samples give 0; the (v, u) image is the transposed (u, v)
    sinogram slice; invert_u flips u."""
    s = np.zeros((4, 2, 4))
    s[:, 0, :] = np.array([[1.7, 300.0, -2.0, 0.0], [2.2, 3.9, 4.0, 5.5], [0.0, 0.0, 0.0, 9.0],
                           [1.0, 2.0, 3.0, 4.0]])
    s[:, 1, :] = 7.0
    f = oi.frames(s, 8, 8)
    assert f.shape == (2, 8, 8) and f.dtype == np.uint8
    # image window: u1 = int(4 - 2) = 2, v1 = 2; frame[v, u] = s[u - 2, j, v - 2]
    np.testing.assert_array_equal(f[0, 2:6, 2:6], np.array([[1, 2, 0, 1], [255, 3, 0, 2], [0, 4, 0, 3],
                                                            [0, 5, 9, 4]], dtype=np.uint8))
    assert f[0].sum() == np.sum(f[0, 2:6, 2:6].astype(int)) and np.all(f[1, 2:6, 2:6] == 7)
    f10 = oi.frames(s, 8, 8, bit_depth=10)
    assert f10[0, 3, 2] == 300 % 256 and f10[0, 2, 2] == 1
    fn = oi.frames(s, 8, 8, normalization_percentile=100, bit_depth=8)
    assert fn[0, 3, 2] == 255 and fn[1, 2, 2] == int(7.0 / 300.0 * 255)
    fu = oi.frames(s, 8, 8, invert_u=True)
    np.testing.assert_array_equal(fu[0, 2:6, 2:6], f[0, 2:6, 2:6][:, ::-1])
    fv = oi.frames(s, 8, 8, invert_v=True)
    np.testing.assert_array_equal(fv[0, 2:6, 2:6], f[0, 2:6, 2:6][::-1, :])


def check_rotate_is_scipy():
    """The rotation is the reference's own scipy call; a quarter turn of the
    (u, v) plane is an exact index permutation."""
    rng = np.random.default_rng(1)
    s = rng.uniform(0, 1, (6, 3, 4))
    r = oi.rotate(s, 90.0)
    assert r.shape == (4, 3, 6)
    np.testing.assert_allclose(r, np.transpose(s, (2, 1, 0))[::-1, :, :], atol=1e-12)


def test_place_known_answer():
    check_place_known_answer()


def test_resize_closed_forms():
    check_resize_closed_forms()


def test_frames_known_answer():
    check_frames_known_answer()


def test_rotate_is_scipy():
    check_rotate_is_scipy()


def test_percentile_two_order_statistics(monkeypatch):
    """The host half of vamtoolbox_amd.imagesequence._percentile (the rank it
    asks the device for and NumPy's linear-method lerp between the two order
    statistics), with the device selection replaced by a sort: equal to
    np.percentile on the lerp branches, the ends, ties and infinities."""
    import warnings
    import torch
    import vamtoolbox_amd.imagesequence as iseq

    def order_stats(flat, rank):   # vamx_order_stats' contract: NaN when any is NaN
        srt = np.sort(flat.numpy())
        if np.isnan(srt).any():
            return float("nan"), float("nan")
        return float(srt[rank]), float(srt[min(rank + 1, srt.size - 1)])

    monkeypatch.setattr(iseq, "_order_stats", order_stats)
    rng = np.random.default_rng(2)
    for a in (rng.uniform(-3, 7, 1001), rng.integers(0, 5, 400).astype(float), np.array([2.5]),
              np.array([np.inf, -np.inf, 1.0, 5.0, -3.0]), np.array([3.0, np.nan, 1.0])):
        t = torch.as_tensor(a)
        for q in (0, 0.1, 12.5, 50, 62.5, 99.5, 99.99, 100):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                got, ref = iseq._percentile(t, q), float(np.percentile(a, q))
            assert got == ref or (np.isnan(got) and np.isnan(ref)), (a.size, q, got, ref)


def test_image_config_and_windows():
    """ImageConfig's checks and defaults (imagesequence.py:14-82) and the host
    placement logic against the oracle's."""
    from vamtoolbox_amd.imagesequence import ImageConfig, _windows
    with pytest.raises(Exception):
        ImageConfig((30, 40))
    c = ImageConfig((64, 48), u_offset=2, v_offset=3, array_num=3, array_offset=12)
    assert c.bit_depth == 8 and c.size_scale == 1.0 and c.normalization_percentile is None
    assert _windows(c, 20, 10) == [(v1, u1) for v1, _, u1, _ in oi.place(64, 48, 20, 10, 2, 3, 3, 12)]
    c = ImageConfig((64, 48), array_num=2, array_offset=10, v_offset=3, u_offset=2)
    assert _windows(c, 20, 10) == [(v1, u1) for v1, _, u1, _ in oi.place(64, 48, 20, 10, 2, 3, 2, 10)]
    # int(a + s / 2) - int(a - s / 2) = s for a >= s / 2: odd sizes fit too
    assert _windows(ImageConfig((64, 48)), 21, 11) == [(18, 21)]
