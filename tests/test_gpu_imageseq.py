"""Image-sequence export on the GPU (vamtoolbox_amd.imagesequence, through
vamx_imgseq_*) against the oracle restatement of imagesequence.py:85-399."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # screen (N_u, N_v), options
    ((64, 48), {}),
    ((64, 48), dict(invert_u=True, invert_v=True, u_offset=3, v_offset=-2)),
    ((96, 80), dict(array_num=3, array_offset=14, v_offset=2)),
    ((96, 80), dict(array_num=4, array_offset=16, invert_v=True)),
    ((64, 48), dict(normalization_percentile=99.5, intensity_scale=0.8)),
    ((64, 48), dict(normalization_percentile=97.0, bit_depth=10)),
    ((64, 48), dict(normalization_percentile=100.0, bit_depth=16, intensity_scale=1.3)),
    ((64, 64), dict(rotate_angle=17.5, invert_u=True, normalization_percentile=99.0)),
    ((64, 64), dict(rotate_angle=-90.0)),
    ((96, 80), dict(size_scale=1.5, invert_v=True, normalization_percentile=99.0)),
    ((64, 48), dict(size_scale=0.7)),
    ((128, 96), dict(rotate_angle=33.0, size_scale=1.25, array_num=2, array_offset=40,
                     normalization_percentile=99.9)),
]


def _sino(seed, nu=24, na=7, nv=20):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0, 300, (nu, na, nv))
    s[rng.uniform(size=s.shape) < 0.05] = 0.0
    return s


@pytest.mark.parametrize("case", range(len(CASES)))
def test_frames_match_oracle(gpu, case):
    import oracle.imageseq as oi
    from vamtoolbox_amd.imagesequence import ImageConfig, ImageSeq
    dims, kw = CASES[case]
    s = _sino(case)
    seq = ImageSeq(ImageConfig(dims, **kw), s)
    got = seq.frames.cpu().numpy()
    ref = oi.frames(s, dims[0], dims[1], **kw)
    assert got.shape == ref.shape and got.dtype == np.uint8
    diff = got.astype(int) - ref.astype(int)
    off = np.count_nonzero(diff)
    print("case %d %s: %d of %d pixels differ (max %d)" % (case, kw, off, got.size, np.abs(diff).max()))
    # the kernels evaluate every float64 expression in the oracle's (and
    # scipy's) operation order, so even rotated and scaled frames agree bit for bit
    assert off == 0
    assert len(seq.images) == s.shape[1] and np.array_equal(seq.images[3], ref[3])


def test_rotate_matches_scipy(gpu):
    """vamx_imgseq_rotate against scipy.ndimage.rotate(axes=(0, 2),
    reshape=True, order=1) itself, the reference's call, on float64 data."""
    import ctypes
    import torch
    from scipy import ndimage
    from vamtoolbox_amd import _lib
    from vamtoolbox_amd.imagesequence import _rotation
    lib = _lib.load()
    rng = np.random.default_rng(4)
    for shape, ang in (((30, 5, 22), 12.0), ((17, 3, 40), -71.3), ((25, 2, 25), 90.0), ((9, 4, 13), 180.0),
                       ((40, 2, 31), 45.0)):
        s = rng.uniform(-1, 1, shape)
        ref = ndimage.rotate(s, ang, axes=(0, 2), reshape=True, order=1)
        (mu, mv), rot, off = _rotation((shape[0], shape[2]), ang)
        assert (mu, shape[1], mv) == ref.shape
        t = torch.tensor(s, device=gpu)
        out = torch.empty(ref.shape, device=gpu, dtype=torch.float64)
        _lib.check(lib.vamx_imgseq_rotate(t.data_ptr(), shape[0], shape[1], shape[2], 0, 0,
                                          (ctypes.c_double * 4)(*rot.ravel()), (ctypes.c_double * 2)(*off),
                                          out.data_ptr(), mu, mv, _lib.stream_ptr(gpu)), "vamx_imgseq_rotate")
        d = np.abs(out.cpu().numpy() - ref)
        print("rotate %s by %g: max abs diff %.2e" % (shape, ang, d.max()))
        assert d.max() <= 1e-12


def test_full_size_sequence(gpu):
    """360 frames of 1920 x 1080 from a 256 x 360 x 256 sinogram (0.75 GB
    of frames), two arrayed copies and 99.5th-percentile normalisation,
    against the oracle on four angles (the normalisation value from the whole
    sinogram, then the oracle on those angles' slices with it)."""
    import torch
    import oracle.imageseq as oi
    from vamtoolbox_amd.imagesequence import ImageConfig, ImageSeq
    rng = np.random.default_rng(7)
    s = rng.uniform(0, 1, (256, 360, 256))
    cfg = ImageConfig((1920, 1080), array_num=2, array_offset=300, normalization_percentile=99.5, u_offset=40)
    seq = ImageSeq(cfg, s)
    torch.cuda.synchronize()
    norm = np.percentile(s, 99.5)
    for j in (0, 91, 200, 359):
        sub = s[:, j:j + 1, :] / norm * 255   # the oracle with the normalisation applied by hand
        ref = oi.frames(sub, 1920, 1080, array_num=2, array_offset=300, u_offset=40)
        np.testing.assert_array_equal(seq.frames[j].cpu().numpy(), ref[0])


def test_percentile_order_statistics_exact(gpu):
    """The percentile normalisation (np.percentile, imagesequence.py:128-131)
    from vamx_order_stats: equal to numpy's value for every case (ties, signed
    zeros, infinities, NaN, one and two elements, the q extremes, the
    benchmark's 256 x 360 x 256 sinogram), and each order statistic equal to
    the sorted array's entry."""
    import torch
    from vamtoolbox_amd.imagesequence import _order_stats, _percentile
    rng = np.random.default_rng(7)
    cases = [
        rng.standard_normal(1000),
        np.repeat(np.arange(10.0), 1001)[rng.permutation(10010)],
        np.full(777, 3.25),
        np.array([0.0, -0.0, 1.0, -1.0, -0.0, 0.0, 2.0]),
        np.array([np.inf, -np.inf, 1.0, 5.0, -3.0]),
        np.array([4.5]),
        np.array([2.0, -7.0]),
        rng.uniform(-1e300, 1e300, 4097),
        np.concatenate([rng.uniform(0, 1, 500), rng.uniform(0, 1e-300, 500), [5e-324, -5e-324]]),
    ]
    for x in cases:
        t = torch.tensor(x, dtype=torch.float64, device=gpu)
        for q in (0, 0.1, 25, 50, 99.5, 100):
            with np.errstate(invalid="ignore"):   # inf - inf in the interpolation, as NumPy's
                ref = float(np.percentile(x, q))
                got = _percentile(t, q)
            assert got == ref or (np.isnan(got) and np.isnan(ref)), (x[:5], q, got, ref)
        srt = np.sort(x)
        for r in sorted({0, len(x) // 3, len(x) - 1}):
            a, b = _order_stats(t, r)
            assert a == srt[r] and b == srt[min(r + 1, len(x) - 1)], (r, a, b)
    # an array starting 8 bytes past a 16-byte boundary (one double per load)
    x = rng.standard_normal(3001)
    t = torch.tensor(x, dtype=torch.float64, device=gpu)[1:]
    assert t.data_ptr() % 16 == 8
    for q in (0, 37.5, 99.5, 100):
        assert _percentile(t, q) == float(np.percentile(x[1:], q))
    x = rng.uniform(0, 1, 1000)
    x[123] = np.nan
    assert np.isnan(_percentile(torch.tensor(x, device=gpu), 50)) and np.isnan(np.percentile(x, 50))
    big = torch.rand((256, 360, 256), dtype=torch.float64, device=gpu, generator=torch.Generator(gpu).manual_seed(3))
    ref = float(np.percentile(big.cpu().numpy(), 99.5))
    assert _percentile(big, 99.5) == ref
