"""Pinning of the oracle's optimisation loops (CPU).

The OSMO / CAL / BCLP loops of ``oracle/optim.py`` are restatements of
``vamtoolbox/optimizer/OSMO.py:36-162``, ``CAL.py:116-218`` and
``BCLP.py:477-526``.  They are checked here by properties that the reference's
loops have and that a plausible one-line mistake breaks:

* the max normalisations make OSMO and CAL invariant to a constant scale of the
  projector pair;
* a target already inside the OSMO band is a fixed point;
* the void step changes the model only on void voxels (downwards), the gel
  step only on gel voxels (upwards), and voxels in (d_l, d_h) are corrected by
  both (a swapped d_h / d_l leaves them alone);
* iteration 0 is the initial state: exactly 1 + 2 (n - 1) forward projections
  for OSMO and 1 + (n - 1) for CAL;
* CAL with lr = 0 keeps its initial iterate; its first momentum step scales
  the gradient by (1 - m);
* BCLP on identity operators with an identity response has closed-form
  iterates: for p = q = 2, g_k - T = (1 - 2 lr)^k (g_0 - T); for p = 2,
  q = 1 the gradient is normalised, so ||g_k - T|| falls by exactly lr per
  iteration.

``tests/test_oracle_mutations.py`` plants one-line mistakes in the oracle and
checks that these tests catch each of them.
"""

import numpy as np

import oracle.optim as oo


# --------------------------------------------------------------------------
# helpers
# --------------------------------------------------------------------------
class Recorder:
    """A linear operator (matrix on the flattened array) that records its inputs."""

    def __init__(self, M, out_shape, scale=1.0):
        self.M, self.out_shape, self.scale = M, out_shape, scale
        self.calls = []

    def __call__(self, x):
        self.calls.append(np.array(x, dtype=np.float64, copy=True))
        return self.scale * (self.M @ np.asarray(x, dtype=np.float64).reshape(-1)).reshape(self.out_shape)


def _tiny_problem(seed=0, n=8, nz=2, n_s=40):
    """A binary target inside the inscribed circle and a random non-negative
    projector pair (B = F^T) of full rank."""
    rng = np.random.default_rng(seed)
    t = np.zeros((n, n, nz))
    t[2:6, 3:6, :] = 1.0
    F = rng.uniform(0.0, 1.0, (n_s, n * n * nz))
    # scaled so that the dose B F T peaks at 1: the thresholds d_h, d_l and
    # CAL's sigmoid then act on unsaturated values
    F /= np.sqrt(np.max(F.T @ (F @ t.reshape(-1))))
    return t, F, rng


def _pair(F, t_shape, scale=1.0):
    fwd = Recorder(F, (F.shape[0],), scale)
    bwd = Recorder(F.T, t_shape, scale)
    return fwd, bwd


class IdentityResponse:
    def map(self, f):
        return np.asarray(f, dtype=np.float64)

    def dmap(self, f):
        return np.ones_like(np.asarray(f, dtype=np.float64))


# --------------------------------------------------------------------------
# OSMO
# --------------------------------------------------------------------------
def check_osmo_scale_invariance():
    t, F, _ = _tiny_problem(1)
    f1, b1 = _pair(F, t.shape)
    f2, b2 = _pair(F, t.shape, scale=3.7)
    bA, xA, eA = oo.osmo(t, f1, b1, 5, filt=None)
    bB, xB, eB = oo.osmo(t, f2, b2, 5, filt=None)
    np.testing.assert_allclose(xB, xA, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(bB, bA, rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(eA, eB)


def check_osmo_fixed_point():
    """Identity operators and no filter: x_0 = T, which lies inside the band
    (void 0 <= d_l, gel 1 >= d_h), so every model update is zero."""
    n, nz = 8, 2
    t = np.zeros((n, n, nz))
    t[2:6, 2:5, :] = 1.0
    eye = np.eye(t.size)
    fwd, bwd = _pair(eye, t.shape)
    fwd.out_shape = t.shape
    b, x, err = oo.osmo(t, fwd, bwd, 4, filt=None)
    np.testing.assert_array_equal(x, t)
    np.testing.assert_array_equal(b, t)
    np.testing.assert_array_equal(err, 0.0)
    for c in fwd.calls:
        np.testing.assert_array_equal(c, t)


def check_osmo_step_order_and_band():
    """Voxels whose dose lies in (d_l, d_h): the void step lowers the model on
    void voxels only, the gel step raises it on gel voxels only."""
    n, nz = 8, 1
    t = np.zeros((n, n, nz))
    t[2:6, 2:6, 0] = 1.0
    gel, void = oo.circle_inds(t)
    # dose operator: a diagonal (0 .. 0.3 over the gel voxels) plus a rank-one
    # term that lifts every voxel: void doses ~0.71 of the maximum, gel doses
    # spread over [0.71, 1]
    N = t.size
    diag = np.full(N, 0.3)
    gi = np.flatnonzero(t.reshape(-1) > 0)
    diag[gi] = np.linspace(0.0, 0.3, gi.size)
    B = np.diag(diag) + np.full((N, N), 1.0 / N) * 3.0
    fwd = Recorder(np.eye(N), t.shape)
    bwd = Recorder(B, t.shape)
    oo.osmo(t, fwd, bwd, 2, filt=None)
    assert len(fwd.calls) == 3 and len(bwd.calls) == 3
    x0 = bwd.scale * (B @ np.clip(t.reshape(-1), 0, None)).reshape(t.shape)
    x0 = x0 / x0.max()
    assert np.all((x0[void] > 0.7) & (x0[void] < 0.8)), "set-up: void doses must sit in (d_l, d_h)"
    xm0, xm1, xm2 = fwd.calls
    d1 = xm1 - xm0
    assert np.all(d1[void] < 0), "void step must lower every void voxel above d_l"
    assert np.all(d1[~void] == 0), "void step must not touch other voxels"
    d2 = xm2 - xm1
    assert np.all(d2[void] == 0) and np.all(d2[~(gel | void)] == 0)
    x1 = bwd.calls[1]  # input of the second dose evaluation: b after the void step
    xv = (B @ x1.reshape(-1)).reshape(t.shape)
    xv = np.clip(xv / xv.max(), 0, None)
    low = gel & (xv < 0.8)
    assert low.any(), "set-up: some gel doses must sit below d_h"
    assert np.all(d2[low] > 0) and np.all(d2[gel & ~low] == 0)


def check_osmo_iteration_count():
    t, F, _ = _tiny_problem(2)
    fwd, bwd = _pair(F, t.shape)
    n_iter = 6
    _, x, err = oo.osmo(t, fwd, bwd, n_iter, filt=None)
    assert len(fwd.calls) == 1 + 2 * (n_iter - 1)
    assert len(bwd.calls) == 1 + 2 * (n_iter - 1)
    # err[0] is the VER of the initial dose (no update before it)
    gel, void = oo.circle_inds(t)
    x0 = (F.T @ np.clip(F @ t.reshape(-1), 0, None)).reshape(t.shape)
    x0 = x0 / x0.max()
    assert err[0] == oo.ver(gel, void, x0)
    assert np.all(np.isfinite(err))


def test_osmo_scale_invariance():
    check_osmo_scale_invariance()


def test_osmo_fixed_point():
    check_osmo_fixed_point()


def test_osmo_step_order_and_band():
    check_osmo_step_order_and_band()


def test_osmo_iteration_count():
    check_osmo_iteration_count()


# --------------------------------------------------------------------------
# CAL
# --------------------------------------------------------------------------
def check_cal_lr_zero():
    t, F, _ = _tiny_problem(3)
    fwd, bwd = _pair(F, t.shape)
    b, x, err = oo.cal(t, fwd, bwd, 5, filt=None, lr=0.0)
    b0 = np.clip(F @ t.reshape(-1), 0, None)
    np.testing.assert_allclose(b, b0 / b0.max(), rtol=1e-13)
    x0 = (F.T @ b0).reshape(t.shape)
    np.testing.assert_allclose(x, x0 / x0.max(), rtol=1e-13)
    assert np.all(err == err[0])


def check_cal_scale_invariance():
    t, F, _ = _tiny_problem(4)
    f1, b1 = _pair(F, t.shape)
    f2, b2 = _pair(F, t.shape, scale=2.5)
    # lr scales with the operator's square so the iterates stay comparable:
    # b ~ c, grad ~ c -> b - lr grad keeps its direction for any c
    bA, xA, eA = oo.cal(t, f1, b1, 5, filt=None, lr=1e-3, d_h=0.8, sigmoid=0.05)
    bB, xB, eB = oo.cal(t, f2, b2, 5, filt=None, lr=1e-3, d_h=0.8, sigmoid=0.05)
    np.testing.assert_allclose(xB, xA, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(bB, bA, rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(eA, eB)


def check_cal_iteration_count_and_direction():
    """Identity operators: each iteration subtracts lr * (sigmoid(x) - T) once;
    where the thresholded dose exceeds the target the sinogram falls."""
    n, nz = 6, 1
    t = np.zeros((n, n, nz))
    t[1:4, 1:4, 0] = 1.0
    eye = np.eye(t.size)
    fwd = Recorder(eye, t.shape)
    bwd = Recorder(eye, t.shape)
    n_iter = 4
    oo.cal(t, fwd, bwd, n_iter, filt=None, lr=0.05, d_h=0.5, sigmoid=0.2)
    assert len(fwd.calls) == 1 + (n_iter - 1)
    assert len(bwd.calls) == 1 + (n_iter - 1) + 1
    # the second dose evaluation's input is b_1 = b_0 - lr * (xt_0 - T)
    b0, b1 = bwd.calls[0], bwd.calls[1]
    xt0 = fwd.calls[1] + t                 # fwd is called with xt - T
    free = (xt0 != t) & (b1 > 0)           # where the clip at 0 does not bind
    assert free.sum() >= 4
    assert np.all(np.sign(b0 - b1)[free] == np.sign(xt0 - t)[free])
    assert np.all(b1[(b0 == 0) & (xt0 > t)] == 0)


def check_cal_momentum_first_step():
    """At the first iteration db_prev = 0, so the update is lr (1 - m) db: one
    iteration with (lr, m) equals one with (lr (1 - m), 0)."""
    t, F, _ = _tiny_problem(5)
    f1, b1 = _pair(F, t.shape)
    f2, b2 = _pair(F, t.shape)
    m = 0.4
    bA, xA, _ = oo.cal(t, f1, b1, 2, filt=None, lr=2e-3, momentum=m, d_h=0.8, sigmoid=0.05)
    bB, xB, _ = oo.cal(t, f2, b2, 2, filt=None, lr=2e-3 * (1 - m), momentum=0.0, d_h=0.8, sigmoid=0.05)
    np.testing.assert_allclose(bA, bB, rtol=1e-12)
    np.testing.assert_allclose(xA, xB, rtol=1e-12)


def test_cal_lr_zero():
    check_cal_lr_zero()


def test_cal_scale_invariance():
    check_cal_scale_invariance()


def test_cal_iteration_count_and_direction():
    check_cal_iteration_count_and_direction()


def test_cal_momentum_first_step():
    check_cal_momentum_first_step()


# --------------------------------------------------------------------------
# BCLP
# --------------------------------------------------------------------------
def _bclp_identity(p, q, lr, n_iter, eps=0.0, seed=6):
    rng = np.random.default_rng(seed)
    T = rng.uniform(0.4, 0.9, (5, 4, 3))
    g0 = T + rng.uniform(-0.3, 0.3, T.shape)
    ident = lambda a: np.asarray(a, dtype=np.float64)
    return T, g0, oo.bclp(T, ident, ident, IdentityResponse(), n_iter, g0, eps=eps, p=p, q=q, lr=lr)


def check_bclp_closed_form_p2_q2():
    lr, n_iter = 0.1, 7
    T, g0, (g, dose, m, losses) = _bclp_identity(2.0, 2.0, lr, n_iter)
    np.testing.assert_allclose(g - T, (1 - 2 * lr) ** n_iter * (g0 - T), rtol=1e-12, atol=1e-15)
    L0 = np.sum((g0 - T) ** 2)
    k = np.arange(n_iter + 1)
    np.testing.assert_allclose(losses, L0 * (1 - 2 * lr) ** (2 * k), rtol=1e-12)
    np.testing.assert_allclose(dose, g, rtol=0, atol=0)


def check_bclp_closed_form_p2_q1():
    lr, n_iter = 0.05, 6
    T, g0, (g, dose, m, losses) = _bclp_identity(2.0, 1.0, lr, n_iter)
    r0 = np.linalg.norm(g0 - T)
    assert r0 > (n_iter + 1) * lr
    np.testing.assert_allclose(losses, r0 - lr * np.arange(n_iter + 1), rtol=1e-12)
    u = (g0 - T) / r0
    np.testing.assert_allclose(g - T, (r0 - n_iter * lr) * u, rtol=1e-10, atol=1e-14)


def check_bclp_band_and_bounds():
    """Voxels whose error is inside the band do not move; the sinogram stays
    within [glb, gub]."""
    rng = np.random.default_rng(7)
    T = rng.uniform(0.2, 0.8, (6, 5, 2))
    g0 = T.copy()
    g0[0] += 0.3          # outside the band
    g0[1] += 0.01         # inside the band (eps = 0.05)
    g0[2] -= 0.15         # outside the band, below the target
    ident = lambda a: np.asarray(a, dtype=np.float64)
    g, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 3, g0, eps=0.05, p=2.0, q=2.0, lr=0.1)
    np.testing.assert_array_equal(g[1], g0[1])
    np.testing.assert_array_equal(g[3:], g0[3:])
    assert np.all(g[0] < g0[0]) and np.all(g[2] > g0[2])
    g, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 3, g0, eps=0.0, p=2.0, q=2.0, lr=5.0,
                         glb=0.1, gub=0.85)
    assert g.min() >= 0.1 and g.max() <= 0.85


def test_bclp_closed_form_p2_q2():
    check_bclp_closed_form_p2_q2()


def test_bclp_closed_form_p2_q1():
    check_bclp_closed_form_p2_q1()


def test_bclp_band_and_bounds():
    check_bclp_band_and_bounds()


def check_bclp_init_fbp_scale():
    """bclp_init: filtered forward projection times pi / (2 n_angles) is a
    filtered back projection when pushed through the (unscaled) back
    projector, so B g0 recovers map_inv(T) in the interior of a smooth disk."""
    import oracle.d1 as d1
    n = 48
    ang = np.arange(0.0, 180.0, 2.0)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    r = np.hypot(xx, yy)
    target = np.clip(1.2 - r / 14.0, 0.0, 1.0)[:, :, None] * 0.8
    ident = IdentityResponse()
    ident.map_inv = lambda m: np.asarray(m, dtype=np.float64)
    g0 = oo.bclp_init(target, lambda v: d1.forward(v, ang), ident, len(ang), glb=None)
    rec = d1.backward(g0, ang)
    inner = r < 10
    rel = np.linalg.norm((rec - target)[inner[:, :, None]]) / np.linalg.norm(target[inner[:, :, None]])
    assert rel < 0.05, rel


def test_bclp_init_fbp_scale():
    check_bclp_init_fbp_scale()


# --------------------------------------------------------------------------
# VER, circle indices, FBP, PM
# --------------------------------------------------------------------------
def check_ver_known_answer():
    """metrics.calcVER (metrics.py:47-87): #(void >= min gel) / (#gel + #void)."""
    gel = np.array([True, True, False, False, False, False])
    void = ~gel
    x = np.array([0.9, 0.6, 0.1, 0.6, 0.7, 0.59])
    # min gel 0.6; void voxels >= 0.6: 0.6 and 0.7 -> 2 of 6
    assert oo.ver(gel, void, x) == 2 / 6


def check_circle_inds():
    """Gel = target > 0 inside the inscribed circle, void = target == 0 inside;
    the circle of radius (n - 1) / 2 holds ~pi ((n - 1) / 2)^2 points and has the
    square's symmetries up to points that rounding puts on the circle."""
    n = 101
    t = np.zeros((n, n, 1))
    t[40:60, 45:70] = 1.0
    gel, void = oo.circle_inds(t)
    inside = gel | void
    assert not np.any(gel & void)
    assert abs(inside.sum() - np.pi * ((n - 1) / 2) ** 2) < 0.005 * n ** 2
    c = inside[:, :, 0]
    # the two axes use the same linspace: exactly symmetric under transposition;
    # mirror images differ only where rounding puts a point on the circle
    assert np.array_equal(c, c.T)
    assert np.sum(c != c[::-1]) <= 8 and np.sum(c != c[:, ::-1]) <= 8
    assert np.array_equal(gel, (t > 0) & inside) and np.array_equal(void, (t == 0) & inside)
    assert inside[0, n // 2, 0] and not inside[0, 0, 0]


def check_fbp_reconstructs():
    """FBP.py:42-63: the normalised filtered back projection of a disk is flat
    inside the disk and well below it outside."""
    import oracle.d1 as d1
    n = 40
    ang = np.arange(0.0, 180.0, 3.0)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    r = np.hypot(xx, yy)
    t = (r < 9).astype(float)[:, :, None]
    fwd, bwd = (lambda v: d1.forward(v, ang)), (lambda s: d1.backward(s, ang))
    b, x, _ = oo.fbp(t, fwd, bwd)
    assert b.min() >= 0 and b.max() == 1.0 and x.max() == 1.0
    # the clip of the filtered sinogram's negative lobes leaves a background
    # below half the disk's dose
    assert x[r < 6].min() > 0.9
    assert x[(r > 12) & (r < 18)].max() < 0.55
    # offset variant: every detector column is lifted by its own minimum
    bo, xo, _ = oo.fbp(t, fwd, bwd, offset=True)
    assert bo.max() == 1.0 and xo.max() == 1.0
    np.testing.assert_allclose(bo.min(axis=0), 0.0, atol=1e-15)


def check_pm_known_answer():
    """PM.py:76-140 by hand on a 3 x 3 slice (identity projectors, p = 1):
    dose = b pi / 180.  Inside the inscribed circle are the centre and the
    four edge midpoints.  The gel centre at dose 0.5 < d_h = 0.8 contributes
    0.3, a void midpoint at 0.9 > d_l = 0.7 contributes 0.2, the other void
    midpoints (dose 0) and the corners (outside) nothing."""
    t = np.zeros((3, 3, 1))
    t[1, 1, 0] = 1.0
    ident = lambda a: np.asarray(a, dtype=np.float64)
    dose = np.zeros((3, 3, 1))
    dose[1, 1, 0], dose[0, 1, 0], dose[2, 2, 0] = 0.5, 0.9, 5.0
    L, G, d = oo.pm_loss_grad(dose * 180 / np.pi, t, ident, ident, d_h=0.8, d_l=0.7, p=1.0)
    np.testing.assert_allclose(d, dose, rtol=1e-15)
    assert abs(L - 0.5) < 1e-12
    expect = np.zeros((3, 3, 1))
    expect[1, 1, 0], expect[0, 1, 0] = -1.0, 1.0
    np.testing.assert_allclose(G, expect * np.pi / 180, rtol=1e-15)


def test_ver_known_answer():
    check_ver_known_answer()


def test_circle_inds():
    check_circle_inds()


def test_fbp_reconstructs():
    check_fbp_reconstructs()


def test_pm_known_answer():
    check_pm_known_answer()


# --------------------------------------------------------------------------
# bit depth (util/data.py:495-523 and its call sites in the loops)
# --------------------------------------------------------------------------
def check_discretize_known_answer():
    """Two bits over [0, 3] give the levels 0, 1, 2, 3.  A value takes the
    largest level not above it; values on or above the top level take it; a
    value below the bottom level gets index -1, which NumPy indexing wraps to
    the top level."""
    x = np.array([-0.5, 0.0, 0.999, 1.0, 1.5, 2.9999, 3.0, 7.0])
    np.testing.assert_array_equal(oo.discretize(x, 2, 0.0, 3.0), [3.0, 0.0, 0.0, 1.0, 1.0, 2.0, 3.0, 3.0])
    # brute force over the 2**6 levels of a shifted range
    rng = np.random.default_rng(11)
    lo, hi = -0.3, 1.7
    levels = lo + (hi - lo) * np.arange(64) / 63
    x = np.concatenate([rng.uniform(lo, hi, 200), levels[::5]])
    out = oo.discretize(x, 6, lo, hi)
    for xi, oi in zip(x, out):
        below = levels[levels <= xi + 1e-15 * abs(xi)]
        assert abs(oi - below.max()) <= 1e-14, (xi, oi)
        assert 0 <= xi - oi < (hi - lo) / 63 + 1e-14


def _levels_of(v, top, bits):
    """True when every value of v is k * top / (2**bits - 1) for an integer k."""
    k = np.asarray(v) * (2 ** bits - 1) / top
    return np.allclose(k, np.round(k), rtol=0, atol=1e-9)


def check_osmo_bit_depth():
    """OSMO.py:61-64, 99-102, 164-165.  With one bit, each normalised
    projection inside the loop becomes two-valued: the entries equal to the
    maximum stay 1, every other entry falls to the projection's minimum.  The
    returned sinogram lies on the levels of [0, max].  Twenty bits reproduce the
    unquantised loop."""
    t, F, _ = _tiny_problem(4)
    fwd, bwd = _pair(F, t.shape)
    oo.osmo(t, fwd, bwd, 3, filt=None, bit_depth=1)
    for xm, b_in in zip(fwd.calls[1:], bwd.calls[1:]):
        raw = F @ xm.reshape(-1)
        raw = raw / raw.max()
        assert raw.min() > 0, "set-up: projections must stay positive"
        expect = np.where(raw == 1.0, 1.0, raw.min())
        np.testing.assert_array_equal(b_in, expect)
    fwd, bwd = _pair(F, t.shape)
    b3, _, _ = oo.osmo(t, fwd, bwd, 3, filt=None, bit_depth=3)
    assert len(np.unique(b3)) <= 8 and _levels_of(b3, b3.max(), 3)
    fwd, bwd = _pair(F, t.shape)
    bq, xq, _ = oo.osmo(t, fwd, bwd, 3, filt=None, bit_depth=20)
    fwd, bwd = _pair(F, t.shape)
    b0, x0, _ = oo.osmo(t, fwd, bwd, 3, filt=None)
    np.testing.assert_allclose(bq, b0, rtol=0, atol=2e-6)
    np.testing.assert_allclose(xq, x0, rtol=0, atol=2e-6)


def check_cal_fbp_bclp_bit_depth():
    """CAL.py:171-173, FBP.py:56-57, BCLP.py:528-533: the sinogram is
    quantised over [0, max] (levels k max / (2**bits - 1)) after each update
    and never moves up by quantisation."""
    t, F, _ = _tiny_problem(5)
    fwd, bwd = _pair(F, t.shape)
    b, _, _ = oo.cal(t, fwd, bwd, 3, filt=None, lr=0.01, bit_depth=3)
    # the returned sinogram is normalised after the last quantisation
    assert _levels_of(b, 1.0, 3)
    for b_in in bwd.calls[1:]:
        assert _levels_of(b_in, b_in.max(), 3) and len(np.unique(b_in)) <= 8
    fwd, bwd = _pair(F, t.shape)
    b2, _, _ = oo.fbp(t, fwd, bwd, filt=None, bit_depth=2)
    fwd, bwd = _pair(F, t.shape)
    b0, _, _ = oo.fbp(t, fwd, bwd, filt=None)
    assert _levels_of(b2, 1.0, 2)
    assert np.all(b2 <= b0 + 1e-15) and np.all(b0 - b2 < 1 / 3)
    T, g0 = np.full((4, 3, 2), 0.5), np.linspace(0.1, 0.9, 24).reshape(4, 3, 2)
    ident = lambda a: np.asarray(a, dtype=np.float64)
    g, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 2, g0, eps=0.0, p=2.0, q=2.0, lr=0.1, bit_depth=2)
    assert _levels_of(g, g.max(), 2) and len(np.unique(g)) <= 4
    gq, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 2, g0, eps=0.0, p=2.0, q=2.0, lr=0.1, bit_depth=20)
    gn, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 2, g0, eps=0.0, p=2.0, q=2.0, lr=0.1)
    np.testing.assert_allclose(gq, gn, rtol=0, atol=2e-6)


def test_discretize_known_answer():
    check_discretize_known_answer()


def test_osmo_bit_depth():
    check_osmo_bit_depth()


def test_cal_fbp_bclp_bit_depth():
    check_cal_fbp_bclp_bit_depth()


# --------------------------------------------------------------------------
# zero-dose rays (the reference's back projection zeroes them in its input)
# --------------------------------------------------------------------------
def check_zero_dose_rays():
    """Every sinogram the loops back project is zero on the masked rays, and
    so is the sinogram they return; an empty mask changes nothing."""
    t, F, rng = _tiny_problem(8)
    mask = rng.uniform(size=F.shape[0]) < 0.3
    none = np.zeros_like(mask)
    ident = lambda a: np.asarray(a, dtype=np.float64)   # noqa: E731

    def runs(fn):
        fwd, bwd = _pair(F, t.shape)
        out = fn(fwd, bwd, mask)
        for b_in in bwd.calls:
            assert np.all(b_in[mask] == 0), "a back-projected sinogram is not zero on the zero-dose rays"
        return out, fn(*_pair(F, t.shape), none), fn(*_pair(F, t.shape), None)

    for name, fn in (("osmo", lambda f, b, m: oo.osmo(t, f, b, 3, filt=None, zero_sino=m)),
                     # d_h near 1: the thresholded dose is ~0, the gradient F(-T) <= 0 and
                     # the update raises the zeroed rays again before each back projection
                     ("cal", lambda f, b, m: oo.cal(t, f, b, 3, d_h=0.99, filt=None, lr=0.01, zero_sino=m)),
                     ("fbp", lambda f, b, m: oo.fbp(t, f, b, filt=None, zero_sino=m))):
        out, empty, plain = runs(fn)
        assert np.all(out[0][mask] == 0), name
        np.testing.assert_array_equal(empty[0], plain[0])
        np.testing.assert_array_equal(empty[1], plain[1])
    T = rng.uniform(0.2, 0.8, F.shape[0])
    g0 = T + 0.2
    calls = []

    def bwd(a):
        calls.append(np.array(a))
        return ident(a)
    g, _, _, _ = oo.bclp(T, ident, bwd, IdentityResponse(), 3, g0, eps=0.0, p=2.0, q=2.0, lr=0.1, zero_sino=mask)
    assert np.all(g[mask] == 0) and all(np.all(c[mask] == 0) for c in calls)
    g_e, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 3, g0, eps=0.0, p=2.0, q=2.0, lr=0.1, zero_sino=none)
    g_n, _, _, _ = oo.bclp(T, ident, ident, IdentityResponse(), 3, g0, eps=0.0, p=2.0, q=2.0, lr=0.1)
    np.testing.assert_array_equal(g_e, g_n)


def test_zero_dose_rays():
    check_zero_dose_rays()
