"""Float64 restatement of the reference's optimisation loops and helpers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``fourier_filter`` / ``filter_sinogram``  vamtoolbox/util/data.py:302-438
* ``filter_target_osmo``                    vamtoolbox/util/data.py:43-113
* ``discretize``                            vamtoolbox/util/data.py:495-523
* ``zero_rays``                             the in-place zeroing of the back projections
                                            (Projector3DParallelCUDA.py:57-58)
* ``Response``                              vamtoolbox/response.py:12-291
* ``osmo``                                  vamtoolbox/optimizer/OSMO.py:9-178
* ``cal``                                   vamtoolbox/optimizer/CAL.py:10-218
* ``cal_threshold``                         vamtoolbox/optimizer/CAL.py:39-110 (d_h None)
* ``bclp``                                  vamtoolbox/optimizer/BCLP.py:138-641
* ``ver``                                   vamtoolbox/metrics.py:47-87
* ``fbp``                                   vamtoolbox/optimizer/FBP.py:9-76
* ``pm_loss_grad`` / ``pm``                 vamtoolbox/optimizer/PM.py:25-229

The projector is passed in as a pair of callables (forward, backward) so the
loops run on the D1 oracle or on an explicit matrix.
"""

from __future__ import annotations

import numpy as np
from scipy.fft import fft, fft2, fftshift, ifft, ifft2, ifftshift


def fourier_filter(size, name="ram-lak"):
    n = np.concatenate((np.arange(1, size / 2 + 1, 2, dtype=int), np.arange(size / 2 - 1, 0, -2, dtype=int)))
    f = np.zeros(size)
    f[0] = 0.25
    f[1::2] = -1 / (np.pi * n) ** 2
    ff = 2 * np.real(fft(f))
    if name == "ram-lak":
        pass
    elif name == "hamming":
        ff *= fftshift(np.hamming(size))
    elif name == "hanning":
        ff *= fftshift(np.hanning(size))
    elif name == "shepp-logan":
        omega = np.pi * np.fft.fftfreq(size)[1:]
        ff[1:] *= np.sin(omega) / omega
    elif name == "cosine":
        ff *= fftshift(np.sin(np.linspace(0, np.pi, size, endpoint=False)))
    else:
        raise ValueError(name)
    return ff


def filter_sinogram(s, name="ram-lak"):
    if name is None:
        return s
    s = np.asarray(s, dtype=np.float64)
    pX = s.shape[0]
    P = max(64, int(2 ** np.ceil(np.log2(2 * pX))))
    ff = fourier_filter(P, name)
    out = np.zeros_like(s)
    for z in range(s.shape[2]):
        img = np.zeros((P, s.shape[1]))
        img[:pX] = s[:, :, z]
        out[:, :, z] = np.real(ifft(fft(img, axis=0) * ff[:, None], axis=0)[:pX, :])
    return out


def filter_target_osmo(x, name="ram-lak"):
    if name is None:
        return x
    ny, nx = x.shape[0], x.shape[1]
    fy, fx = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    f = np.sqrt(fx ** 2 + fy ** 2)
    if name == "hamming":
        w = np.abs(np.hamming(nx))
        f = f * np.abs(np.sqrt(np.outer(w, w)))
    elif name != "ram-lak":
        raise ValueError(name)
    X = fftshift(fft2(x, axes=(0, 1)), axes=(0, 1))
    return np.real(ifft2(ifftshift(X * f[:, :, None], axes=(0, 1)), axes=(0, 1)))


class Response:
    """Generalised logistic response, analytical or as a 512-point table."""

    def __init__(self, kind="interpolation", A=0.0, K=1.0, B=25.0, M=0.5, nu=1.0, n_pts=512):
        self.kind = kind
        self.A, self.K, self.B, self.M, self.nu = A, K, B, M, nu
        if kind == "interpolation":
            self.f0 = np.linspace(0, 1, n_pts)
            self.m0 = self._map(self.f0)
            self.d0 = self._dmap(self.f0)

    def _map(self, f):
        return self.A + (self.K - self.A) / (1 + np.exp(-self.B * (f - self.M))) ** (1 / self.nu)

    def _dmap(self, f):
        c1 = (1 / (self.K - self.A)) ** self.nu
        c2 = self.B / self.nu
        c3 = (self._map(f) - self.A) ** (self.nu + 1)
        return c1 * c2 * c3 * np.exp(-self.B * (f - self.M))

    def map(self, f):
        if self.kind == "analytical":
            return self._map(f)
        return np.interp(f, self.f0, self.m0, left=self.m0[0], right=self.m0[-1])

    def dmap(self, f):
        if self.kind == "analytical":
            return self._dmap(f)
        return np.interp(f, self.f0, self.d0, left=self.d0[0], right=self.d0[-1])

    def map_inv(self, m):
        if self.kind == "analytical":
            with np.errstate(divide="ignore", invalid="ignore"):
                return -np.log(((self.K - self.A) / (m - self.A)) ** self.nu - 1) / self.B + self.M
        return np.interp(m, self.m0, self.f0, left=self.f0[0], right=self.f0[-1])


def discretize(x, bit_depth, lo, hi):
    """util/data.py:495-523: 2**bit_depth equally spaced levels from lo to hi
    (both included); each value becomes the largest level <= it, values at or
    above hi become hi, and values below lo take index -1, which NumPy's
    indexing wraps to the last level (hi)."""
    levels = np.linspace(lo, hi, 2 ** bit_depth, endpoint=True)
    k = np.digitize(x, bins=levels) - 1
    return levels[k]


def zero_rays(b, zero_sino):
    """The sinogram as the reference's back projection leaves it: the rays
    through the zero-dose body set to 0, in place in the reference
    (Projector3DParallelCUDA.py:57-58, Projector3DParallel.py:43-44); a new
    array here.  The loops below zero their sinogram before each back
    projection, so the back projection and the rest of the loop see it zeroed."""
    return b if zero_sino is None else np.where(zero_sino, 0.0, b)


def circle_inds(target):
    n0, n1, n2 = target.shape
    cy, cx, _ = np.meshgrid(np.linspace(-1, 1, n0), np.linspace(-1, 1, n1), np.linspace(-1, 1, n2))
    inside = cx ** 2 + cy ** 2 <= 1
    return np.logical_and(target > 0, inside), np.logical_and(target == 0, inside)


def ver(gel, void, x):
    return np.sum(x[void] >= np.min(x[gel])) / (np.sum(gel) + np.sum(void))


def osmo(target, fwd, bwd, n_iter, d_h=0.8, d_l=0.7, filt="ram-lak", inhibition=0.0, bit_depth=None,
         zero_sino=None):
    """OSMO.py:111-178 (no exit_param); ``bit_depth`` discretises each
    normalised projection over [min, max] (OSMO.py:61-64, 99-102) and the
    returned sinogram over [0, max] (OSMO.py:164-165)."""
    gel, void = circle_inds(target)
    err = np.full(n_iter, np.nan)
    xm = np.real(filter_target_osmo(target, filt))
    b = zero_rays(np.clip(fwd(xm), 0, None), zero_sino)
    x = bwd(b)
    x = x / np.amax(x)
    err[0] = ver(gel, void, x)

    def step(xm):
        b = fwd(xm)
        b = b / np.amax(b)
        if bit_depth is not None:
            b = discretize(b, bit_depth, np.min(b), np.max(b))
        b[b < inhibition] = inhibition
        b = zero_rays(b, zero_sino)
        xu = bwd(b)
        xu = xu / np.amax(xu)
        xu[xu < 0] = 0
        return xu, b

    for it in range(1, n_iter):
        vd = x[void] - d_l
        vd[vd < 0] = 0
        xm = xm.copy()
        xm[void] = xm[void] - vd
        x, _ = step(xm)
        gd = d_h - x[gel]
        gd[gd < 0] = 0
        xm[gel] = xm[gel] + gd
        x, b = step(xm)
        err[it] = ver(gel, void, x)
    if bit_depth is not None:
        b = discretize(b, bit_depth, 0.0, np.max(b))
    return b, x, err


def cal_threshold(x, gel, void, sigmoid):
    """CAL.py:39-110 (d_h None): the threshold of np.linspace(min, max, 100)
    maximising (gel voxels above it) / #gel - (void voxels above it) / #void,
    and the sigmoid 1 / (1 + exp(-(sigmoid * (x - threshold))))."""
    ths = np.linspace(np.amin(x), np.amax(x), 100)
    score = np.zeros(len(ths))
    for i, t in enumerate(ths):
        above = np.where(x > t, 1, 0)
        score[i] = np.sum(above[gel]) / np.count_nonzero(gel) - np.sum(above[void]) / np.count_nonzero(void)
    th = ths[np.argmax(score)]
    return 1 / (1 + np.exp(-(float(sigmoid) * (x - th)))), th


def cal(target, fwd, bwd, n_iter, d_h=0.8, filt="ram-lak", lr=0.01, momentum=0.0, sigmoid=0.01, bit_depth=None,
        zero_sino=None):
    """CAL.py:116-218: sigmoid threshold at d_h, or (d_h None) the score-maximising
    threshold of cal_threshold; ``bit_depth`` discretises each updated
    sinogram over [0, max] (CAL.py:171-173)."""
    gel, void = circle_inds(target)
    err = np.full(n_iter, np.nan)
    b = zero_rays(np.clip(filter_sinogram(fwd(target), filt), 0, None), zero_sino)
    db_prev = np.zeros_like(b)
    x = bwd(b)
    x = x / np.max(x)
    err[0] = ver(gel, void, x)
    b_opt = b
    for it in range(1, n_iter):
        if d_h is None:
            xt, _ = cal_threshold(x, gel, void, sigmoid)
        else:
            xt = 1 / (1 + np.exp(-(x - d_h) * (1 / sigmoid)))
        db = fwd(xt - target)
        grad = db * (1 - momentum) + db_prev * momentum / (1 - momentum ** (it + 1))
        b = np.clip(b - lr * grad, 0, None)
        if bit_depth is not None:
            b = discretize(b, bit_depth, 0.0, np.max(b))
        b = zero_rays(b, zero_sino)
        x = bwd(b)
        x = x / np.amax(x)
        db_prev = db
        b_opt = b
        err[it] = ver(gel, void, x)
    b_opt = np.clip(b_opt, 0, None)
    b_opt = b_opt / np.amax(b_opt)          # already zero on the zero-dose rays
    x_opt = bwd(b_opt)
    return b_opt, x_opt / np.amax(x_opt), err


def bclp_state(g, target, bwd, resp, eps, weight, p, q, dvol=1.0):
    """BCLPNorm.updateVariables / computeLoss / operand (BCLP.py:277-367)."""
    dose = bwd(g)
    m = resp.map(dose)
    eT = m - target
    eb = np.abs(eT) - eps
    v = eb > 0
    loss = (np.sum(v * weight * np.abs(eb) ** p) * dvol) ** (q / p)
    operand = v * weight * (np.abs(eb) ** (p - 1)) * np.sign(eT) * resp.dmap(dose)
    return dose, m, loss, operand


def bclp(target, fwd, bwd, resp, n_iter, g0, eps=0.1, weight=1.0, p=2.0, q=1.0, lr=0.01, glb=0.0, gub=None,
         grad_op=None, bit_depth=None, zero_sino=None):
    """BCLPNorm.gradientDescent (BCLP.py:477-526) from a given initial sinogram.

    ``grad_op`` (default ``fwd``) maps the operand to sinogram space.
    """
    grad_op = fwd if grad_op is None else grad_op

    def constrain(g):                       # imposeSinogramConstraints, BCLP.py:528-533
        if bit_depth is not None:
            g = discretize(g, bit_depth, 0.0, np.amax(g))
        return np.clip(g, glb, gub)

    g = constrain(np.asarray(g0, dtype=np.float64))
    losses = np.full(n_iter + 1, np.nan)
    for k in range(n_iter):
        g = zero_rays(g, zero_sino)
        dose, m, loss, operand = bclp_state(g, target, bwd, resp, eps, weight, p, q)
        losses[k] = loss
        G = grad_op(operand)
        if loss != 0.0:
            G = (q * loss ** ((q - p) / q)) * G
        g = constrain(g - lr * G)
    g = zero_rays(g, zero_sino)
    dose, m, loss, _ = bclp_state(g, target, bwd, resp, eps, weight, p, q)
    losses[n_iter] = loss
    return g, dose, m, losses


def bclp_init(target, fwd, resp, n_angles, filt="ram-lak", glb=0.0, gub=None):
    """BCLP.py:258-271 (projector without inverseBackward)."""
    g0 = fwd(resp.map_inv(target))
    g0 = filter_sinogram(g0, filt) * (np.pi / (2 * n_angles))
    return np.clip(g0, glb, gub)


def fbp(target, fwd, bwd, filt="ram-lak", offset=False, bit_depth=None, zero_sino=None):
    """FBP.py:42-63: filtered forward projection, non-negative, normalised
    (``bit_depth``: discretised over [0, max], FBP.py:56-57); dose normalised."""
    gel, void = circle_inds(target)
    b = filter_sinogram(fwd(target), filt)
    if offset:
        b = b + np.broadcast_to(np.abs(np.min(b, axis=0)), b.shape)
    else:
        b = np.clip(b, 0, None)
    b = b / np.max(b)
    if bit_depth is not None:
        b = discretize(b, bit_depth, 0.0, np.max(b))
    b = zero_rays(b, zero_sino)
    x = bwd(b)
    x = x / np.max(x)
    return b, x, ver(gel, void, x)


def pm_loss_grad(b, target, bwd, fwd, d_h=0.8, d_l=0.7, rho_1=1.0, rho_2=1.0, p=1.0):
    """Loss and gradient of PM.py:76-140 at sinogram b (gradient through the
    forward projector, both scaled by pi/180)."""
    gel, void = circle_inds(target)
    core = np.zeros(target.shape, dtype=bool)
    core[gel] = True
    background = np.zeros(target.shape, dtype=bool)
    background[void] = True
    dose = bwd(b) * np.pi / 180
    V1 = core & ~(dose >= d_h)
    V2 = background & ~(dose <= d_l)
    loss = float(np.sum(rho_1 * np.power(dose - d_l, p) * V2 + rho_2 * np.power(d_h - dose, p) * V1))
    gd = rho_1 * V2 * p * np.power(dose - d_l, p - 1) - rho_2 * V1 * p * np.power(d_h - dose, p - 1)
    return loss, fwd(gd) * (np.pi / 180), dose


def pm(target, fwd, bwd, n_iter, filt="ram-lak", **kw):
    """PM.py:168-229: L-BFGS-B from the filtered initial sinogram, bounds [0, 1]."""
    from scipy import optimize
    b0 = np.clip(filter_sinogram(fwd(target), filt), 0, None)
    b0 = b0 / np.max(b0)
    shape = b0.shape

    def f(bf):
        return pm_loss_grad(bf.reshape(shape), target, bwd, fwd, **kw)[0]

    def g(bf):
        return pm_loss_grad(bf.reshape(shape), target, bwd, fwd, **kw)[1].reshape(-1)

    res = optimize.minimize(fun=f, x0=b0.reshape(-1), jac=g, method="L-BFGS-B",
                            bounds=np.tile(np.array([[0.0, 1.0]]), (b0.size, 1)),
                            options={"maxiter": n_iter, "ftol": 1e-12, "disp": False, "maxcor": 10, "maxls": 30})
    b = res.x.reshape(shape)
    return b, bwd(b) * np.pi / 180
