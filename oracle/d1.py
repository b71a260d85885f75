"""Float64 restatement of the reference's CPU parallel-beam projector ("D1").

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference: ``vamtoolbox/projector/Projector3DParallel.py`` class
``Projector3DParallelPython``:

* forward  (lines 113-156): for every z-slice and angle, rotate the slice with
  ``skimage.transform.warp(slice, R, clip=True)`` (bilinear, zero padding,
  output clipped to the slice's value range) using the inverse-map matrix R of
  lines 137-143, then sum the rotated image over its rows (line 155).
* backward (lines 158-207): voxel-driven; for every angle the detector
  coordinate ``t = x cos - y sin`` of each voxel (lines 184-185) is
  interpolated in the projection with ``np.interp(..., left=0, right=0)``
  (lines 187-190), summed over angles (line 201) and finally clipped to the
  inscribed circle with ``util.data.clipToCircle`` (line 207;
  ``vamtoolbox/util/data.py:14-40``).

Build-defined extensions (absent from the reference's Python projector; they
reduce exactly to the reference when their parameter is zero):

* ``det_offset`` (delta): detector bin k sits at lateral coordinate
  ``t_k = k - c - delta``; an integer delta is an exact index shift.
* ``mu`` (per-voxel absorption inside a cylinder of radius R): a sample at
  lateral coordinate t and along-ray coordinate s is weighted by
  ``exp(-mu * max(0, sqrt(R^2 - t^2) - s) / cos(phi))`` when |t| < R and by 0
  otherwise -- the per-ray form of ``astra3Dabsorption.getExpabs``
  (``vamtoolbox/projector/astra3Dabsorption.py:129-145``).  mu = 0 means no
  absorption at all (weight exactly 1, no cylinder).
* ``incl_deg`` (phi): rays rise by ``tan(phi)`` slices per in-plane unit step;
  the in-plane sample positions are those of D1, the z position of the sample
  of detector row j is ``j + (r - c) tan(phi)`` (linear, zero padded).  At
  phi = 0 the z weight is exactly 1 on slice j.

All geometry (sample coordinates, floor decisions, weights) is computed in
float64 with the same operation order as the reference's routines, so that
every discontinuous decision (which taps, which detector bin, inside/outside)
is reproduced exactly; the HIP kernels do the same.
"""

from __future__ import annotations

import numpy as np


# --------------------------------------------------------------------------
# circle clip: util/data.py:14-40 (clipToCircle), geometry.py:682-711 (getCircleMask)
# --------------------------------------------------------------------------
def outside_circle_mask(ny: int, nx: int) -> np.ndarray:
    """Boolean (ny, nx) mask of points *outside* the inscribed circle.

    util/data.py:29-33: ``circle_y, circle_x = np.meshgrid(linspace(-1,1,shape0),
    linspace(-1,1,shape1))`` ('xy' indexing, so the arrays have shape
    (shape1, shape0)) and ``x[circle_x**2 + circle_y**2 > 1] = 0``.
    """
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx))
    return circle_x ** 2 + circle_y ** 2 > 1


def clip_to_circle(x: np.ndarray) -> np.ndarray:
    """Return a copy of ``x`` with everything outside the inscribed circle zeroed.

    util/data.py:14-40.  The reference mutates its argument in place; the
    oracle returns a copy.
    """
    x = np.array(x, dtype=np.float64, copy=True)
    mask = outside_circle_mask(x.shape[0], x.shape[1])
    if x.ndim == 2:
        x[mask] = 0
    else:
        x[np.broadcast_to(mask[..., None], x.shape)] = 0
    return x


def inside_circle_mask(ny: int, nx: int) -> np.ndarray:
    """getCircleMask, geometry.py:682-711 (``R <= 1``)."""
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx))
    return circle_x ** 2 + circle_y ** 2 <= 1


# --------------------------------------------------------------------------
# warp matrix and bilinear sampling: Projector3DParallel.py:135-145
# --------------------------------------------------------------------------
def warp_matrix(angle_deg: float, center: int) -> np.ndarray:
    """Inverse-map matrix R of Projector3DParallel.py:135-143."""
    angle = np.deg2rad(angle_deg)
    cos_a, sin_a = np.cos(angle), np.sin(angle)
    return np.array(
        [
            [cos_a, sin_a, -center * (cos_a + sin_a - 1)],
            [-sin_a, cos_a, -center * (cos_a - sin_a - 1)],
            [0, 0, 1],
        ]
    )


def sample_coords(M: np.ndarray, kcol: np.ndarray, rrow: np.ndarray):
    """Input (row, col) for output (row rrow, col kcol) under inverse map M.

    skimage's affine transform: ``c = M00*x + M01*y + M02``,
    ``r = M10*x + M11*y + M12`` with x = output column, y = output row,
    evaluated left to right without fused multiply-add.
    """
    col = M[0, 0] * kcol + M[0, 1] * rrow + M[0, 2]
    row = M[1, 0] * kcol + M[1, 1] * rrow + M[1, 2]
    return row, col


def bilinear_zero_pad(img: np.ndarray, row: np.ndarray, col: np.ndarray) -> np.ndarray:
    """Bilinear interpolation with constant-0 padding, skimage ``order=1``.

    ``img`` is (ny, nx, nz); ``row``/``col`` any common shape S; returns S+(nz,).
    Neighbours are floor/ceil per axis; out-of-range neighbours read 0;
    ``top = (1-dc)*tl + dc*tr``, ``bottom = (1-dc)*bl + dc*br``,
    ``out = (1-dr)*top + dr*bottom``.
    """
    ny, nx = img.shape[0], img.shape[1]
    minr = np.floor(row)
    minc = np.floor(col)
    maxr = np.ceil(row)
    maxc = np.ceil(col)
    dr = (row - minr)[..., None]
    dc = (col - minc)[..., None]

    def pix(rr, cc):
        rr = rr.astype(np.int64)
        cc = cc.astype(np.int64)
        ok = (rr >= 0) & (rr < ny) & (cc >= 0) & (cc < nx)
        v = img[np.clip(rr, 0, ny - 1), np.clip(cc, 0, nx - 1)]
        return np.where(ok[..., None], v, 0.0)

    top = (1 - dc) * pix(minr, minc) + dc * pix(minr, maxc)
    bottom = (1 - dc) * pix(maxr, minc) + dc * pix(maxr, maxc)
    return (1 - dr) * top + dr * bottom


def clip_warp_output(slice_min: np.ndarray, slice_max: np.ndarray, out: np.ndarray) -> np.ndarray:
    """skimage ``warp(..., clip=True)`` output clipping, per z-slice.

    The rotated image is clipped to the input's [min, max]; when the constant
    value 0 lies outside that range, output pixels exactly equal to 0 are kept
    at 0.  ``out`` has z as its last axis.
    """
    out = np.array(out, copy=True)
    for z in range(out.shape[-1]):
        lo, hi = slice_min[z], slice_max[z]
        o = out[..., z]
        preserve = not (lo <= 0.0 <= hi)
        if preserve:
            zero = o == 0.0
        o = np.clip(o, lo, hi)
        if preserve:
            o[zero] = 0.0
        out[..., z] = o
    return out


# --------------------------------------------------------------------------
# extension weights
# --------------------------------------------------------------------------
def default_radius(n: int) -> float:
    return (n - 1) / 2.0


def absorption_weight(t, s, mu: float, radius: float, incl_deg: float = 0.0):
    """exp(-mu * max(0, sqrt(R^2-t^2) - s) / cos(phi)) for |t| < R, else 0; 1 if mu == 0."""
    t = np.asarray(t, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    if mu == 0.0:
        return np.ones(np.broadcast(t, s).shape)
    inside = np.abs(t) < radius
    half = np.sqrt(np.where(inside, radius * radius - t * t, 0.0))
    path = np.maximum(0.0, half - s) / np.cos(np.deg2rad(incl_deg))
    return np.where(inside, np.exp(-mu * path), 0.0)


def _z_taps(zpos: np.ndarray, nz: int):
    """Linear interpolation taps in z with zero padding: [(index, weight, valid), ...]."""
    z0 = np.floor(zpos)
    w1 = zpos - z0
    z0 = z0.astype(np.int64)
    taps = []
    for zi, w in ((z0, 1.0 - w1), (z0 + 1, w1)):
        ok = (zi >= 0) & (zi < nz)
        taps.append((np.clip(zi, 0, nz - 1), np.where(ok, w, 0.0)))
    return taps


# --------------------------------------------------------------------------
# D1 forward: Projector3DParallel.py:113-156
# --------------------------------------------------------------------------
def forward(vol: np.ndarray, angles_deg, det_offset: float = 0.0, mu: float = 0.0,
            radius: float | None = None, incl_deg: float = 0.0,
            clip_output: bool = True, occ: np.ndarray | None = None) -> np.ndarray:
    """Sinogram (N, n_angles, nz) of volume (N, N, nz).

    ``clip_output`` applies skimage's per-slice output clip (the reference
    always does, ``warp(..., clip=True)``); the adjoint is defined for the
    linear part, ``clip_output=False``.  ``occ`` is an occlusion sinogram
    (``occlusion_sinogram``): samples behind the first occluding row of their
    column are dropped (Projector3DParallel.py:147-153).
    """
    vol = np.asarray(vol, dtype=np.float64)
    if vol.ndim == 2:
        vol = vol[:, :, None]
    N, nx, nz = vol.shape
    assert N == nx, "D1 assumes a square slice (Projector3DParallel.py:64-67 uses nY only)"
    angles_deg = np.atleast_1d(np.asarray(angles_deg, dtype=np.float64))
    c = N // 2
    R = default_radius(N) if radius is None else radius
    kk = np.arange(N, dtype=np.float64) - det_offset      # output column (detector)
    rr = np.arange(N, dtype=np.float64)                   # output row (along ray)
    K, RR = np.meshgrid(kk, rr, indexing="xy")            # shape (r, k)
    t = (np.arange(N) - c) - det_offset
    s = np.arange(N) - c
    W = absorption_weight(t[None, :], s[:, None].astype(np.float64), mu, R, incl_deg)  # (r, k)
    smin = vol.min(axis=(0, 1))
    smax = vol.max(axis=(0, 1))
    tan_phi = np.tan(np.deg2rad(incl_deg)) if incl_deg != 0.0 else 0.0

    out = np.zeros((N, angles_deg.size, nz))
    for i, a in enumerate(angles_deg):
        M = warp_matrix(a, c)
        row, col = sample_coords(M, K, RR)
        if incl_deg == 0.0:
            rot = bilinear_zero_pad(vol, row, col)            # (r, k, nz)
            if clip_output:
                rot = clip_warp_output(smin, smax, rot)
        else:
            plane = bilinear_zero_pad(vol, row, col)          # (r, k, nz) per source slice
            rot = np.zeros_like(plane)
            j = np.arange(nz, dtype=np.float64)
            zpos = j[None, :] + s[:, None] * tan_phi          # (r, nz)
            for zi, w in _z_taps(zpos, nz):
                rot += w[:, None, :] * np.take_along_axis(
                    plane, np.broadcast_to(zi[:, None, :], plane.shape), axis=2)
        if occ is not None:
            if incl_deg != 0.0:
                raise NotImplementedError("occlusion with inclination")
            col = occ[:, i, :]                                # (k, nz)
            # the shadow applies only when the column holds a finite non-zero
            # first hit (count_nonzero - isnan != 0, Projector3DParallel.py:150)
            active = (np.count_nonzero(col, axis=0) - np.isnan(col).sum(axis=0)) != 0
            with np.errstate(invalid="ignore"):
                shadow = s[:, None, None] > col[None, :, :]   # y > occ (NaN: False)
            rot = rot * ~(shadow & active[None, None, :])
        out[:, i, :] = (W[..., None] * rot).sum(axis=0)
    return out


# --------------------------------------------------------------------------
# occlusion sinogram: Projector3DParallel.py:77-111
# --------------------------------------------------------------------------
def occlusion_sinogram(insert: np.ndarray, angles_deg, det_offset: float = 0.0) -> np.ndarray:
    """First occluding row coordinate y = r - N//2 per detector column.

    ``insert`` (N, N, nz) is 1 where the attenuation field is > 0.  Per slice
    and angle the insert is rotated like the forward projection (bilinear,
    zero padded, clipped); a sample occludes when its rotated value is > 0,
    i.e. when a tap of positive weight is an occluding voxel.  Returns
    (N, n_angles, nz) float64 with NaN where a column meets no occluder.
    """
    ins = (np.asarray(insert) > 0).astype(np.float64)
    if ins.ndim == 2:
        ins = ins[:, :, None]
    N, nx, nz = ins.shape
    angles_deg = np.atleast_1d(np.asarray(angles_deg, dtype=np.float64))
    c = N // 2
    kk = np.arange(N, dtype=np.float64) - det_offset
    rr = np.arange(N, dtype=np.float64)
    K, RR = np.meshgrid(kk, rr, indexing="xy")
    smin = ins.min(axis=(0, 1))
    smax = ins.max(axis=(0, 1))
    out = np.full((N, angles_deg.size, nz), np.nan)
    for i, a in enumerate(angles_deg):
        row, col = sample_coords(warp_matrix(a, c), K, RR)
        rot = clip_warp_output(smin, smax, bilinear_zero_pad(ins, row, col))   # (r, k, nz)
        hit = rot > 0
        first = np.argmax(hit, axis=0)                                          # (k, nz)
        out[:, i, :] = np.where(hit.any(axis=0), (first - c).astype(np.float64), np.nan)
    return out


# --------------------------------------------------------------------------
# D1 backward: Projector3DParallel.py:158-207
# --------------------------------------------------------------------------
def _interp_taps(t: np.ndarray, xp0: float, xp_last: float, n: int, c: int, det_offset: float):
    """np.interp(t, xp, fp, left=0, right=0) as two taps for xp = xp0 + arange(n).

    The nodes are ``xp[k] = (k - c) - det_offset``; ``xp0``/``xp_last`` are the
    first and last node values as numpy holds them;
    t is valid when ``xp0 <= t <= xp_last`` (numpy compares t with the nodes).
    Returns (j, w, valid) with value = fp[j] + w*(fp[j1]-fp[j]), j1 = min(j+1, n-1):
    numpy evaluates ``slope*(t - xp[j]) + fp[j]`` on the interval containing t
    and returns ``fp[-1]`` exactly at the last node (w = 0 there).
    """
    u = t - xp0
    valid = (t >= xp0) & (t <= xp_last)
    j = np.clip(np.floor(u), 0, n - 1).astype(np.int64)
    w = np.where(j == n - 1, 0.0, t - ((j - c) - det_offset))
    return j, w, valid


def backward(sino: np.ndarray, angles_deg, det_offset: float = 0.0, mu: float = 0.0,
             radius: float | None = None, incl_deg: float = 0.0,
             clip: bool = True, occ: np.ndarray | None = None) -> np.ndarray:
    """Reconstruction (N, N, nz) of sinogram (N, n_angles, nz).

    With an occlusion sinogram ``occ`` a voxel gets nothing from an angle when
    ``s > floor(np.interp(t, xp, occ column, left=nan, right=nan))``
    (Projector3DParallel.py:192-199, 209-216; NaN compares False).
    """
    sino = np.asarray(sino, dtype=np.float64)
    if sino.ndim == 2:
        sino = sino[:, :, None]
    N, na, nz = sino.shape
    angles_deg = np.atleast_1d(np.asarray(angles_deg, dtype=np.float64))
    assert angles_deg.size == na
    radius_c = N // 2
    y, x = np.mgrid[:N, :N] - radius_c                   # line 65
    c = N // 2
    xp0 = float(-c) - det_offset                          # proj_t = arange(N) - N//2 (line 67)
    xp_last = float(N - 1 - c) - det_offset
    R = default_radius(N) if radius is None else radius
    tan_phi = np.tan(np.deg2rad(incl_deg)) if incl_deg != 0.0 else 0.0
    out = np.zeros((N, N, nz))
    for i, a in enumerate(angles_deg):
        ang = np.deg2rad(a)
        cos_a, sin_a = np.cos(ang), np.sin(ang)
        t = x * cos_a - y * sin_a                         # line 184
        s = x * sin_a + y * cos_a                         # line 185
        j, w, valid = _interp_taps(t, xp0, xp_last, N, c, det_offset)
        b = sino[:, i, :]                                 # (N, nz)
        b0 = b[j]
        b1 = b[np.minimum(j + 1, N - 1)]
        val = b0 + w[..., None] * (b1 - b0)
        val = np.where(valid[..., None], val, 0.0)        # (N, N, nz) per detector row
        if incl_deg != 0.0:
            zz = np.arange(nz, dtype=np.float64)
            vpos = zz[None, None, :] - s[..., None] * tan_phi
            acc = np.zeros_like(val)
            for zi, wz in _z_taps(vpos, nz):
                acc += wz * np.take_along_axis(val, zi, axis=2)
            val = acc
        if mu != 0.0:
            val = val * absorption_weight(t, s, mu, R, incl_deg)[..., None]
        if occ is not None:
            if incl_deg != 0.0:
                raise NotImplementedError("occlusion with inclination")
            xp = (np.arange(N) - c) - det_offset
            for z in range(nz):
                oz = np.interp(t, xp, occ[:, i, z], left=np.nan, right=np.nan)
                with np.errstate(invalid="ignore"):
                    shadow = s > np.floor(oz)
                val[:, :, z] = val[:, :, z] * ~shadow
        out += val
    if clip:
        out = clip_to_circle(out)
    return out


# --------------------------------------------------------------------------
# exact adjoints (transposes of the linear parts)
# --------------------------------------------------------------------------
def forward_adjoint(sino: np.ndarray, angles_deg, det_offset: float = 0.0, mu: float = 0.0,
                    radius: float | None = None, incl_deg: float = 0.0) -> np.ndarray:
    """Transpose of ``forward(..., clip_output=False)``: scatter each sample's
    bilinear taps back with the same weights."""
    sino = np.asarray(sino, dtype=np.float64)
    if sino.ndim == 2:
        sino = sino[:, :, None]
    N, na, nz = sino.shape
    angles_deg = np.atleast_1d(np.asarray(angles_deg, dtype=np.float64))
    c = N // 2
    R = default_radius(N) if radius is None else radius
    kk = np.arange(N, dtype=np.float64) - det_offset
    rr = np.arange(N, dtype=np.float64)
    K, RR = np.meshgrid(kk, rr, indexing="xy")
    t = (np.arange(N) - c) - det_offset
    s = np.arange(N) - c
    W = absorption_weight(t[None, :], s[:, None].astype(np.float64), mu, R, incl_deg)
    tan_phi = np.tan(np.deg2rad(incl_deg)) if incl_deg != 0.0 else 0.0
    out = np.zeros((N, N, nz))
    for i, a in enumerate(angles_deg):
        M = warp_matrix(a, c)
        row, col = sample_coords(M, K, RR)                # (r, k)
        y = W[..., None] * sino[None, :, i, :]            # (r, k, nz): value per sample per det row
        if incl_deg != 0.0:
            j = np.arange(nz, dtype=np.float64)
            zpos = j[None, :] + s[:, None] * tan_phi      # (r, nz): source z of det row j
            y_src = np.zeros_like(y)
            rI = np.arange(N)[:, None]
            kI = np.arange(N)[None, :]
            for zi, wz in _z_taps(zpos, nz):
                for jj in range(nz):
                    np.add.at(y_src, (rI, kI, zi[:, jj][:, None]),
                              wz[:, jj][:, None] * y[:, :, jj])
            y = y_src
        minr = np.floor(row)
        minc = np.floor(col)
        maxr = np.ceil(row)
        maxc = np.ceil(col)
        dr = row - minr
        dc = col - minc
        for rsel, wr in ((minr, 1 - dr), (maxr, dr)):
            for csel, wc in ((minc, 1 - dc), (maxc, dc)):
                ri = rsel.astype(np.int64)
                ci = csel.astype(np.int64)
                ok = (ri >= 0) & (ri < N) & (ci >= 0) & (ci < N)
                w = np.where(ok, wr * wc, 0.0)
                np.add.at(out, (np.clip(ri, 0, N - 1), np.clip(ci, 0, N - 1)), w[..., None] * y)
    return out


def backward_adjoint(vol: np.ndarray, angles_deg, det_offset: float = 0.0, mu: float = 0.0,
                     radius: float | None = None, incl_deg: float = 0.0,
                     clip: bool = True) -> np.ndarray:
    """Transpose of ``backward``: each voxel scatters its value to the two
    interpolation nodes with the np.interp weights."""
    vol = np.asarray(vol, dtype=np.float64)
    if vol.ndim == 2:
        vol = vol[:, :, None]
    N, _, nz = vol.shape
    angles_deg = np.atleast_1d(np.asarray(angles_deg, dtype=np.float64))
    if clip:
        vol = clip_to_circle(vol)
    c = N // 2
    y, x = np.mgrid[:N, :N] - c
    xp0 = float(-c) - det_offset
    xp_last = float(N - 1 - c) - det_offset
    R = default_radius(N) if radius is None else radius
    tan_phi = np.tan(np.deg2rad(incl_deg)) if incl_deg != 0.0 else 0.0
    out = np.zeros((N, angles_deg.size, nz))
    for i, a in enumerate(angles_deg):
        ang = np.deg2rad(a)
        cos_a, sin_a = np.cos(ang), np.sin(ang)
        t = x * cos_a - y * sin_a
        s = x * sin_a + y * cos_a
        j, w, valid = _interp_taps(t, xp0, xp_last, N, c, det_offset)
        v = vol
        if mu != 0.0:
            v = v * absorption_weight(t, s, mu, R, incl_deg)[..., None]
        if incl_deg != 0.0:
            zz = np.arange(nz, dtype=np.float64)
            vpos = zz[None, None, :] - s[..., None] * tan_phi
            vd = np.zeros_like(v)                          # value per detector row
            for zi, wz in _z_taps(vpos, nz):
                flat_i = np.broadcast_to(np.arange(N)[:, None, None], v.shape)
                flat_k = np.broadcast_to(np.arange(N)[None, :, None], v.shape)
                np.add.at(vd, (flat_i, flat_k, zi), wz * v)
            v = vd
        v = np.where(valid[..., None], v, 0.0)
        col = np.zeros((N, nz))
        np.add.at(col, j.ravel(), ((1 - w)[..., None] * v).reshape(-1, nz))
        np.add.at(col, np.minimum(j + 1, N - 1).ravel(), (w[..., None] * v).reshape(-1, nz))
        out[:, i, :] = col
    return out
