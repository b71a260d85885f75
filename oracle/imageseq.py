"""NumPy restatement of the reference's image-sequence export
(vamtoolbox/imagesequence.py:14-82 ImageConfig, 85-182 ImageSeq.__init__,
300-399 the helpers).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``place``          _insertImage (300-333) and _arrayInsertImage (336-361):
                     the (v1, v2, u1, u2) window of each arrayed copy, in the
                     reference's order (a later copy overwrites an earlier one)
* ``rotate``         _rotate (397-399): ``scipy.ndimage.rotate(axes=(0, 2),
                     reshape=True, order=1)``, the routine the reference calls
* ``resize_linear``  _scaleSize (369-380): OpenCV's ``resize`` with
                     INTER_LINEAR restated (half-pixel centres, source index
                     clamped at both edges); OpenCV is absent here, so this
                     function is pinned by closed forms only
* ``percentile``     ``np.percentile`` (linear method), the routine the
                     reference calls
* ``frames``         the whole pipeline: flips, rotation, resize, percentile
                     normalisation, intensity scale, truncation, cast to the
                     bit depth's unsigned type, insertion into uint8 frames
"""

from __future__ import annotations

import numpy as np


def place(n_u, n_v, s_u, s_v, u_offset=0, v_offset=0, array_num=1, array_offset=0):
    """Windows (v1, v2, u1, u2) of the copies in insertion order.

    One copy: ImageSeq passes an extra shift of 0 and _insertImage subtracts
    image_config.v_offset once, so the window is centred at N/2 - offset.
    Arrayed copies: the extra shift of copy k is int(array_offset * d / 2 +
    v_offset) (even array_num, d = 1, -1, 2, -2, ...) or
    int(array_offset * d + v_offset) (odd array_num, d = 0, 1, -1, 2, ...),
    so v_offset counts twice there, as in the reference.
    """
    def window(extra):
        u1 = int(n_u / 2 - u_offset - s_u / 2)
        u2 = int(n_u / 2 - u_offset + s_u / 2)
        v1 = int(n_v / 2 - v_offset - extra - s_v / 2)
        v2 = int(n_v / 2 - v_offset - extra + s_v / 2)
        if u1 < 0 or u2 > n_u:
            raise Exception("u window out of the screen")
        if v1 < 0 or v2 > n_v:
            raise Exception("v window out of the screen")
        return v1, v2, u1, u2

    if array_num == 1:
        return [window(0)]
    out = []
    for k in range(array_num):
        if array_num % 2 == 0:
            d = int(np.ceil((k + 1) / 2) * (-1) ** k)
            extra = int(array_offset * d / 2 + v_offset)
        else:
            d = int(0 + np.ceil(k / 2) * (-1) ** (k + 1))
            extra = int(array_offset * d + v_offset)
        out.append(window(extra))
    return out


def rotate(s, angle_deg):
    from scipy import ndimage
    return ndimage.rotate(s, angle_deg, axes=(0, 2), reshape=True, order=1)


def resize_linear(img, new_w, new_h):
    """cv2.resize(img, (new_w, new_h), interpolation=INTER_LINEAR) on a 2-D
    float array: destination pixel d samples source position
    (d + 0.5) * (src / dst) - 0.5; the integer part is clamped to
    [0, src - 1] with a zero fraction at either edge."""
    h, w = img.shape

    def taps(n_src, n_dst):
        f = (np.arange(n_dst) + 0.5) * (n_src / n_dst) - 0.5
        i0 = np.floor(f).astype(np.int64)
        t = f - i0
        t = np.where(i0 < 0, 0.0, t)
        i0 = np.where(i0 < 0, 0, i0)
        t = np.where(i0 >= n_src - 1, 0.0, t)
        i0 = np.where(i0 >= n_src - 1, n_src - 1, i0)
        i1 = np.minimum(i0 + 1, n_src - 1)
        return i0, i1, t

    y0, y1, ty = taps(h, new_h)
    x0, x1, tx = taps(w, new_w)
    top = img[y0][:, x0] * (1 - tx) + img[y0][:, x1] * tx
    bot = img[y1][:, x0] * (1 - tx) + img[y1][:, x1] * tx
    return top * (1 - ty)[:, None] + bot * ty[:, None]


def percentile(a, p):
    return np.percentile(a, p)


def to_unsigned(x, bit_depth):
    """astype of the truncated values to the bit depth's unsigned type
    (imagesequence.py:132-143): truncation toward zero.  The reference's cast
    of a negative or NaN value is undefined in C; here (and in the product)
    such values become 0."""
    x = np.where(np.isnan(x) | (x < 0), 0.0, x)
    if bit_depth <= 8:
        return x.astype(np.uint8)
    if bit_depth <= 16:
        return x.astype(np.uint16)
    return x.astype(np.uint32)


def frames(sino, n_u, n_v, rotate_angle=0.0, u_offset=0, v_offset=0, invert_u=False, invert_v=False,
           array_num=1, array_offset=0, normalization_percentile=None, bit_depth=8, intensity_scale=1.0,
           size_scale=1.0):
    """ImageSeq.__init__ (imagesequence.py:85-182): one uint8 frame per angle."""
    m = np.array(sino, dtype=np.float64, copy=True)
    if invert_u:
        m = np.flip(m, axis=0)
    if invert_v:
        m = np.flip(m, axis=2)
    if rotate_angle != 0.0:
        m = rotate(m, rotate_angle)
    if size_scale != 1.0:
        nh, nw = int(m.shape[0] * size_scale), int(m.shape[2] * size_scale)
        r = np.zeros((nh, m.shape[1], nw))
        for i in range(m.shape[1]):
            r[:, i, :] = resize_linear(m[:, i, :], nw, nh)
        m = r
    maxv = 2 ** bit_depth - 1
    if normalization_percentile is not None:
        m = m / percentile(m, normalization_percentile) * maxv
    if intensity_scale != 1.0:
        m = m * intensity_scale
    m = to_unsigned(np.minimum(m, maxv), bit_depth)
    wins = place(n_u, n_v, m.shape[0], m.shape[2], u_offset, v_offset, array_num, array_offset)
    out = np.zeros((m.shape[1], n_v, n_u), dtype=np.uint8)
    for j in range(m.shape[1]):
        img = m[:, j, :].T
        for v1, v2, u1, u2 in wins:
            out[j, v1:v2, u1:u2] = img      # uint16 / uint32 wrap modulo 256, as numpy's assignment
    return out
