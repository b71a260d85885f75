"""Image-sequence export: the frames a DLP projector shows, one per angle
(vamtoolbox/imagesequence.py).

``ImageConfig`` keeps the reference's options and their checks (14-82).
``ImageSeq`` builds the frames on the GPU (libvamx_hip.so): the flips, the
rotation (vamx_imgseq_rotate: scipy.ndimage.rotate(axes=(0, 2),
reshape=True, order=1), as the reference calls it), the size scaling
(vamx_imgseq_resize: OpenCV's INTER_LINEAR resize), and one fused kernel
(vamx_imgseq_compose) for the percentile normalisation, the intensity scale,
the truncation, the cast to the bit depth's unsigned type and the insertion of
the arrayed copies into uint8 frames.  The percentile is NumPy's linear one,
taken from two order statistics of the device array.

``frames`` is the (n_angles, N_v, N_u) uint8 device tensor; ``images`` the
reference's list of per-angle NumPy frames (copied to the host on first use).

Differences from the reference, all at its edges: a negative or NaN sample
becomes 0 (the reference's cast of such floats to an unsigned type is
undefined behaviour in C); ``save`` / ``loadImageSeq`` use a NumPy ``.npz``
container instead of a dill pickle; ``saveAsVideo`` and ``preview`` need
OpenCV, which this package does not depend on, and raise.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._device import require_gpu
from .geometry import Sinogram


class ImageConfig:
    """imagesequence.py:14-82."""

    def __init__(self, image_dims, **kwargs):
        self.N_u, self.N_v = image_dims
        if self.N_u % 4 != 0 or self.N_v % 4 != 0:
            raise Exception("Image dimensions should be divisible by 4.")
        self.rotate_angle = kwargs.get("rotate_angle", 0.0)
        self.u_offset = kwargs.get("u_offset", 0)
        self.v_offset = kwargs.get("v_offset", 0)
        self.invert_u = kwargs.get("invert_u", False)
        self.invert_v = kwargs.get("invert_v", False)
        self.array_num = kwargs.get("array_num", 1)
        self.array_offset = kwargs.get("array_offset", 0)
        self.normalization_percentile = kwargs.get("normalization_percentile", None)
        self.intensity_scale = kwargs.get("intensity_scale", 1.0)
        self.bit_depth = kwargs.get("bit_depth", 8)
        self.size_scale = kwargs.get("size_scale", 1.0)


def _windows(cfg, s_u, s_v):
    """(v1, u1) of each copy, in the reference's insertion order
    (_insertImage 300-333, _arrayInsertImage 336-361), with its checks."""
    n_u, n_v = cfg.N_u, cfg.N_v

    def window(extra):
        u1 = int(n_u / 2 - cfg.u_offset - s_u / 2)
        u2 = int(n_u / 2 - cfg.u_offset + s_u / 2)
        v1 = int(n_v / 2 - cfg.v_offset - extra - s_v / 2)
        v2 = int(n_v / 2 - cfg.v_offset - extra + s_v / 2)
        if u1 < 0 or u2 > n_u:
            raise Exception(
                "Image could not be inserted because it is either too large in the u-dimension or the offset "
                "causes it to extend out of the input screen size")
        if v1 < 0 or v2 > n_v:
            raise Exception(
                "Image could not be inserted because it is either too large in the v-dimension or the offset "
                "causes it to extend out of the input screen size")
        return v1, u1

    if cfg.array_num == 1:
        return [window(0)]
    out = []
    for k in range(cfg.array_num):
        if cfg.array_num % 2 == 0:
            d = int(np.ceil((k + 1) / 2) * (-1) ** k)          # 1, -1, 2, -2, ...
            extra = int(cfg.array_offset * d / 2 + cfg.v_offset)
        else:
            d = int(0 + np.ceil(k / 2) * (-1) ** (k + 1))      # 0, 1, -1, 2, -2, ...
            extra = int(cfg.array_offset * d + cfg.v_offset)
        out.append(window(extra))
    return out


def _rotation(shape_uv, angle_deg):
    """Output plane shape, matrix and offset of scipy.ndimage.rotate with
    reshape=True for the (u, v) plane of the given shape."""
    from scipy import special
    c, s = special.cosdg(angle_deg), special.sindg(angle_deg)
    rot = np.array([[c, s], [-s, c]])
    iy, ix = shape_uv
    bounds = rot @ [[0, 0, iy, iy], [0, ix, 0, ix]]
    out_shape = (np.ptp(bounds, axis=1) + 0.5).astype(int)
    out_center = rot @ ((out_shape - 1) / 2)
    in_center = (np.asarray(shape_uv) - 1) / 2
    return (int(out_shape[0]), int(out_shape[1])), rot, in_center - out_center


_OS_WORK = {}


def _order_stats(flat, rank):
    """sorted(flat)[rank], sorted(flat)[rank + 1] of a float64 device vector
    (vamx_order_stats: an exact radix selection; NaN when it holds a NaN)."""
    import ctypes
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    dev = flat.device
    ws = _OS_WORK.get(dev)
    if ws is None:
        nb = ctypes.c_int64()
        _lib.check(lib.vamx_order_stats_workspace_bytes(ctypes.byref(nb)), "vamx_order_stats_workspace_bytes")
        ws = _OS_WORK[dev] = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    out = torch.empty(2, dtype=torch.float64, device=dev)
    _lib.check(lib.vamx_order_stats(flat.data_ptr(), flat.numel(), int(rank), out.data_ptr(), ws.data_ptr(),
                                    _lib.stream_ptr(dev)), "vamx_order_stats")
    a, b = out.tolist()
    return a, b


def _percentile(t, q):
    """np.percentile(a, q) (linear method) of a float64 device tensor: the
    order statistics lo and lo + 1 on the device, numpy's interpolation on
    the host."""
    flat = t.reshape(-1).contiguous()
    n = flat.numel()
    vi = (n - 1) * (np.float64(q) / 100)
    if vi < 0:
        return _order_stats(flat, 0)[0]
    # at q = 100 numpy interpolates the last element with itself (an
    # infinite one gives NaN)
    lo = min(int(np.floor(vi)), n - 1)
    a, b = _order_stats(flat, lo)
    if a != a:
        return float("nan")
    g = vi - lo
    d = b - a
    return b - d * (1 - g) if g >= 0.5 else a + d * g


class ImageSeq:
    """imagesequence.py:85-182 on the GPU."""

    def __init__(self, image_config: ImageConfig, sinogram, device=None):
        if isinstance(sinogram, Sinogram):
            sinogram = sinogram.array
        elif not isinstance(sinogram, (np.ndarray, torch.Tensor)):
            raise TypeError("sinogram not specified.")
        cfg = image_config
        self.file_extension = ".imgseq"
        self.image_config = cfg
        if cfg.bit_depth > 32:
            raise ValueError("Bit depth higher than 32-bit is not supported.")
        dev = require_gpu(device if device is not None else
                          (sinogram.device if isinstance(sinogram, torch.Tensor) and sinogram.is_cuda else None))
        lib = _lib.load()
        m = torch.as_tensor(sinogram, device=dev).to(torch.float64)
        if m.ndim == 2:
            m = m[:, :, None]
        m = m.contiguous()
        flip_u, flip_v = int(bool(cfg.invert_u)), int(bool(cfg.invert_v))
        with torch.cuda.device(dev):
            st = _lib.stream_ptr(dev)
            if cfg.rotate_angle != 0.0:
                nu, na, nv = m.shape
                (mu, mv), rot, off = _rotation((nu, nv), cfg.rotate_angle)
                out = torch.empty((mu, na, mv), device=dev, dtype=torch.float64)
                mat = (ctypes.c_double * 4)(*rot.ravel())
                offs = (ctypes.c_double * 2)(*off)
                _lib.check(lib.vamx_imgseq_rotate(m.data_ptr(), nu, na, nv, flip_u, flip_v, mat, offs,
                                                  out.data_ptr(), mu, mv, st), "vamx_imgseq_rotate")
                m, flip_u, flip_v = out, 0, 0
            if cfg.size_scale != 1.0:
                nu, na, nv = m.shape
                mu, mv = int(nu * cfg.size_scale), int(nv * cfg.size_scale)
                out = torch.empty((mu, na, mv), device=dev, dtype=torch.float64)
                _lib.check(lib.vamx_imgseq_resize(m.data_ptr(), nu, na, nv, flip_u, flip_v, out.data_ptr(), mu, mv,
                                                  st), "vamx_imgseq_resize")
                m, flip_u, flip_v = out, 0, 0
            mu, na, mv = m.shape
            wins = _windows(cfg, mu, mv)
            if len(wins) > _lib.IMGSEQ_MAX_COPIES:
                raise ValueError("at most %d arrayed copies" % _lib.IMGSEQ_MAX_COPIES)
            prm = _lib.ImgSeqParams()
            prm.n_u, prm.n_v = int(cfg.N_u), int(cfg.N_v)
            prm.u1 = wins[0][1]
            prm.n_copies = len(wins)
            for k, (v1, _) in enumerate(wins):
                prm.v1[k] = v1
            prm.bit_depth = int(cfg.bit_depth)
            prm.normalize = int(cfg.normalization_percentile is not None)
            prm.norm_value = _percentile(m, cfg.normalization_percentile) if prm.normalize else 1.0
            prm.intensity_scale = float(cfg.intensity_scale)
            prm.flip_u, prm.flip_v = flip_u, flip_v
            self.frames = torch.empty((na, int(cfg.N_v), int(cfg.N_u)), device=dev, dtype=torch.uint8)
            _lib.check(lib.vamx_imgseq_compose(m.data_ptr(), mu, na, mv, ctypes.byref(prm),
                                               self.frames.data_ptr(), st), "vamx_imgseq_compose")
        self._images = None

    @property
    def images(self):
        if self._images is None:
            host = self.frames.cpu().numpy()
            self._images = [host[j] for j in range(host.shape[0])]
        return self._images

    def save(self, name: str):
        """Save as ``name.imgseq`` (a NumPy .npz container: frames and options)."""
        cfg = {k: v for k, v in vars(self.image_config).items() if v is not None}
        with open(name + self.file_extension, "wb") as f:
            np.savez(f, frames=np.stack(self.images), **{"cfg_" + k: np.asarray(v) for k, v in cfg.items()})

    def saveAsImages(self, save_dir: str, image_prefix: str = "image", image_type: str = ".png"):
        """imagesequence.py:276-289."""
        from PIL import Image
        for k, image in enumerate(self.images):
            save_path = os.path.join(save_dir, image_prefix + "%s" % str(k).zfill(4) + image_type)
            Image.fromarray(image).save(save_path, subsampling=0, quality=100)

    def saveAsVideo(self, *args, **kwargs):
        raise NotImplementedError("saveAsVideo needs OpenCV's VideoWriter, which vamtoolbox_amd does not use")

    def preview(self):
        raise NotImplementedError("preview needs a display player, which vamtoolbox_amd does not provide")


class _LoadedImageSeq:
    def __init__(self, frames, cfg):
        self.images = [frames[j] for j in range(frames.shape[0])]
        self.image_config = cfg
        self.file_extension = ".imgseq"

    saveAsImages = ImageSeq.saveAsImages


def loadImageSeq(file_name: str):
    """Load a sequence written by ImageSeq.save (frames as host arrays)."""
    with np.load(file_name, allow_pickle=False) as z:
        frames = z["frames"]
        opts = {k[4:]: z[k].item() for k in z.files if k.startswith("cfg_")}
    n_u, n_v = opts.pop("N_u"), opts.pop("N_v")
    return _LoadedImageSeq(frames, ImageConfig((n_u, n_v), **opts))
