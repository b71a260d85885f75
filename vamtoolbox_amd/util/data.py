"""Sinogram / target helpers of the optimisation path.

Mirrors ``vamtoolbox/util/data.py``: ``sigmoid`` (9-11), ``clipToCircle``
(14-40), ``filterTargetOSMO`` (43-113), ``filterSinogram`` (302-370),
``_get_fourier_filter`` (373-438) and ``discretize`` (495-523).  Every
function accepts a NumPy array (computed in float64 on the host, as the
reference) or a torch tensor (computed on the tensor's device; FFTs in
float64 with torch.fft).  Unlike the reference, ``clipToCircle`` returns a
new array instead of zeroing its argument in place.
"""

from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _is_t(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


def sigmoid(x, g):
    """util/data.py:9-11."""
    if _is_t(x):
        return 1.0 / (1.0 + torch.exp(-x * (1.0 / g)))
    y = 1 / (1 + np.exp(-x * (1 / g)))
    return y.astype(float)


def outside_mask(ny: int, nx: int) -> np.ndarray:
    """util/data.py:29-31: True outside the inscribed circle (shape (nx, ny))."""
    circle_y, circle_x = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx))
    return circle_x ** 2 + circle_y ** 2 > 1


def clipToCircle(x):
    """Zero everything outside the inscribed circle (util/data.py:14-40); returns a copy."""
    mask = outside_mask(x.shape[0], x.shape[1])
    if _is_t(x):
        m = torch.as_tensor(mask, device=x.device)
        if x.ndim == 3:
            m = m[..., None]
        return torch.where(m, torch.zeros((), dtype=x.dtype, device=x.device), x)
    x = np.array(x, copy=True)
    if x.ndim == 2:
        x[mask] = 0
    else:
        x[np.broadcast_to(mask[..., np.newaxis], x.shape)] = 0
    return x


def filterTargetOSMO(x, filter_name):
    """OSMO target pre-filter in 2-D Fourier space (util/data.py:43-113)."""
    filter_types = ("ram-lak", "shepp-logan", "cosine", "hamming", "hanning", None)
    if filter_name not in filter_types:
        raise ValueError("Unknown filter: %s" % filter_name)
    if filter_name is None or filter_name.lower() == "none":
        return x
    n_y, n_x = x.shape[0], x.shape[1]
    f_y, f_x = np.meshgrid(np.linspace(-1, 1, n_y), np.linspace(-1, 1, n_x), indexing="ij")
    f = np.sqrt(f_x ** 2 + f_y ** 2)
    fourier_filter = np.real(f).copy()
    if filter_name == "shepp-logan":
        omega = np.pi / 2 * f
        nz = omega != 0
        fourier_filter[nz] *= np.sin(omega[nz]) / omega[nz]
    elif filter_name == "cosine":
        fourier_filter *= np.cos(np.pi / 2 * f)
    elif filter_name == "hamming":
        w = np.abs(np.hamming(n_x))
        fourier_filter *= np.abs(np.sqrt(np.outer(w, w)))
    elif filter_name == "hanning":
        w = np.abs(np.hanning(n_x))
        fourier_filter *= np.abs(np.sqrt(np.outer(w, w)))
    if _is_t(x):
        xd = x.to(torch.float64)
        ff = torch.as_tensor(fourier_filter, device=x.device)
        if x.ndim == 3:
            ff = ff[:, :, None]
        X = torch.fft.fftshift(torch.fft.fft2(xd, dim=(0, 1)), dim=(0, 1))
        out = torch.fft.ifft2(torch.fft.ifftshift(X * ff, dim=(0, 1)), dim=(0, 1))
        return out.real
    from scipy.fft import fft2, fftshift, ifft2, ifftshift
    X = fftshift(fft2(x, axes=(0, 1)), axes=(0, 1))
    ff = fourier_filter if x.ndim == 2 else fourier_filter[:, :, np.newaxis]
    return np.real(ifft2(ifftshift(X * ff, axes=(0, 1)), axes=(0, 1)))


def _get_fourier_filter(size: int, filter_name: str):
    """Ram-Lak from the spatial-domain kernel, with windows (util/data.py:373-438)."""
    from scipy.fft import fft, fftfreq, fftshift
    n = np.concatenate((np.arange(1, size / 2 + 1, 2, dtype=int), np.arange(size / 2 - 1, 0, -2, dtype=int)))
    f = np.zeros(size)
    f[0] = 0.25
    f[1::2] = -1 / (np.pi * n) ** 2
    fourier_filter = 2 * np.real(fft(f))
    if filter_name == "ram-lak":
        pass
    elif filter_name == "shepp-logan":
        omega = np.pi * fftfreq(size)[1:]
        fourier_filter[1:] *= np.sin(omega) / omega
    elif filter_name == "cosine":
        freq = np.linspace(0, np.pi, size, endpoint=False)
        fourier_filter *= fftshift(np.sin(freq))
    elif filter_name == "hamming":
        fourier_filter *= fftshift(np.hamming(size))
    elif filter_name == "hanning":
        fourier_filter *= fftshift(np.hanning(size))
    elif filter_name == "ram-lak_freq":
        fourier_filter /= (np.pi / 4) ** 2
    elif filter_name == "ram-lak_scikit":
        fourier_filter *= np.pi / 2
    elif filter_name is None or filter_name.lower() == "none":
        fourier_filter[:] = 1
    return fourier_filter[:, np.newaxis]


def filterSinogram(sinogram, filter_name):
    """Ramp-filter each projection along the detector axis (util/data.py:302-370)."""
    filter_types = ("ram-lak", "shepp-logan", "cosine", "hamming", "hanning", "ram-lak_freq", "ram-lak_scikit", None)
    if filter_name is None or filter_name.lower() == "none":
        return sinogram
    if filter_name.lower() not in filter_types:
        raise ValueError("Unknown filter: %s" % filter_name)
    if sinogram.ndim not in (2, 3):
        raise Exception("The sinogram provided does not have compatible number of dimension (2 or 3).")
    pX = sinogram.shape[0]
    padded = max(64, int(2 ** np.ceil(np.log2(2 * pX))))
    ff = _get_fourier_filter(padded, filter_name)[:, 0]
    if _is_t(sinogram):
        s = sinogram.to(torch.float64)
        pad = torch.zeros((padded,) + tuple(s.shape[1:]), dtype=torch.float64, device=s.device)
        pad[:pX] = s
        fft_f = torch.as_tensor(ff, device=s.device).reshape((padded,) + (1,) * (s.ndim - 1))
        out = torch.fft.ifft(torch.fft.fft(pad, dim=0) * fft_f, dim=0).real[:pX]
        return out
    from scipy.fft import fft, ifft
    s = np.asarray(sinogram)
    pad_width = ((0, padded - pX),) + ((0, 0),) * (s.ndim - 1)
    img = np.pad(s, pad_width, mode="constant", constant_values=0)
    fshape = (padded,) + (1,) * (s.ndim - 1)
    return np.real(ifft(fft(img, axis=0) * ff.reshape(fshape), axis=0)[:pX])


def discretize(x, bit_depth: int, range: list, output_dtype=float):
    """util/data.py:495-523 (np.digitize semantics, index -1 wraps to the last bin)."""
    assert len(range) == 2, "range argument should have 2 elements [min,max]"
    if _is_t(x):
        r0 = float(range[0]) if not _is_t(range[0]) else range[0].item()
        r1 = float(range[1]) if not _is_t(range[1]) else range[1].item()
        bins = torch.as_tensor(np.linspace(r0, r1, 2 ** bit_depth, endpoint=True), device=x.device, dtype=torch.float64)
        inds = torch.bucketize(x.to(torch.float64).contiguous(), bins, right=True) - 1
        inds = torch.remainder(inds, bins.numel())
        return bins[inds].to(x.dtype)
    bins = np.linspace(range[0], range[1], 2 ** bit_depth, endpoint=True)
    inds = np.digitize(x, bins=bins) - 1
    return bins[inds].astype(output_dtype)
