"""GPU device selection: the product path runs on a HIP device or not at all."""

from __future__ import annotations

import torch


def require_gpu(device=None) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("vamtoolbox_amd needs a HIP GPU (torch.cuda.is_available() is False); "
                           "there is no CPU implementation of the propagators")
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)
