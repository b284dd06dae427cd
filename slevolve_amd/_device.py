"""Torch plumbing: device tensors, stream handles, host <-> device conversion."""
from __future__ import annotations

import numpy as np
import torch


def device(dev=None) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("slevolve_amd needs a HIP device (MI355X); none is visible")
    return torch.device(dev) if dev is not None else torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def as_device(a, dtype, dev, shape=None) -> torch.Tensor:
    """Device-resident, contiguous tensor of `dtype` for a numpy array or tensor (no copy if it
    already is one)."""
    if isinstance(a, torch.Tensor):
        t = a
        if t.device != dev or t.dtype != dtype or not t.is_contiguous():
            t = t.to(device=dev, dtype=dtype).contiguous()
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t
