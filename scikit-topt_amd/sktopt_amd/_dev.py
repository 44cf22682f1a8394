"""Device plumbing: torch owns the HBM buffers and the stream; the kernels
are the library's.  Every compute entry point requires a visible GPU."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("sktopt_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def to_dev(x, dtype=torch.float64) -> torch.Tensor:
    """numpy / torch -> contiguous device tensor of ``dtype`` (no copy if it
    already is one)."""
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=dtype)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)
    return t.contiguous()


def is_host(x) -> bool:
    return not (isinstance(x, torch.Tensor) and x.is_cuda)
