"""Device plumbing: torch tensors carry device memory and the stream; every
computation is a libmivis kernel."""

from __future__ import annotations

import numpy as np
import torch

from . import _capi


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("mlx_vis_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def as_device_f32(X) -> torch.Tensor:
    """(n, d) float32 contiguous device tensor from numpy / torch input."""
    if isinstance(X, torch.Tensor):
        t = X.detach()
        if t.dtype != torch.float32:
            t = t.float()
        if t.device.type != "cuda":
            t = t.to(device(), non_blocking=False)
        return t.contiguous()
    arr = np.ascontiguousarray(np.asarray(X, dtype=np.float32))
    return torch.from_numpy(arr).to(device())


def require_device(what: str, *tensors) -> None:
    """Raises ValueError unless every given tensor (None skipped) lives on the
    GPU: the kernels take raw device pointers, and a host pointer would fault."""
    for t in tensors:
        if t is not None and (not isinstance(t, torch.Tensor) or t.device.type != "cuda"):
            raise ValueError(f"{what}: expected device (cuda) tensors, got {getattr(t, 'device', type(t))}")


def f32_rows(t: torch.Tensor) -> torch.Tensor:
    """The float32, row-major contiguous form of a device tensor (a copy only
    when the dtype or the strides differ): the kernels read raw row-major
    float32 memory, so a half-precision tensor or a strided view must never
    reach them as it is."""
    if t.dtype != torch.float32:
        t = t.float()
    return t.contiguous()


def i32_vec(t: torch.Tensor | None) -> torch.Tensor | None:
    """The int32 contiguous form of an index tensor (None passes through)."""
    if t is None:
        return None
    return t.to(torch.int32).contiguous()


def workspace(nbytes: int) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


def stream() -> int | None:
    return _capi.stream_handle()
