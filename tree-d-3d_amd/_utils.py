"""Argument checks and array conversions (the reference's _utils.py:11-63).

The reference works on the arrays of cunumpy's active backend (NumPy or CuPy). Here every
computation runs on the GPU: NumPy (or list) inputs are copied to torch tensors on the
current device, and results come back in the caller's array type (NumPy for NumPy
inputs, torch tensors on the device for torch inputs).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from . import _abi


def _finite(name: str, value: float, *, minimum=None, positive=False) -> float:
    """_utils.py:11-20."""
    value = float(value)
    if not np.isfinite(value):
        raise ValueError(f"{name} must be finite")
    if minimum is not None and value < minimum:
        raise ValueError(f"{name} must be >= {minimum}")
    if positive and value <= 0:
        raise ValueError(f"{name} must be positive")
    return value


def _coincident() -> ValueError:
    return ValueError("Coincident charges require positive softening")


def _check_code(code: int) -> None:
    """_utils.py:26-35: raise for the failure a kernel reported in ``status[0]``."""
    if code > 0:
        raise _coincident()
    if code < 0:
        raise RuntimeError(
            "the quadtree is too deep for the traversal stack of the GPU kernels; "
            "lower max_depth"
        )


def _check(status: torch.Tensor) -> None:
    _check_code(int(status[0].item()))


def _is_torch(array: Any) -> bool:
    return isinstance(array, torch.Tensor)


def _device(array, dtype=torch.float64) -> torch.Tensor:
    """A contiguous tensor on the GPU (copying host data)."""
    dev = _abi.device()
    if isinstance(array, torch.Tensor):
        return array.to(device=dev, dtype=dtype).contiguous()
    np_dtype = np.float64 if dtype == torch.float64 else np.int64
    host = np.ascontiguousarray(np.asarray(array, dtype=np_dtype))
    return torch.from_numpy(host).to(dev)


def _points(array, name: str) -> torch.Tensor:
    """``array`` as a contiguous component-major ``(2, N)`` float64 GPU tensor."""
    points = _device(array)
    if points.ndim != 2 or points.shape[0] != 2:
        raise ValueError(
            f"{name} must have shape (2, N) (component-major), got {tuple(points.shape)}"
        )
    return points


def _host(array) -> np.ndarray:
    """A NumPy array (from either kind)."""
    if isinstance(array, torch.Tensor):
        return array.detach().cpu().numpy()
    return np.asarray(array)


def _out(tensor: torch.Tensor, like_torch: bool):
    """Return ``tensor`` in the caller's array type."""
    return tensor if like_torch else tensor.cpu().numpy()
