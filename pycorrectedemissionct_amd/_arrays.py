"""Host <-> device array plumbing.

corrct's API is NumPy in, NumPy out (fresh arrays the caller may mutate,
``projectors.py:276, 318``). The HIP backend accepts NumPy arrays (copied to
the current device) or torch device tensors (used in place), and returns the
same kind it was given.
"""

from __future__ import annotations

import numpy as np
import torch


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(x, dtype=torch.float32) -> tuple[torch.Tensor, bool]:
    """(contiguous float32 device tensor, whether the input was host data)."""
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            t = x
            if t.dtype != dtype:
                t = t.to(dtype)
            return t.contiguous(), False
        return x.to(device=device(), dtype=dtype).contiguous(), True
    arr = np.ascontiguousarray(np.asarray(x), dtype=np.float32 if dtype == torch.float32 else None)
    return torch.from_numpy(arr).to(device()), True


def from_dev(t: torch.Tensor, to_host: bool):
    if to_host:
        return t.detach().cpu().numpy()
    return t


def is_host(x) -> bool:
    return not (isinstance(x, torch.Tensor) and x.is_cuda)
