"""Device-memory plumbing: torch (ROCm) tensors hold every field that the
HIP kernels read or write."""
from __future__ import annotations

import numpy as np


def torch_mod():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('the HIP path needs a GPU (torch.cuda.is_available() is False)')
    return torch


def as_device(x, dtype=None):
    """numpy array or torch tensor -> contiguous CUDA tensor (float64 or
    complex128 unless ``dtype`` is given)."""
    torch = torch_mod()
    if isinstance(x, torch.Tensor):
        t = x
    else:
        a = np.asarray(x)
        if dtype is None:
            dtype = torch.complex128 if np.iscomplexobj(a) else torch.float64
        t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is None:
        dtype = torch.complex128 if t.is_complex() else torch.float64
    return t.to(device='cuda', dtype=dtype).contiguous()


def like_input(t, ref):
    """Return ``t`` (device) as numpy when ``ref`` was numpy."""
    if isinstance(ref, np.ndarray):
        return t.detach().cpu().numpy()
    return t
