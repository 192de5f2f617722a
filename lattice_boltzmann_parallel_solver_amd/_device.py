"""Moving reference-style arrays to and from the GPU for the drop-in functions.

numpy inputs are copied to the device, computed on, and copied back (the
parity path); float64/float32 torch device tensors are used in place (the
performance path).  torch is used only for device memory and the current
stream.  There is no CPU fallback.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def torch_cuda():
    """The torch module, after checking that a HIP device is visible."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the lattice-Boltzmann kernels have no CPU fallback")
    return torch


def to_device(a, dtype=None):
    """(contiguous device tensor, 'numpy'|'torch') for an array-like input."""
    torch = torch_cuda()
    if isinstance(a, torch.Tensor):
        t = a.detach()
        if t.device.type != "cuda":
            t = t.to("cuda")
        want = dtype or (t.dtype if t.dtype in (torch.float64, torch.float32) else torch.float64)
        return t.to(want).contiguous(), "torch"
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    t = torch.from_numpy(arr).to("cuda")
    if dtype is not None:
        t = t.to(dtype)
    return t, "numpy"


def back(t, kind):
    """Return a device result in the caller's array kind."""
    if kind == "torch":
        return t
    return t.cpu().numpy()


def dtype_code(t) -> int:
    torch = torch_cuda()
    if t.dtype == torch.float64:
        return _lib.DT_F64
    if t.dtype == torch.float32:
        return _lib.DT_F32
    raise TypeError(f"unsupported dtype {t.dtype}")


def stream() -> int:
    return torch_cuda().cuda.current_stream().cuda_stream


def write_back(dst, t) -> None:
    """Copy a device result into the caller's array in place (numpy or tensor)."""
    torch = torch_cuda()
    if isinstance(dst, torch.Tensor):
        if dst.data_ptr() != t.data_ptr():
            dst.copy_(t)
    else:
        dst[...] = t.cpu().numpy()
