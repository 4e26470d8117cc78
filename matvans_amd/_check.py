"""Argument checks shared by every Python binding that hands a tensor's address to the
library.  The C ABI takes plain device pointers, so a host tensor, a tensor on another
device, the wrong element type, a short buffer or a strided view would otherwise reach a
kernel as a bad address or a wrong extent.  Each check raises ValueError naming the
argument."""
from __future__ import annotations


def _torch():
    import torch
    return torch


def is_device_tensor(t) -> bool:
    """True for a tensor in HIP device memory (torch's "cuda" device type on ROCm)."""
    return t.device.type == "cuda"


def tensor(name, t, *, dtype, device=None, shape=None, min_numel=None, dtypes=None):
    """Require a torch tensor on a HIP device (``device`` if given), of ``dtype`` (or one
    of ``dtypes``), with ``shape`` (None entries match any extent) or at least
    ``min_numel`` elements, whose last dimension is contiguous and whose rows (2-D) do not
    overlap.  Returns ``t``."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if not is_device_tensor(t):
        raise ValueError(f"{name}: must be a device tensor (got {t.device})")
    if device is not None and t.device != torch.device(device):
        raise ValueError(f"{name}: on {t.device}, expected {device}")
    ok = (t.dtype in dtypes) if dtypes else (t.dtype == dtype)
    if not ok:
        want = " or ".join(str(d) for d in dtypes) if dtypes else str(dtype)
        raise ValueError(f"{name}: dtype {t.dtype}, expected {want}")
    if shape is not None:
        if t.dim() != len(shape) or any(s is not None and s != d for s, d in zip(shape, t.shape)):
            want = "[" + ", ".join("*" if s is None else str(s) for s in shape) + "]"
            raise ValueError(f"{name}: shape {list(t.shape)}, expected {want}")
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError(f"{name}: {t.numel()} elements, at least {min_numel} needed")
    if t.numel() > 0 and t.dim() >= 1 and t.shape[-1] > 1 and t.stride(-1) != 1:
        raise ValueError(f"{name}: the last dimension must be contiguous (stride "
                         f"{t.stride(-1)})")
    if t.dim() == 2 and t.shape[0] > 1 and t.stride(0) < t.shape[1]:
        raise ValueError(f"{name}: rows overlap (row stride {t.stride(0)} < {t.shape[1]})")
    if t.dim() > 2:
        raise ValueError(f"{name}: expected at most 2 dimensions, got {t.dim()}")
    return t


def vectors(name, v, *, cols, device, dtype, max_vec=8):
    """A [cols] or [n_vec, cols] batch of vectors (1 <= n_vec <= max_vec); returns the 2-D
    view."""
    tensor(name, v, dtype=dtype, device=device)
    vv = v.unsqueeze(0) if v.dim() == 1 else v
    if vv.dim() != 2 or vv.shape[-1] != cols:
        raise ValueError(f"{name}: shape {list(v.shape)}, expected [{cols}] or [n_vec, {cols}]")
    if not 1 <= vv.shape[0] <= max_vec:
        raise ValueError(f"{name}: {vv.shape[0]} vectors, expected 1..{max_vec}")
    return vv


def output(name, out, *, shape, dtype, device):
    """A caller-supplied output of exactly ``shape`` (2-D) or its 1-D squeeze."""
    torch = _torch()
    tensor(name, out, dtype=dtype, device=device)
    want = torch.Size(shape)
    if out.shape != want and not (shape[0] == 1 and out.shape == want[1:]):
        raise ValueError(f"{name}: shape {list(out.shape)}, expected {list(shape)}")
    if out.dim() == 2 and shape[0] > 1 and out.stride(0) != shape[1]:
        raise ValueError(f"{name}: must be contiguous")
    return out
