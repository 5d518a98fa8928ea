"""Tensor ingress/egress for the drop-in API.

The reference picks a backend from the input arrays (ot/backend.py:245-262)
and returns results of the same type, dtype and device
(test/test_bregman.py:351-387 assert this).  Here every input is moved once
to a ROCm device tensor; results go back to the caller's type:

* NumPy arrays and lists -> NumPy arrays of M's dtype;
* torch tensors -> torch tensors on the caller's device and dtype.

float32 and float64 run natively; other float dtypes compute in float32.
"""

import threading

import numpy as np
import torch

from . import _lib

_tls = threading.local()


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pot_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream


def host_ctrl():
    """Per-thread pinned host buffer used by the solver's stop polling."""
    buf = getattr(_tls, "ctrl", None)
    if buf is None:
        buf = torch.zeros(1024, dtype=torch.uint8, pin_memory=True)
        _tls.ctrl = buf
    return buf


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def is_empty(x):
    if x is None:
        return True
    if isinstance(x, list):
        return len(x) == 0
    return len(x) == 0


class Ingress:
    """Converts inputs to device tensors and outputs back to the input kind."""

    def __init__(self, *arrays):
        ref = None
        for x in arrays:
            if x is None or isinstance(x, list):
                continue
            if not is_empty(x):
                ref = x
                break
        if ref is None:
            ref = np.zeros(0)
        self.is_torch = isinstance(ref, torch.Tensor)
        if self.is_torch:
            self.out_dtype = ref.dtype
            self.out_device = ref.device
        else:
            ref_np = np.asarray(ref)
            self.out_dtype = ref_np.dtype if ref_np.dtype.kind == "f" else np.dtype(np.float64)
        self.device = require_gpu()
        if self.is_torch:
            cd = ref.dtype if ref.dtype in (torch.float32, torch.float64) else torch.float32
        else:
            cd = torch.float64 if self.out_dtype == np.float64 else torch.float32
            if self.out_dtype not in (np.float32, np.float64):
                cd = torch.float32
        self.dtype = cd
        self.code = _lib.F32 if cd == torch.float32 else _lib.F64

    def dev(self, x, dtype=None):
        """Device tensor (contiguous) of x in dtype (default: compute dtype)."""
        dtype = dtype or self.dtype
        if isinstance(x, torch.Tensor):
            t = x.detach()
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64)))
        return t.to(device=self.device, dtype=dtype).contiguous()

    def out(self, t):
        """Device tensor -> the caller's array kind."""
        if self.is_torch:
            return t.to(device=self.out_device, dtype=self.out_dtype)
        return t.detach().cpu().numpy().astype(self.out_dtype, copy=False)

    def scalar(self, val):
        """Scalar of the caller's kind (numpy scalar or 0-d tensor)."""
        if self.is_torch:
            return torch.tensor(val, dtype=self.out_dtype, device=self.out_device)
        return self.out_dtype.type(val)
