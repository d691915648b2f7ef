"""Host-side helpers: input coercion to device tensors and dtype bookkeeping."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L

_CPLX = {torch.float64: torch.complex128, torch.float32: torch.complex64,
         torch.complex128: torch.complex128, torch.complex64: torch.complex64}
_REAL = {torch.complex128: torch.float64, torch.complex64: torch.float32}

KIND = {torch.float64: 0, torch.float32: 1, torch.complex128: 2, torch.complex64: 3}


def capture_graphs(dev, fn, n=1):
    """Record n successive calls of fn() (launches only) as HIP graphs.

    Captured on a side stream of default priority.  The cyclic garbage
    collector is held off until the captures end: a collection inside a
    capture could destroy another graph or free the device memory of an
    unreachable solver state while the stream is capturing.
    """
    import gc
    was = gc.isenabled()
    gc.disable()
    graphs = []
    try:
        cur = torch.cuda.current_stream(dev)
        cap = torch.cuda.Stream(dev)
        cap.wait_stream(cur)
        with torch.cuda.stream(cap):
            for _ in range(n):
                g = torch.cuda.CUDAGraph()
                g.capture_begin()
                try:
                    fn()
                finally:
                    g.capture_end()
                graphs.append(g)
        cur.wait_stream(cap)
    finally:
        if was:
            gc.enable()
    return graphs


def default_device():
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cuda", 0)


def to_tensor(x, device=None, dtype=None) -> torch.Tensor:
    """Any array-like -> contiguous device tensor (optionally cast)."""
    device = device or default_device()
    if isinstance(x, torch.Tensor):
        t = x
        if t.device != torch.device(device):
            t = t.to(device)
    else:
        a = np.asarray(x)
        if a.dtype.kind == "c" and a.dtype.itemsize > 16:
            a = a.astype(np.complex128)
        elif a.dtype.kind == "f" and a.dtype.itemsize > 8:
            a = a.astype(np.float64)
        elif a.dtype.kind in "iub":
            a = a.astype(np.float64)
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    if dtype is not None and t.dtype != dtype:
        t = convert(t, dtype)
    return t.contiguous()


def convert(t: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """Precision / kind conversion on the device through libhfx (no torch compute)."""
    if t.dtype == dtype:
        return t
    if not t.is_cuda or t.dtype not in KIND or dtype not in KIND:
        return t.to(dtype)
    src = t.contiguous()
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    L.call("hfx_convert", L.ptr(src), KIND[src.dtype], L.ptr(out), KIND[dtype], src.numel(),
           L.stream(src.device))
    return out


def complex_dtype_of(t: torch.Tensor) -> torch.dtype:
    return _CPLX[t.dtype]


def real_dtype_of(cdt: torch.dtype) -> torch.dtype:
    return _REAL[cdt]


def host_array(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
