"""Device context: one HIP stream driven by one host thread (the sbx_ctx of the C ABI)."""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib


class Context:
    """Owns an sbx_ctx bound to a dedicated (non-default, default-priority) torch stream.

    All library launches go to `self.stream`; torch work that must be ordered with them should be
    issued `with torch.cuda.stream(ctx.stream)` or separated by `sync()`.
    """

    def __init__(self, device: int = 0):
        import torch

        self.lib = _lib.load()
        self.device = device
        torch.cuda.set_device(device)
        self.stream = torch.cuda.Stream(device=device)
        h = C.c_void_p()
        rc = self.lib.sbx_create(C.byref(h), device, C.c_void_p(self.stream.cuda_stream))
        if rc != _lib.SBX_OK:
            raise _lib.SoftboxError(rc, "sbx_create failed")
        self.h = h
        self._children = weakref.WeakSet()   # networks built on this context: closed before it

    def adopt(self, obj) -> None:
        self._children.add(obj)

    # ---------------------------------------------------------------- errors
    def check(self, rc: int) -> int:
        if rc < 0:
            msg = self.lib.sbx_last_error(self.h)
            raise _lib.SoftboxError(rc, (msg or b"").decode())
        return rc

    def clear_error(self) -> None:
        self.lib.sbx_clear_error(self.h)

    def sync(self) -> None:
        self.check(self.lib.sbx_sync(self.h))

    def close(self) -> None:
        for ch in list(getattr(self, "_children", ())):
            ch.close()
        if getattr(self, "h", None):
            self.lib.sbx_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter teardown order
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- device arrays
    def to_device(self, a: np.ndarray):
        """Copy a host array into a new device tensor (synchronous)."""
        import torch

        t = torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{self.device}")
        torch.cuda.synchronize(self.device)
        return t

    @staticmethod
    def to_host(t) -> np.ndarray:
        import torch

        torch.cuda.synchronize(t.device)
        return t.detach().cpu().numpy()


def ptr(t) -> C.c_void_p:
    """Device address of a contiguous tensor (None -> NULL)."""
    if t is None:
        return C.c_void_p(0)
    if not t.is_contiguous():
        raise ValueError("device arrays must be contiguous")
    return C.c_void_p(t.data_ptr())


def fptr(a) -> C.POINTER(C.c_float):
    """Host float32 array -> float* (the reference's tiny params arrays)."""
    arr = np.ascontiguousarray(a, dtype=np.float32)
    return arr.ctypes.data_as(C.POINTER(C.c_float)), arr


def iptr(a) -> C.POINTER(C.c_int):
    arr = np.ascontiguousarray(a, dtype=np.int32)
    return arr.ctypes.data_as(C.POINTER(C.c_int)), arr
