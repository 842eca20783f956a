"""ctypes binding of the C ABI (include/mp_isdf_amd.h) for device-pointer work.

The pybind11 module ``mp_isdf`` covers the reference's host-array surface;
this thin binding exposes the device-level entry points (AO tiles, sketch
stages) to callers that hold device buffers (torch tensors, the benchmark).
Every call raises on a non-zero status with the library's message.
"""
from __future__ import annotations

import ctypes

from . import LIBRARY

_lib = None

_ERRORS = {1: ValueError, 2: OverflowError}


def lib():
    global _lib
    if _lib is None:
        import mp_isdf_amd  # noqa: F401  (torch-first import order)
        _lib = ctypes.CDLL(str(LIBRARY))
        _lib.mpisdf_last_error.restype = ctypes.c_char_p
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise _ERRORS.get(status, RuntimeError)(lib().mpisdf_last_error().decode())


def ptr(x) -> ctypes.c_void_p:
    """Device or host address of a torch tensor / numpy array (None -> NULL)."""
    if x is None:
        return ctypes.c_void_p(0)
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    return x.ctypes.data_as(ctypes.c_void_p)


def call(name: str, *args) -> None:
    conv = []
    for a in args:
        if isinstance(a, bool):
            conv.append(ctypes.c_int(int(a)))
        elif isinstance(a, int):
            conv.append(ctypes.c_int64(a))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        elif a is None or hasattr(a, "data_ptr") or hasattr(a, "ctypes"):
            conv.append(ptr(a))
        else:
            conv.append(a)
    check(getattr(lib(), name)(*conv))


class Context:
    """Owning handle of mpisdf_context for device-pointer AO evaluation."""

    def __init__(self, atm, bas, env, cart: bool = False):
        import numpy as np
        self._atm = np.ascontiguousarray(atm, dtype=np.int32)
        self._bas = np.ascontiguousarray(bas, dtype=np.int32)
        self._env = np.ascontiguousarray(env, dtype=np.float64)
        self.handle = ctypes.c_void_p()
        check(lib().mpisdf_context_create(ptr(self._atm), ctypes.c_int64(len(self._atm)), ptr(self._bas),
                                          ctypes.c_int64(len(self._bas)), ptr(self._env),
                                          ctypes.c_int64(len(self._env)), ctypes.c_int(int(cart)),
                                          ctypes.byref(self.handle)))
        n = ctypes.c_int64()
        check(lib().mpisdf_context_nao(self.handle, ctypes.byref(n)))
        self.nao = n.value

    def ao_eval(self, xyz, out, mode: int = 0, dtype: int = 0, ld: int | None = None, cstride: int | None = None,
                stream=None):
        n = xyz.shape[0]
        ld = self.nao if ld is None else ld
        cstride = n * ld if cstride is None else cstride
        check(lib().mpisdf_ao_eval_device(self.handle, ptr(xyz), ctypes.c_int64(n), ctypes.c_int(mode),
                                          ctypes.c_int(dtype), ptr(out), ctypes.c_int64(ld), ctypes.c_int64(cstride),
                                          ctypes.c_void_p(stream or 0)))

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value and _lib is not None:
            _lib.mpisdf_context_destroy(self.handle)
            self.handle = ctypes.c_void_p()
