"""roctx ranges around the engine's phases (the warm-up stages and the
production chunks), seen by ``rocprofv3 --marker-trace`` next to the kernels
they enqueue.  Enabled by ``PYGFMC_ROCTX=1``; without it, or without the
ROCm roctx library, ``span`` does nothing."""
from __future__ import annotations

import contextlib
import ctypes
import os

_LIB = None
_TRIED = False
_NAMES = ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4",
          "libroctx64.so")


def library():
    """The roctx library when tracing is enabled and it loads, else None."""
    global _LIB, _TRIED
    if os.environ.get("PYGFMC_ROCTX", "0") in ("", "0"):
        return None
    if not _TRIED:
        _TRIED = True
        for name in _NAMES:
            for path in (os.path.join("/opt/rocm/lib", name), name):
                try:
                    lib = ctypes.CDLL(path)
                except OSError:
                    continue
                lib.roctxRangePushA.restype = ctypes.c_int
                lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                lib.roctxRangePop.restype = ctypes.c_int
                lib.roctxRangePop.argtypes = []
                _LIB = lib
                break
            if _LIB is not None:
                break
    return _LIB


@contextlib.contextmanager
def span(name):
    """A roctx range named ``name`` around the block (nested ranges nest)."""
    lib = library()
    if lib is None:
        yield None
        return
    level = lib.roctxRangePushA(name.encode())
    try:
        yield level
    finally:
        lib.roctxRangePop()
