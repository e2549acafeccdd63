"""Device state of the diagonal forward-walking measurement: the operator
columns (gfmc_diag_create), the ancestry ring and the per-step workspace
(driver.py:240-288, _diag_ring_buffers :810-838)."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .observables import lower_diag_operators

__all__ = ["DiagRecorder"]


class DiagRecorder:
    """``ops`` ({name: operator}) on ``n_sites`` sites, measured at the lags
    ``m_list`` (p // 2 of the sorted windows) over ``n_walkers`` walkers.
    The ring is reset by :meth:`reset` (a fresh ring: identity parents, zero
    values, counter 0); every step loop that is given this recorder appends
    one row [n_cols][n_p] per step."""

    def __init__(self, ops, n_sites, m_list, n_walkers, device):
        self.lowering = lower_diag_operators(ops, n_sites)
        self.m_list = np.ascontiguousarray(m_list, dtype=np.int32)
        if np.any(np.diff(self.m_list) < 0) or np.any(self.m_list < 0):
            raise ValueError("m_list must be non-decreasing and >= 0")
        L = self.lowering
        self._keep = [L.kind, L.edge_off, L.edge_i, L.edge_j, L.edge_c, L.a, L.b, L.c,
                      self.m_list]
        d = _lib.DiagDesc()
        d.n_sites = int(n_sites)
        d.n_cols = L.n_cols
        i32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        f64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        d.kind, d.edge_off = i32(L.kind), i32(L.edge_off)
        d.edge_i, d.edge_j, d.edge_c = i32(L.edge_i), i32(L.edge_j), f64(L.edge_c)
        d.a, d.b, d.c = f64(L.a), f64(L.b), f64(L.c)
        d.n_p = len(self.m_list)
        d.m = i32(self.m_list)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gfmc_diag_create(ctypes.byref(d), ctypes.byref(h)), "gfmc_diag_create")
        self.handle = h
        self.n_walkers = int(n_walkers)
        self.n_cols = L.n_cols
        self.n_p = len(self.m_list)
        self.m_max = int(self.m_list[-1])
        nbytes = int(lib.gfmc_diag_ring_bytes(h, self.n_walkers))
        self.ring = torch.empty(nbytes, dtype=torch.uint8, device=device)
        nwork = int(lib.gfmc_diag_work_doubles(h, self.n_walkers))
        self.work = torch.empty(nwork, dtype=torch.float64, device=device)
        self.reset()

    @property
    def names(self):
        return self.lowering.names

    def reset(self):
        _lib.check(_lib.load().gfmc_diag_ring_init(self.handle, self.n_walkers,
                                                   _lib.ptr(self.ring), _lib.stream()),
                   "gfmc_diag_ring_init")

    def series(self, n, device):
        return torch.empty((n, self.n_cols, self.n_p), dtype=torch.float64, device=device)

    def values(self, spins):
        """(n_walkers, n_cols) device values of packed walkers."""
        out = torch.empty((spins.shape[0], self.n_cols), dtype=torch.float64,
                          device=spins.device)
        _lib.check(_lib.load().gfmc_diag_values(self.handle, spins.shape[0], _lib.ptr(spins),
                                                _lib.ptr(out), _lib.stream()),
                   "gfmc_diag_values")
        return out

    def record(self, spins, parents, row):
        """One step's record (gfmc_diag_record) into the device row
        [n_cols][n_p]."""
        _lib.check(_lib.load().gfmc_diag_record(self.handle, spins.shape[0], _lib.ptr(spins),
                                                _lib.ptr(parents), _lib.ptr(self.ring),
                                                _lib.ptr(row), _lib.ptr(self.work),
                                                _lib.stream()), "gfmc_diag_record")

    def split(self, rows):
        """Host rows (n, n_cols, n_p) -> {name: (n, n_p) | (n, n_vals, n_p)}."""
        return {k: np.ascontiguousarray(v) for k, v in self.lowering.split(rows).items()}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().gfmc_diag_destroy(self.handle)
        except Exception:
            pass
