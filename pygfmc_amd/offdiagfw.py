"""Device state of the off-diagonal forward-walking measurement: one shadow
chain per operator and measured step (forward_walk_operator,
evolution.py:892-1101; the production-step call, driver.py:290-331), the
measurement grid (_measurement_plan / _measured_steps, driver.py:478-566)
and the operator positivity check (estimators.py:631-681)."""
from __future__ import annotations

import ctypes
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from .hamiltonian import SpinHalf
from .observables import Identity, SzSzStructureFactor

__all__ = ["OffDiagMeasurer", "measurement_plan", "measured_steps", "check_even_p"]

N_POSITIVITY_CHECK = 100


def check_even_p(off_ops, p_sorted):
    """Every p even when an operator has off-diagonal slots (driver.py:1451-1461)."""
    if any(int(getattr(op, "M_total", 1)) > 1 for op in off_ops.values()):
        bad = [p for p in p_sorted if p % 2 != 0]
        if bad:
            raise ValueError(f"All p values must be divisible by 2 when off_diag_operators have "
                             f"off-diagonal elements (M_total > 1); got {bad}")


def measurement_plan(n_steps, p_sorted, n_off_ops, off_diag_stride, return_raw):
    """The window grid and the off-diagonal measurement grid of one call
    (_measurement_plan, driver.py:478-551): with the raw series returned the
    measured steps are every stride-th; without them (equal windows) only
    the steps every window can use, [p_max, n_steps - p_max // 2 - 1]."""
    sorted_p = list(p_sorted)
    m_array = [p // 2 for p in sorted_p]
    equal_windows = not bool(return_raw)
    p_max = sorted_p[-1] if sorted_p else 0
    m_max_od = p_max // 2
    if p_max >= n_steps:
        raise ValueError(f"every p value must be smaller than n_steps: got n_steps={n_steps} "
                         f"and max(p)={p_max}")
    if n_off_ops > 0 and n_steps <= p_max + m_max_od:
        raise ValueError(f"off-diagonal operators need n_steps > 1.5*max(p) (max(p) + "
                         f"max(p)//2 = {p_max + m_max_od}: every window p needs p past and "
                         f"p//2 future steps in this call), got n_steps={n_steps}")
    gate = n_off_ops > 0 and (off_diag_stride > 1 or equal_windows)
    if equal_windows and n_off_ops > 0:
        offset = next((k for k in range(p_max, p_max + off_diag_stride)
                       if (k + 1) % off_diag_stride == 0), p_max)
        if not any((k + 1) % off_diag_stride == 0 for k in range(p_max, n_steps - m_max_od)):
            raise ValueError(
                f"equal off-diagonal windows (return_raw=False) measure only the steps "
                f"{p_max}..{n_steps - m_max_od - 1} (n_steps={n_steps} minus the lead-in "
                f"max(p)={p_max} and the tail max(p)//2={m_max_od}), and none of them is a "
                f"multiple of off_diag_stride={off_diag_stride} (counted from 1); raise "
                f"n_steps or lower the stride")
    else:
        offset = off_diag_stride - 1
    return SimpleNamespace(p_array=sorted_p, m_array=m_array, p_max=p_max,
                           m_max=max(m_array) if m_array else 0, n_off_ops=n_off_ops,
                           off_diag_stride=off_diag_stride, equal_windows=equal_windows,
                           gate_off_diag=gate, off_diag_offset=offset)


def measured_steps(n_steps, plan):
    """Boolean (n_steps,): the 0-based steps that measure the off-diagonal
    operators, (t + 1) % stride == 0, restricted with equal windows to the
    band every p can use (_measured_steps, driver.py:554-566)."""
    t = np.arange(n_steps)
    measured = (t + 1) % plan.off_diag_stride == 0
    if plan.equal_windows and plan.n_off_ops > 0:
        measured &= (t >= plan.p_max) & (t <= n_steps - plan.p_max // 2 - 1)
    return measured


class OffDiagMeasurer:
    """``ops`` ({name: operator}) measured by forward walking on the chains
    of Hamiltonian problem ``h`` (a :class:`~pygfmc_amd.problem.Problem`)
    with guide ``guide``, at the lags ``m_list`` over ``n_walkers`` walkers.

    A SpinHalf with off-diagonal slots is lowered like a Hamiltonian with the
    run's guide and applied by gfmc_apply_operator; a purely diagonal
    operator (a SpinHalf without off-diagonal slots, :class:`Identity`)
    starts its shadow chain from its per-walker values
    (gfmc_diag_values)."""

    def __init__(self, ops, h, guide, m_list, n_walkers, device):
        from .diagfw import DiagRecorder
        from .problem import get_problem
        self.names = list(ops)
        self.h = h
        self.m_list = [int(m) for m in m_list]
        if not self.m_list or len(self.m_list) > 64:
            raise ValueError("between 1 and 64 windows are supported")
        self.n_p = len(self.m_list)
        self.m_max = max(self.m_list)
        self.n_walkers = int(n_walkers)
        self._m = (ctypes.c_int32 * self.n_p)(*self.m_list)
        self.entries = []
        lib = _lib.load()
        nbytes = 0
        for name, op in ops.items():
            if not isinstance(name, str):
                raise TypeError(f"operator names must be strings; got {name!r}")
            if isinstance(op, SzSzStructureFactor):
                raise TypeError(f"off-diagonal operator {name!r}: a vector-valued operator is "
                                "measured as a diagonal operator (diag_operators)")
            if isinstance(op, SpinHalf) and op.n_off > 0:
                if op.N != h.n_sites:
                    raise ValueError(f"off-diagonal operator {name!r} has {op.N} sites, the "
                                     f"Hamiltonian {h.n_sites}")
                prob = get_problem(op, guide, h.precision)
                prob.reserve(self.n_walkers)
                self.entries.append(SimpleNamespace(name=name, op=op, prob=prob, rec=None))
                handle = prob.handle
            elif isinstance(op, (SpinHalf, Identity)):
                rec = DiagRecorder({name: op}, h.n_sites, [0], self.n_walkers, device)
                self.entries.append(SimpleNamespace(name=name, op=op, prob=None, rec=rec))
                handle = None
            else:
                raise TypeError(f"off-diagonal operator {name!r} of type {type(op).__name__} "
                                "is not supported on the GPU (SpinHalf and Identity are)")
            nbytes = max(nbytes, int(lib.gfmc_fw_operator_work_bytes(
                h.handle, handle, self.n_walkers, self.m_max)))
        self._work = torch.empty(nbytes + 256, dtype=torch.uint8, device=device)
        off = (-self._work.data_ptr()) % 256
        self.work = self._work[off:off + nbytes]

    @property
    def n_ops(self):
        return len(self.entries)

    def rows(self, n, device):
        """Device buffers (n, n_ops, n_p) for the values and the log norms."""
        return (torch.empty((n, self.n_ops, self.n_p), dtype=torch.float64, device=device),
                torch.empty((n, self.n_ops, self.n_p), dtype=torch.float64, device=device))

    def measure(self, spins, beta, alw, seed, step, fw_row, norm_row):
        """One measurement at main-chain position ``step`` on the
        post-resampling walkers ``spins``: every operator's (val, log_norm)
        at the lags into the device rows (n_ops, n_p)."""
        lib = _lib.load()
        st = _lib.stream()
        for i, e in enumerate(self.entries):
            vals = None
            if e.rec is not None:
                vals = e.rec.values(spins)
            _lib.check(lib.gfmc_forward_walk_operator(
                self.h.handle, e.prob.handle if e.prob is not None else None,
                self.n_walkers, _lib.ptr(spins), _lib.ptr(vals), self.n_p, self._m,
                float(beta), float(alw), int(seed) & 0xFFFFFFFFFFFFFFFF,
                int(step) & 0xFFFFFFFF, _lib.TAG_OPERATOR + i, _lib.TAG_PROPAGATE,
                _lib.TAG_RECONFIG, _lib.ptr(self.work), _lib.ptr(fw_row[i]),
                _lib.ptr(norm_row[i]), st), "gfmc_forward_walk_operator")

    def check_positivity(self, spins, verbose=False):
        """Every guide-transformed rate of every operator with off-diagonal
        slots non-negative and finite on the first 100 walkers
        (check_off_diag_operator_positivity, estimators.py:631-681)."""
        n = min(N_POSITIVITY_CHECK, spins.shape[0])
        for e in self.entries:
            if e.prob is None:
                continue
            r, _, _ = e.prob.rates(spins[:n].contiguous(), want_eloc=False)
            r = r.cpu().numpy()
            if not (np.all(np.isfinite(r)) and np.all(r >= 0)):
                raise ValueError(f"Operator '{e.name}': negative rate detected (min = "
                                 f"{float(np.nanmin(r)):.6f}). All rates must be >= 0.")
        if verbose:
            print(f"[estimators] - [check_off_diag_operator_positivity] - all operators passed "
                  f"positivity check ({n} configs, {len(self.entries)} operators)", flush=True)
