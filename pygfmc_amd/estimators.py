"""Windowed energy estimator and its blocked jackknife (host, FP64 NumPy).

The per-step series (log G_n, E_n) come off the GPU once per chunk; the
estimator and the jackknife are O(n_steps) host work, as in the reference
(estimators.py:268-305, io.py:921-1157).
"""
from __future__ import annotations

import warnings

import numpy as np

__all__ = ["validated_p_array", "estimate_energy", "estimate_diagonal", "estimate_off_diagonal",
           "evaluate_observables", "DeviceEnergyEstimator", "validated_int",
           "check_population"]


def validated_int(name, value, minimum):
    """Integer-valued control parameter >= minimum (_validation.py:7-19)."""
    if isinstance(value, (bool, np.bool_)):
        raise ValueError(f"{name} must be an integer >= {minimum}, got {value!r}.")
    try:
        result = int(value)
        ok = float(result) == float(value)
    except (TypeError, ValueError, OverflowError):
        ok, result = False, minimum - 1
    if not ok or result < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}, got {value!r}.")
    return result


def validated_p_array(p_array):
    """Sorted unique positive integer windows (estimators.py:30-42)."""
    if p_array is None or len(p_array) == 0:
        raise ValueError("p_array must contain at least one positive window size.")
    out = []
    for p in p_array:
        p_int = int(p)
        if float(p_int) != float(p):
            raise ValueError(f"p_array values must be integers, got {p!r}.")
        if p_int <= 0:
            raise ValueError(f"p_array values must be > 0, got {p_int}.")
        out.append(p_int)
    return sorted(set(out))


def estimate_energy(p_array, log_weights, E):
    """Mixed estimator <E>_p = sum_i G_p(i) E(i+p) / sum_i G_p(i) with
    log G_p(i) = C[i+p] - C[i], C the cumulative sum of log G_n, shifted by
    its max before the exponential (estimators.py:268-305, 595)."""
    w = np.asarray(log_weights, dtype=np.float64)
    E = np.asarray(E, dtype=np.float64)
    if w.ndim != 1 or E.ndim != 1:
        raise ValueError(f"log_weights and E must be one-dimensional; got {w.shape} and {E.shape}.")
    if len(w) != len(E):
        raise ValueError(f"log_weights and E must have equal lengths; got {len(w)} and {len(E)}.")
    if len(w) == 0:
        raise ValueError("log_weights and E must contain at least one sample.")
    sorted_p = validated_p_array(p_array)
    n = len(w)
    if not np.all(np.isfinite(E)):
        raise ValueError("Energy samples contain nonfinite values.")
    too_large = [p for p in sorted_p if p >= n]
    if too_large:
        raise ValueError(f"Each p value must be smaller than the number of available "
                         f"samples ({n}); invalid values: {too_large}.")
    if not np.all(np.isfinite(w)):
        raise ValueError("Branch weights contain nonfinite values.")
    C = np.cumsum(w)
    out = {}
    for p in sorted_p:
        log_G_p = C[p:] - C[:-p]
        shift = np.max(log_G_p)
        G_p = np.exp(log_G_p - shift)
        out[p] = float(np.sum(G_p * E[p:])) / float(np.sum(G_p))
    return out


def estimate_diagonal(p_array, log_weights, diag, m_max_common=0):
    """Forward-walking estimates of diagonal operators,
    obs[name][p] = sum_i G_p(i) D(i+p) / sum_i G_p(i) over the windows
    i >= m_max_common, with D the operator's series at p's column and G_p
    the energy estimator's window weights (estimators.py:268-305, 540-622).
    ``diag``: {name: (n, n_p) | (n, n_vals, n_p)}; a scalar per p for the
    former, an (n_vals,) array for the latter.  ``m_max_common`` is the
    reference's common window trim, max(m) when off-diagonal operators share
    the run, else 0."""
    w = np.asarray(log_weights, dtype=np.float64)
    sorted_p = validated_p_array(p_array)
    n = len(w)
    too_large = [p for p in sorted_p if p >= n]
    if too_large:
        raise ValueError(f"Each p value must be smaller than the number of available "
                         f"samples ({n}); invalid values: {too_large}.")
    if not np.all(np.isfinite(w)):
        raise ValueError("Branch weights contain nonfinite values.")
    m_max_common = int(m_max_common)
    short = [p for p in sorted_p if n - p - m_max_common <= 0]
    if short:
        raise ValueError(f"Insufficient samples for the window trim at p={short}; each "
                         f"requires n_samples > p + {m_max_common}.")
    arrs = {}
    for name, arr in diag.items():
        a = np.asarray(arr, dtype=np.float64)
        if a.ndim not in (2, 3):
            raise ValueError(f"diag['{name}'] must be (n_samples, n_p) for a scalar operator "
                             f"or (n_samples, n_vals, n_p) for a vector-valued one; got "
                             f"{a.shape}.")
        if a.shape[0] != n or a.shape[-1] != len(sorted_p):
            raise ValueError(f"diag['{name}'] has shape {a.shape}; its first axis must be the "
                             f"{n} samples and its last the {len(sorted_p)} p values.")
        if not np.all(np.isfinite(a)):
            raise ValueError(f"diag['{name}'] contains nonfinite values.")
        arrs[name] = a
    C = np.cumsum(w)
    out = {name: {} for name in arrs}
    for j, p in enumerate(sorted_p):
        log_G = (C[p:] - C[:-p])[m_max_common:]
        G = np.exp(log_G - np.max(log_G))
        G_sum = float(np.sum(G))
        for name, a in arrs.items():
            if a.ndim == 2:
                out[name][p] = float(np.sum(G * a[:, j][p:][m_max_common:])) / G_sum
            else:
                out[name][p] = np.sum(G[:, None] * a[:, :, j][p:][m_max_common:],
                                      axis=0) / G_sum
    return out


def _dense(series, names, what):
    """{name: (n_rows, n_p)} in ``names`` order (or a dense array) ->
    (n_rows, n_ops, n_p) (dense_and_names, estimators.py:120-150)."""
    if isinstance(series, dict):
        if names is None:
            names = list(series)
        if set(series) != set(names):
            raise ValueError(f"{what} is keyed {sorted(series)}, expected {sorted(names)}")
        return np.stack([np.asarray(series[k], dtype=np.float64) for k in names], axis=1), names
    a = np.asarray(series, dtype=np.float64)
    if names is None:
        names = [f"off_diag_{k}" for k in range(a.shape[1])]
    return a, list(names)


def estimate_off_diagonal(p_array, log_weights, off_diag, off_diag_norm, off_diag_steps,
                          m_array=None, m_max_common=0, op_names=None):
    """Forward-walking pure estimates of the off-diagonal operators
    (_off_diagonal_estimates, estimators.py:308-355).  Per measured step k
    and window p (m = p // 2), with L the cumulative log G of the main chain
    (L[0] = 0):
        numerator term   = exp(L[k+1] - L[k-m+1] + log_norm(k)) val(k),
        denominator term = exp(L[k+m+1] - L[k-m+1]),
    summed over the measured k with p + max(0, m_max_common - m) <= k <
    n - m, exponentials shifted by the larger maximum of the two.
    ``off_diag`` / ``off_diag_norm``: {name: (n_rows, n_p)} (or dense
    (n_rows, n_ops, n_p) with ``op_names``); ``off_diag_steps`` the
    increasing 0-based step of each row.  Returns {name: {p: value}}."""
    w = np.asarray(log_weights, dtype=np.float64)
    sorted_p = validated_p_array(p_array)
    m_arr = [p // 2 for p in sorted_p] if m_array is None else [int(m) for m in m_array]
    if len(m_arr) != len(sorted_p):
        raise ValueError(f"m_array must have one entry per distinct p value; got "
                         f"{len(m_arr)} entries for {len(sorted_p)} windows.")
    vals, names = _dense(off_diag, op_names, "off_diag")
    norms, _ = _dense(off_diag_norm, names, "off_diag_norm")
    n = len(w)
    if vals.ndim != 3 or vals.shape[1] != len(names) or vals.shape[2] != len(sorted_p):
        raise ValueError(f"off_diag must have shape (n_measurements, n_operators, n_p); got "
                         f"{vals.shape}.")
    if norms.shape != vals.shape:
        raise ValueError(f"off_diag_norm must match off_diag exactly; got {norms.shape} and "
                         f"{vals.shape}.")
    if not (np.all(np.isfinite(vals)) and np.all(np.isfinite(norms))):
        raise ValueError("off_diag and off_diag_norm must contain only finite values.")
    if not np.all(np.isfinite(w)):
        raise ValueError("Branch weights contain nonfinite values.")
    k_all = np.asarray(off_diag_steps, dtype=np.int64)
    if k_all.shape != (vals.shape[0],):
        raise ValueError(f"off_diag_steps must hold one step per row ({vals.shape[0]}); got "
                         f"{k_all.shape}.")
    if k_all.size and (np.any(np.diff(k_all) <= 0) or k_all[0] < 0 or k_all[-1] >= n):
        raise ValueError(f"off_diag_steps must be strictly increasing steps in [0, {n}).")
    short = [p for p, m in zip(sorted_p, m_arr) if n - m - p <= 0]
    if short:
        raise ValueError(f"Insufficient samples for off-diagonal windows {short}; each "
                         "requires n_samples > p + m.")
    L = np.concatenate([[0.0], np.cumsum(w)])
    out = {name: {} for name in names}
    for j, (p, m) in enumerate(zip(sorted_p, m_arr)):
        lo_i = max(0, int(m_max_common) - m)
        lo, hi = np.searchsorted(k_all, [p + lo_i, n - m])
        rows = slice(lo, hi)
        i_sel = k_all[rows] - p
        if len(i_sel) == 0:
            raise ValueError(f"No stored off-diagonal measurement is usable for p={p}, m={m}.")
        log_num = L[p + 1 + i_sel] - L[p - m + 1 + i_sel]
        log_den = L[p + m + 1 + i_sel] - L[p - m + 1 + i_sel]
        for idx, name in enumerate(names):
            fw = vals[rows, idx, j]
            log_terms = log_num + norms[rows, idx, j]
            fin = np.isfinite(log_terms)
            shift = (max(np.max(log_terms[fin]), np.max(log_den)) if fin.any()
                     else np.max(log_den))
            num = float(np.sum(np.exp(log_terms - shift) * fw))
            den = float(np.sum(np.exp(log_den - shift)))
            out[name][p] = num / den
    return out


def evaluate_observables(raw, p_array=None, n_blocks=None):
    """Re-estimate the energy and the operators of ``raw`` (``log_weights``,
    ``energy``, ``p_array``, and the operator series ``diag``, ``off_diag``,
    ``off_diag_norm``, ``off_diag_steps`` when present), optionally with the
    blocked jackknife error: delete each of ``n_blocks`` contiguous blocks,
    recompute the whole estimator on the rest (off-diagonal rows of the
    deleted block dropped, the others renumbered into the shortened chain),
    err = sqrt((nb-1)/nb * sum (x_b - mean)^2) (io.py:921-1157).  Trailing
    samples that do not fill a block are dropped from both the estimate and
    the error.  The operators are estimated on the run's own grid; a
    different ``p_array`` re-estimates the energy only.  With both diagonal
    and off-diagonal operators every window starts at i >= max(m) (the
    common trim of estimators.py:595-611).  Returns ``(energy, obs)`` or
    ``(energy, obs, energy_err, obs_err)``."""
    if raw.get("energy") is None:
        raise ValueError("raw holds no 'energy' series, so nothing can be re-estimated.")
    if raw.get("log_weights") is None:
        raise ValueError("raw holds no 'log_weights' series; the windowed estimator "
                         "cannot be rebuilt without the per-step branch weights.")
    E = np.asarray(raw["energy"], dtype=np.float64)
    lw = np.asarray(raw["log_weights"], dtype=np.float64)
    stored = raw.get("p_array")
    stored = validated_p_array(stored) if stored is not None else None
    grid = validated_p_array(p_array) if p_array is not None else stored
    if grid is None:
        raise ValueError("no p grid: pass p_array or a raw dict that carries one")
    diag = dict(raw.get("diag") or {})
    off = raw.get("off_diag")
    has_off = off is not None and (not isinstance(off, dict) or len(off) > 0)
    if (diag or has_off) and stored is None:
        raise ValueError("the operator series are stored per p of the run; raw must carry "
                         "its p_array")
    m_arr = raw.get("m_array")
    m_arr = [p // 2 for p in stored] if m_arr is None and stored is not None else m_arr
    vals = norms = steps = names = None
    if has_off:
        vals, names = _dense(off, raw.get("off_diag_operators"), "off_diag")
        if raw.get("off_diag_norm") is None:
            raise ValueError("off_diag_norm is required whenever off_diag is supplied.")
        if raw.get("off_diag_steps") is None:
            raise ValueError("off_diag_steps is required whenever off_diag is supplied.")
        norms, _ = _dense(raw["off_diag_norm"], names, "off_diag_norm")
        steps = np.asarray(raw["off_diag_steps"], dtype=np.int64)
        if steps.shape != (vals.shape[0],) or (steps.size and (
                np.any(np.diff(steps) <= 0) or steps[0] < 0 or steps[-1] >= len(E))):
            raise ValueError("off_diag_steps must be strictly increasing steps in "
                             f"[0, {len(E)}), one per row.")
    m_common = max(m_arr) if (diag and has_off) else 0

    def obs_of(lw_, diag_, vals_, norms_, steps_):
        o = {}
        if diag_:
            o.update(estimate_diagonal(stored, lw_, diag_, m_common))
        if vals_ is not None:
            o.update(estimate_off_diagonal(stored, lw_, vals_, norms_, steps_, m_arr,
                                           m_common, names))
        return o

    if n_blocks is None:
        return estimate_energy(grid, lw, E), obs_of(lw, diag, vals, norms, steps)
    n_blocks = validated_int("n_blocks", n_blocks, 1)
    n = E.shape[0]
    blen = n // n_blocks
    if n_blocks < 2 or blen < 1:
        raise ValueError(f"n_blocks={n_blocks} over {n} steps leaves {blen} steps per block; "
                         "a jackknife needs at least 2 usable blocks.")
    usable = blen * n_blocks
    if usable != n:
        warnings.warn(f"Discarding {n - usable} trailing samples for equal jackknife blocks; "
                      "point estimates and errors both use the retained samples.", stacklevel=2)
        E, lw = E[:usable], lw[:usable]
        diag = {k: np.asarray(v)[:usable] for k, v in diag.items()}
        if vals is not None:
            keep = steps < usable
            vals, norms, steps = vals[keep], norms[keep], steps[keep]
        n = usable
    if n_blocks < 20:
        warnings.warn(f"n_blocks={n_blocks} is few: the error on the error is then "
                      "comparable to the error itself. ~20 or more is advisable.", stacklevel=2)
    est = estimate_energy(grid, lw, E)
    obs = obs_of(lw, diag, vals, norms, steps)
    reps, obs_reps = [], {name: {p: [] for p in stored} for name in obs}
    for b in range(n_blocks):
        m = np.ones(n, dtype=bool)
        m[b * blen:(b + 1) * blen] = False
        e = estimate_energy(grid, lw[m], E[m])
        if vals is not None:
            new_pos = np.cumsum(m) - 1
            rows = m[steps]
            o = obs_of(lw[m], {k: np.asarray(v)[m] for k, v in diag.items()}, vals[rows],
                       norms[rows], new_pos[steps[rows]])
        else:
            o = obs_of(lw[m], {k: np.asarray(v)[m] for k, v in diag.items()}, None, None,
                       None)
        reps.append([e[p] for p in grid])
        for name in o:
            for p in stored:
                obs_reps[name][p].append(np.asarray(o[name][p], dtype=float))

    def jk(x):
        x = np.asarray(x, dtype=float)
        return np.sqrt((n_blocks - 1) / n_blocks * np.sum((x - x.mean(axis=0)) ** 2, axis=0))

    err = jk(reps)
    obs_err = {name: {p: (float(v) if np.ndim(v) == 0 else v)
                      for p, v in ((p, jk(obs_reps[name][p])) for p in stored)}
               for name in obs_reps}
    return est, obs, dict(zip(grid, (float(x) for x in err))), obs_err


def check_population(log_G_chunk, step0, E_chunk=None):
    """Raise on a non-finite log G_n (extinct population) or E_n
    (driver.py:913-937)."""
    bad = np.flatnonzero(~np.isfinite(log_G_chunk))
    if bad.size:
        raise ValueError(
            f"[gfmc] population extinct at step {step0 + int(bad[0])}: the mean branch "
            "weight of the step is not finite, i.e. every walker's weight is zero (or one "
            "overflowed).  The ensemble cannot be renormalised; check lam against the "
            "diagonal elements (discrete methods), the step size and the guide.")
    if E_chunk is not None:
        bad = np.flatnonzero(~np.isfinite(E_chunk))
        if bad.size:
            raise ValueError(
                f"[gfmc] non-finite energy at step {step0 + int(bad[0])}: the weighted mean "
                "of the local energies is not finite while the branch weights are, i.e. a "
                "walker's local energy overflowed (a guide ratio or a rate of its "
                "configuration is inf or NaN); check the guide and the Hamiltonian on the "
                "configurations the walk reaches.")


class DeviceEnergyEstimator:
    """The windowed energy estimator of :func:`estimate_energy`, streamed on
    the GPU over the device series (gfmc_fw_update): a ring buffer of the
    last max(p) + 1 cumulative log normalisation factors and, per p, a
    running maximum with rescaled sums.  Agrees with the host formula to
    rounding (the maximum is running instead of global)."""

    def __init__(self, p_array, device):
        import torch

        from . import _lib
        self.p = validated_p_array(p_array)
        if len(self.p) > 64:
            raise ValueError("at most 64 p values")
        self.p_max = int(self.p[-1])
        n = int(_lib.load().gfmc_fw_state_doubles(len(self.p), self.p_max))
        self.state = torch.empty(n, dtype=torch.float64, device=device)
        self.p_dev = torch.as_tensor(np.asarray(self.p, dtype=np.int32), device=device)
        _lib.check(_lib.load().gfmc_fw_init(_lib.ptr(self.state), len(self.p), self.p_max,
                                            _lib.stream()), "gfmc_fw_init")

    def update(self, rows):
        """Consume (n, 2) device rows (log G - log n, E) of consecutive steps."""
        from . import _lib
        if rows.shape[0] == 0:
            return
        _lib.check(_lib.load().gfmc_fw_update(_lib.ptr(self.p_dev), len(self.p), self.p_max,
                                              _lib.ptr(rows), int(rows.shape[0]),
                                              _lib.ptr(self.state), _lib.stream()),
                   "gfmc_fw_update")

    def result(self):
        """{p: E_p} (host); raises ValueError when a non-finite value was
        consumed."""
        st = self.state.cpu().numpy()
        if st[2] != 0.0:
            raise ValueError("Nonfinite log weights or energies in the streamed series.")
        n_p = len(self.p)
        num, den = st[-2 * n_p:-n_p], st[-n_p:]
        return {p: float(num[j] / den[j]) for j, p in enumerate(self.p)}

