"""Windowed energy estimator and its blocked jackknife (host, FP64 NumPy).

The per-step series (log G_n, E_n) come off the GPU once per chunk; the
estimator and the jackknife are O(n_steps) host work, as in the reference
(estimators.py:268-305, io.py:921-1157).
"""
from __future__ import annotations

import warnings

import numpy as np

__all__ = ["validated_p_array", "estimate_energy", "estimate_diagonal", "evaluate_observables",
           "DeviceEnergyEstimator", "validated_int", "check_population"]


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


def evaluate_observables(raw, p_array=None, n_blocks=None):
    """Re-estimate the energy from ``raw`` (``log_weights``, ``energy``,
    ``p_array``), optionally with the blocked jackknife error: delete each of
    ``n_blocks`` contiguous blocks, recompute the whole estimator on the rest,
    err = sqrt((nb-1)/nb * sum (x_b - mean)^2) (io.py:921-1157).  Trailing
    samples that do not fill a block are dropped from both the estimate and
    the error.  Returns ``(energy, obs)`` or ``(energy, obs, energy_err,
    obs_err)``; ``obs`` holds the diagonal operators of ``raw['diag']``
    (empty without them)."""
    if raw.get("energy") is None:
        raise ValueError("raw holds no 'energy' series, so nothing can be re-estimated.")
    if raw.get("log_weights") is None:
        raise ValueError("raw holds no 'log_weights' series; the windowed estimator "
                         "cannot be rebuilt without the per-step branch weights.")
    E = np.asarray(raw["energy"], dtype=np.float64)
    lw = np.asarray(raw["log_weights"], dtype=np.float64)
    grid = p_array if p_array is not None else raw.get("p_array")
    if grid is None:
        raise ValueError("no p grid: pass p_array or a raw dict that carries one")
    grid = validated_p_array(grid)
    diag = raw.get("diag") or {}
    if diag and p_array is not None and list(grid) != list(validated_p_array(raw["p_array"])):
        raise ValueError("the diagonal series are stored per p of the run; re-estimate them "
                         "on the run's own p_array")

    def estimate(lw_, E_, diag_):
        e = estimate_energy(grid, lw_, E_)
        return e, (estimate_diagonal(grid, lw_, diag_) if diag_ else {})

    if n_blocks is None:
        return estimate(lw, E, diag)
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
        n = usable
    if n_blocks < 20:
        warnings.warn(f"n_blocks={n_blocks} is few: the error on the error is then "
                      "comparable to the error itself. ~20 or more is advisable.", stacklevel=2)
    est, obs = estimate(lw, E, diag)
    reps, obs_reps = [], {name: {p: [] for p in grid} for name in obs}
    for b in range(n_blocks):
        m = np.ones(n, dtype=bool)
        m[b * blen:(b + 1) * blen] = False
        e, o = estimate(lw[m], E[m], {k: np.asarray(v)[m] for k, v in diag.items()})
        reps.append([e[p] for p in grid])
        for name in o:
            for p in grid:
                obs_reps[name][p].append(np.asarray(o[name][p], dtype=float))

    def jk(x):
        x = np.asarray(x, dtype=float)
        return np.sqrt((n_blocks - 1) / n_blocks * np.sum((x - x.mean(axis=0)) ** 2, axis=0))

    err = jk(reps)
    obs_err = {name: {p: (float(v) if np.ndim(v) == 0 else v)
                      for p, v in ((p, jk(obs_reps[name][p])) for p in grid)}
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
            "overflowed).  The ensemble cannot be renormalised; check the step size and "
            "the guide.")
    if E_chunk is not None:
        bad = np.flatnonzero(~np.isfinite(E_chunk))
        if bad.size:
            raise ValueError(
                f"[gfmc] non-finite energy at step {step0 + int(bad[0])}: the weighted mean "
                "of the local energies is not finite while the branch weights are, i.e. a "
                "walker's local energy overflowed; check the guide and the Hamiltonian on "
                "the configurations the walk reaches.")


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

