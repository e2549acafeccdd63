"""Many independent lattices of uneven size in one launch.

``run_batch`` runs GFMC on a list of (Hamiltonian, guide, walkers) triples
with one propagation launch, one segmented reconfiguration and one gather
per step for the whole batch (gfmc_batch_* in include/gfmc.h).  The
reference has no batch: it runs one Hamiltonian per ``run_gfmc`` call
(driver.py:1148-1547), so a batch replaces a loop of runs.  Lattice ``l``
follows exactly the single-lattice run with ``rng_key = (seed + l,
position)``; warm-up and estimators are those of ``run_gfmc`` applied per
lattice.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .engine import (METHODS, StepGraph, require_lam, require_method, require_step_size,
                     warmup_schedule)
from .estimators import check_population, estimate_energy, validated_int, validated_p_array
from .problem import Problem, precision_of
from .walkers import Walker, as_key

__all__ = ["Batch", "BatchChain", "run_batch"]


class Batch:
    """Device batch of lattices; keeps its problems alive."""

    def __init__(self, hamiltonians, guides, n_walkers, precision=None, method="cont_time",
                 lam=None):
        if len(hamiltonians) != len(guides) or len(hamiltonians) != len(n_walkers):
            raise ValueError("hamiltonians, guides and n_walkers must have equal lengths")
        if len(hamiltonians) == 0:
            raise ValueError("a batch needs at least one lattice")
        require_method(method)
        prec = precision_of(hamiltonians[0], precision)
        self.problems = [Problem(H, G, prec) for H, G in zip(hamiltonians, guides)]
        lams = lam if np.ndim(lam) else [lam] * len(self.problems)
        for p, lv in zip(self.problems, lams):
            lv = require_lam(method, lv)
            _lib.check(_lib.load().gfmc_problem_set_method(p.handle, METHODS[method],
                                                           0.0 if lv is None else lv),
                       "gfmc_problem_set_method")
        self.n_lat = len(self.problems)
        self.n_walkers = [int(n) for n in n_walkers]
        lib = _lib.load()
        handles = (ctypes.c_void_p * self.n_lat)(*[p.handle.value for p in self.problems])
        nws = (ctypes.c_int64 * self.n_lat)(*self.n_walkers)
        h = ctypes.c_void_p()
        _lib.check(lib.gfmc_batch_create(handles, self.n_lat, nws, ctypes.byref(h)),
                   "gfmc_batch_create")
        self.handle = h
        nt, st, wd = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        _lib.check(lib.gfmc_batch_get_info(h, ctypes.byref(nt), ctypes.byref(st),
                                           ctypes.byref(wd)), "gfmc_batch_get_info")
        self.n_total, self.stride, self.work_doubles = nt.value, st.value, wd.value
        self.offsets = np.concatenate([[0], np.cumsum(self.n_walkers)]).astype(np.int64)
        self.lanes = 64

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().gfmc_batch_destroy(self.handle)
        except Exception:
            pass

    @property
    def lanes_ok(self):
        m = int(_lib.load().gfmc_batch_lanes_ok(self.handle))
        return [g for g in (64, 32, 16) if m & g]

    def set_lanes(self, lanes):
        _lib.check(_lib.load().gfmc_batch_set_lanes(self.handle, int(lanes)),
                   "gfmc_batch_set_lanes")
        self.lanes = int(lanes)


class BatchChain:
    """Device buffers of a batch's walker populations."""

    def __init__(self, batch, walkers_list):
        if len(walkers_list) != batch.n_lat:
            raise ValueError("one Walker population per lattice")
        self.b = batch
        dev = _lib.device()
        n, W = batch.n_total, batch.stride
        self.spins = torch.zeros((n, W), dtype=torch.int32, device=dev)
        self.log_w = torch.zeros(n, dtype=torch.float64, device=dev)
        self.e_loc = torch.zeros(n, dtype=torch.float64, device=dev)
        for l, w in enumerate(walkers_list):
            p = batch.problems[l]
            if w.n_sites != p.n_sites or w.n_walkers != batch.n_walkers[l]:
                raise ValueError(f"lattice {l}: walkers do not match the batch "
                                 f"({w.n_walkers} x {w.n_sites} vs {batch.n_walkers[l]} x "
                                 f"{p.n_sites})")
            a, z = batch.offsets[l], batch.offsets[l + 1]
            self.spins[a:z, :w.spins.shape[1]] = w.spins
            self.log_w[a:z] = w.log_weight_dev
            self.e_loc[a:z] = w.e_loc_dev
        self.spins_b = torch.empty_like(self.spins)
        self.e_loc_b = torch.empty_like(self.e_loc)
        self.parents = torch.empty(n, dtype=torch.int32, device=dev)
        self.work = torch.empty(max(1, batch.work_doubles), dtype=torch.float64, device=dev)
        self.evals = torch.zeros(n, dtype=torch.int32, device=dev)

    def steps(self, n, beta, alw, seed, step0, warmup=False, series=None, evals=False):
        """Enqueue ``n`` steps; ``alw`` a device [n_lat] tensor; returns the
        (n, n_lat, 2) device series."""
        if series is None:
            series = torch.empty((n, self.b.n_lat, 2), dtype=torch.float64,
                                 device=self.spins.device)
        tp = _lib.TAG_WARMUP_PROPAGATE if warmup else _lib.TAG_PROPAGATE
        tr = _lib.TAG_WARMUP_RECONFIG if warmup else _lib.TAG_RECONFIG
        _lib.check(_lib.load().gfmc_batch_run_steps(
            self.b.handle, _lib.ptr(self.spins), _lib.ptr(self.log_w), _lib.ptr(self.e_loc),
            _lib.ptr(self.spins_b), _lib.ptr(self.e_loc_b), _lib.ptr(self.parents),
            _lib.ptr(self.work), int(n), float(beta), _lib.ptr(alw),
            int(seed) & 0xFFFFFFFFFFFFFFFF, int(step0) & 0xFFFFFFFF, tp, tr, _lib.ptr(series),
            _lib.ptr(self.evals) if evals else None, _lib.stream()), "gfmc_batch_run_steps")
        return series

    def walkers(self, l):
        a, z = self.b.offsets[l], self.b.offsets[l + 1]
        p = self.b.problems[l]
        return Walker.from_device(self.spins[a:z, :p.n_words].contiguous(),
                                  self.log_w[a:z].clone(), self.e_loc[a:z].clone(), p.n_sites)


def _lse(v):
    m = np.max(v)
    return m + np.log(np.sum(np.exp(v - m)))


def run_batch(walkers_list, n_steps, hamiltonians, step_size, guides, p_array, *,
              method="cont_time", rng_key=0, warmup_steps=1000, precision=None, verbosity=0,
              number_of_outputs=10, use_graph=True, lanes="auto", lam=None):
    """GFMC on every lattice of the batch.  Returns ``(energies, raws,
    stats)``: ``energies[l]`` {p: E_p}, ``raws[l]`` with ``log_weights`` and
    ``energy`` series, ``stats`` with the batch throughput.  ``lanes``: lanes
    per walker (64, 32, 16) or "auto" (pygfmc_amd.autotune)."""
    require_method(method)
    require_step_size(method, step_size)
    n_steps = validated_int("n_steps", n_steps, 1)
    warmup_steps = validated_int("warmup_steps", warmup_steps, 1)
    p_sorted = validated_p_array(p_array)
    if p_sorted[-1] >= n_steps:
        raise ValueError(f"every p value must be smaller than n_steps: got n_steps={n_steps} "
                         f"and max(p)={p_sorted[-1]}")
    key = as_key(rng_key)
    b = Batch(hamiltonians, guides, [w.n_walkers for w in walkers_list], precision, method,
              lam)
    lams = lam if np.ndim(lam) else [lam] * b.n_lat
    chain = BatchChain(b, walkers_list)
    dev = chain.spins.device
    tune_times = None
    if lanes == "auto":
        from .autotune import tune_batch
        _, tune_times = tune_batch(chain, float(step_size),
                                   torch.zeros(b.n_lat, dtype=torch.float64, device=dev))
    else:
        b.set_lanes(int(lanes))
    alw = np.empty(b.n_lat)
    for l in range(b.n_lat):
        a, z = b.offsets[l], b.offsets[l + 1]
        p = b.problems[l]
        _, e, _ = p.rates(chain.spins[a:z, :p.n_words].contiguous(), want_rates=False)
        alw[l] = float(torch.mean(-e).item() if method == "cont_time" else
                       torch.mean(torch.log(float(lams[l]) - e)).item())
    step_f = float(step_size)
    counter = key.position
    for step_, n_w in warmup_schedule(step_f, warmup_steps, method):
        series = chain.steps(n_w, step_, torch.as_tensor(alw, device=dev), key.seed, counter,
                             warmup=True)
        counter += n_w
        lg = series[:, :, 0].cpu().numpy()
        for l in range(b.n_lat):
            check_population(lg[:, l], 0)
            alw[l] = alw[l] + (_lse(lg[:, l]) - np.log(n_w)) / step_
    alw_dev = torch.as_tensor(alw, device=dev)
    series = torch.empty((n_steps, b.n_lat, 2), dtype=torch.float64, device=dev)
    chunk = max(1, n_steps // max(1, number_of_outputs))
    torch.cuda.synchronize()
    t0 = time.time()
    s0 = 0
    graph = None
    while s0 < n_steps:
        n = min(chunk, n_steps - s0)
        if use_graph and n == chunk and n_steps >= 2 * chunk:
            if graph is None:
                graph = StepGraph(chain, n, step_f, alw_dev, key.seed, evals=True)
            graph.launch(key.position + s0, out=series[s0:s0 + n])
        else:
            chain.steps(n, step_f, alw_dev, key.seed, key.position + s0,
                        series=series[s0:s0 + n], evals=True)
        s0 += n
    torch.cuda.synchronize()
    runtime = time.time() - t0
    arr = series.cpu().numpy()
    energies, raws = [], []
    for l in range(b.n_lat):
        lw, E = arr[:, l, 0].copy(), arr[:, l, 1].copy()
        check_population(lw, 0, E)
        energies.append(estimate_energy(p_sorted, lw, E))
        raws.append(dict(log_weights=lw, energy=E, p_array=list(p_sorted),
                         m_array=[q // 2 for q in p_sorted]))
    evals = chain.evals.to(torch.float64)
    stats = dict(runtime=runtime, it_per_sec=n_steps / runtime,
                 walker_steps_per_sec=b.n_total * n_steps / runtime,
                 evals_per_step_mean=float(evals.mean().item()) / n_steps,
                 avg_log_weight=alw.copy(), n_walkers_total=b.n_total, n_lattices=b.n_lat,
                 hip_graph=graph is not None, lanes_per_walker=b.lanes,
                 autotune_seconds=tune_times)
    if verbosity:
        print(f"[gfmc] - [run_batch] - {b.n_lat} lattices, {n_steps} steps @ "
              f"{stats['it_per_sec']:.1f} it/s", flush=True)
    return energies, raws, stats
