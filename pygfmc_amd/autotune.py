"""Autotuner of the propagation kernel: lanes per walker (64, 32 or 16,
i.e. one, two or four walkers per wavefront).

Each usable value is timed on copies of the caller's walkers (one untimed
launch, then the best of ``reps`` timed launches, HIP events on the launch
stream); the smallest time wins, with a 5 % margin in favour of more lanes
per walker so that near-ties resolve the same way every time.  Decisions are
kept per (device, problem, walker-count bucket) in memory and in a JSON file
(``$PYGFMC_AMD_CACHE``, default ``~/.cache/pygfmc_amd/autotune.json``; set
``PYGFMC_AMD_AUTOTUNE_CACHE=0`` to keep them in memory only), so later runs
on the same machine reuse them.  The choice changes the summation order of
the off-diagonal weights (blocked over the group's lanes): pin ``lanes`` for
bit-reproducible runs.  No reference counterpart (the reference's vmapped
XLA loop has no launch shape to tune).
"""
from __future__ import annotations

import hashlib
import json
import os

import torch

from . import _lib

__all__ = ["tune_problem", "tune_batch", "clear_cache"]

_MEM = {}
MARGIN = 0.05


def _cache_path():
    return os.environ.get("PYGFMC_AMD_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pygfmc_amd",
                                       "autotune.json"))


def _use_file():
    return os.environ.get("PYGFMC_AMD_AUTOTUNE_CACHE", "1") != "0"


def _load_file():
    if not _use_file():
        return {}
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    _MEM[key] = value
    if not _use_file():
        return
    data = _load_file()
    data[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def clear_cache():
    _MEM.clear()


def _bucket(n):
    return 1 << max(0, int(n) - 1).bit_length()


def _key(kind, fingerprint, n_walkers, extra=""):
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    h = hashlib.sha1(fingerprint).hexdigest()[:16]
    return f"{kind}|{dev}|{h}|{_bucket(n_walkers)}|{extra}"


def _pick(times):
    """times: {lanes: seconds}; the fastest, preferring more lanes within
    MARGIN."""
    best = None
    for g in sorted(times, reverse=True):
        if best is None or times[g] < times[best] * (1.0 - MARGIN):
            best = g
    return best


def _time(launch, reps):
    launch()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        launch()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return min(ts)


def tune_problem(problem, spins, beta, alw, reps=3, seed=0x5EED):
    """Pick and set the lanes per walker of ``problem`` for the walkers
    ``spins`` (device, packed); returns (lanes, {lanes: seconds} or None
    when cached)."""
    cands = problem.lanes_ok
    if len(cands) == 1:
        problem.set_lanes(cands[0])
        return cands[0], None
    fp = problem.lowered_fingerprint()
    key = _key("problem", fp, spins.shape[0], str(problem.precision))
    hit = _MEM.get(key) or _load_file().get(key)
    if hit in cands:
        problem.set_lanes(hit)
        _MEM[key] = hit
        return hit, None
    nw = spins.shape[0]
    times = {}
    for g in cands:
        problem.set_lanes(g)
        s = spins.clone()
        lw = torch.zeros(nw, dtype=torch.float64, device=spins.device)
        e = torch.zeros_like(lw)
        times[g] = _time(lambda: problem.propagate(s, lw, e, beta, alw, seed, 0,
                                                   _lib.TAG_WARMUP_PROPAGATE), reps)
    best = _pick(times)
    problem.set_lanes(best)
    _store(key, best)
    return best, times


def tune_batch(chain, beta, alw, reps=3, seed=0x5EED):
    """Pick and set the lanes per walker of a batch chain's batch (timed on
    copies of its buffers)."""
    b = chain.b
    cands = b.lanes_ok
    if len(cands) == 1:
        b.set_lanes(cands[0])
        return cands[0], None
    h = hashlib.sha1()
    for p in b.problems:
        h.update(p.lowered_fingerprint())
    h.update(repr(b.n_walkers).encode())
    key = _key("batch", h.digest(), b.n_total, str(b.problems[0].precision))
    hit = _MEM.get(key) or _load_file().get(key)
    if hit in cands:
        b.set_lanes(hit)
        _MEM[key] = hit
        return hit, None
    lib = _lib.load()
    times = {}
    for g in cands:
        b.set_lanes(g)
        s = chain.spins.clone()
        lw = torch.zeros_like(chain.log_w)
        e = torch.zeros_like(chain.e_loc)

        def launch():
            _lib.check(lib.gfmc_batch_propagate(b.handle, _lib.ptr(s), _lib.ptr(lw), _lib.ptr(e),
                                                float(beta), _lib.ptr(alw), seed, 0,
                                                _lib.TAG_WARMUP_PROPAGATE, None, _lib.stream()),
                       "gfmc_batch_propagate")
        times[g] = _time(launch, reps)
    best = _pick(times)
    b.set_lanes(best)
    _store(key, best)
    return best, times
