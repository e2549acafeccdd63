"""Autotuner of the propagation kernel's launch shape: lanes per walker
(64, 32 or 16, i.e. one, two or four walkers per wavefront), wavefronts per
workgroup (1, 2 or 4 on the record paths) and the XCD-contiguous mapping of
workgroups to walkers (gfmc_problem_set_launch).

Every usable combination is timed on a scratch copy of the caller's walkers
with carried rows, as production steps run: the copy is restored, one
untimed branching step fills the rows, then the best of ``reps``
propagation launches (HIP events on the launch stream) counts; every
combination replays the same steps.  The smallest time wins, with a 5 %
margin in favour of the earlier combination in the order (more lanes, fewer
wavefronts per workgroup, no remapping first) so that near-ties resolve the
same way every time.  Decisions are kept per (device, problem, walker-count
bucket) in memory and in a JSON file (``$PYGFMC_AMD_CACHE``, default
``~/.cache/pygfmc_amd/autotune.json``; set ``PYGFMC_AMD_AUTOTUNE_CACHE=0`` to
keep them in memory only), so later runs on the same machine reuse them.
The lanes change the summation order of the off-diagonal weights (blocked
over the group's lanes): pin ``lanes`` for bit-reproducible runs; the other
two do not change any result.  No reference counterpart (the reference's
vmapped XLA loop has no launch shape to tune).
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


def _pick(times, order=None):
    """times: {candidate: seconds}; the fastest, preferring a candidate
    earlier in ``order`` (default: more lanes first) within MARGIN."""
    best = None
    for g in (order if order is not None else sorted(times, reverse=True)):
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


def _candidates(problem):
    return [(g, w, x) for g in problem.lanes_ok for w in problem.wpb_ok for x in (0, 1)]


def _apply(problem, c):
    problem.set_lanes(c[0])
    problem.set_launch(c[1], c[2])


def tune_problem(problem, spins, beta, alw, reps=3, seed=0x5EED):
    """Pick and set the launch shape of ``problem`` for the walkers
    ``spins`` (device, packed); returns ((lanes, waves per workgroup, XCD
    remap), {"lanes/waves/remap": seconds} or None when cached or when
    there is one candidate)."""
    cands = _candidates(problem)
    if len(cands) == 1:
        _apply(problem, cands[0])
        return cands[0], None
    fp = problem.lowered_fingerprint()
    key = _key("launch", fp, spins.shape[0], str(problem.precision))
    hit = _MEM.get(key) or _load_file().get(key)
    if hit is not None and tuple(hit) in cands:
        _apply(problem, tuple(hit))
        _MEM[key] = list(hit)
        return tuple(hit), None
    from .engine import Chain
    from .walkers import Walker
    nw, n_sites = spins.shape[0], problem.n_sites
    zeros = torch.zeros(nw, dtype=torch.float64, device=spins.device)
    ch = Chain(problem, Walker.from_device(spins, zeros, zeros.clone(), n_sites))
    s0, lw0, e0 = ch.spins.clone(), ch.log_w.clone(), ch.e_loc.clone()
    out2 = torch.empty(2, dtype=torch.float64, device=spins.device)
    times = {}
    for c in cands:
        _apply(problem, c)
        ch.spins.copy_(s0)
        ch.log_w.copy_(lw0)
        ch.e_loc.copy_(e0)
        ch.drop_carry()
        ch.step(beta, alw, seed, 0, out2, evals=False)
        ts = []
        for r in range(reps):
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ch.step(beta, alw, seed, 1 + r, out2, events=ev, evals=False)
            ev[1].synchronize()
            ts.append(ev[0].elapsed_time(ev[1]) * 1e-3)
        times[c] = min(ts)
    del ch
    best = _pick(times, cands)
    _apply(problem, best)
    _store(key, list(best))
    return best, {f"{g}/{w}/{x}": t for (g, w, x), t in times.items()}


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
