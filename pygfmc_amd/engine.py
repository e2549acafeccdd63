"""The GFMC driver: ``run`` / ``run_gfmc``, ``evolve_walkers``,
``reconfiguration`` (codes/pygfmc/engine/driver.py, evolution.py,
walkers/reconfiguration.py), driving the HIP kernels.

Methods: continuous time (``"cont_time"``, evolution.py:559-721) and the
discrete ``"single_step"`` / ``"multi_step"`` propagators with the shift
``lam`` (evolution.py:322-556).  Diagonal operators are measured by forward
walking on the device (driver.py:240-288), off-diagonal operators by
shadow chains (driver.py:290-331, evolution.py:848-1101; ``offdiagfw``).
"""
from __future__ import annotations

import ctypes
import hashlib
import math
import time

import numpy as np
import torch

from . import _lib
from .estimators import (check_population, estimate_diagonal, estimate_energy,
                         estimate_off_diagonal, validated_int,
                         validated_p_array)
from .problem import get_problem
from .trace import span
from .walkers import PhiloxKey, Walker, as_key

__all__ = ["run", "run_gfmc", "evolve_walkers", "reconfiguration", "systematic_parents",
           "sample_index", "estimate_average_log_branch_weight", "DEFAULT_WARMUP_STEPS",
           "Chain", "StepGraph", "set_method"]

DEFAULT_WARMUP_STEPS = 1000


def require_step_size(method, step_size):
    """evolution.py:744-758."""
    if not math.isfinite(float(step_size)):
        raise ValueError(f"step_size must be finite; got step_size={step_size!r}")
    if not float(step_size) > 0:
        raise ValueError(f"step_size must be > 0; got step_size={step_size!r}")
    if method != "cont_time" and float(int(step_size)) != float(step_size):
        raise ValueError(f"method={method!r} requires an integer step_size; "
                         f"got step_size={step_size!r}")


METHODS = {"cont_time": 0, "single_step": 1, "multi_step": 2}


def require_method(method):
    if method not in METHODS:
        raise ValueError(f"Unknown method: {method!r}")


def require_lam(method, lam):
    """evolution.py:724-741: a finite float for the discrete methods, None
    for cont_time."""
    if method == "cont_time":
        return None
    if lam is None:
        raise ValueError(
            f"method={method!r} needs the GFMC shift: pass lam=<float> with lam >= max_x "
            f"H_xx over the reachable configurations (cont_time is the only method that "
            f"does without it).")
    value = float(lam)
    if not math.isfinite(value):
        raise ValueError(f"lam must be finite; got lam={lam!r}")
    return value


def set_method(problem, method, lam):
    """Select the propagator of ``problem``'s kernels (and its shift)."""
    lam = require_lam(method, lam)
    _lib.check(_lib.load().gfmc_problem_set_method(problem.handle, METHODS[method],
                                                   0.0 if lam is None else lam),
               "gfmc_problem_set_method")
    return lam


def warmup_schedule(step_f, warmup_steps, method="cont_time"):
    """(step size, count) of the four warm-up ramp stages with counts
    divmod(warmup_steps, 4): beta/8, beta/4, beta/2, beta, or for the
    discrete methods max(1, n_bra // 8), ... , n_bra (driver.py:567-584)."""
    if warmup_steps <= 0:
        return []
    base, extra = divmod(warmup_steps, 4)
    counts = [base + (i < extra) for i in range(4)]
    if method == "cont_time":
        sizes = [step_f / 8.0, step_f / 4.0, step_f / 2.0, step_f]
    else:
        n = int(step_f)
        sizes = [float(max(1, n // 8)), float(max(1, n // 4)), float(max(1, n // 2)),
                 float(n)]
    return [(s, c) for s, c in zip(sizes, counts) if c > 0]


class Chain:
    """Device buffers of a walker population for repeated branching steps."""

    def __init__(self, problem, walkers):
        if walkers.n_sites != problem.n_sites:
            raise ValueError(f"walkers have {walkers.n_sites} sites, the Hamiltonian "
                             f"{problem.n_sites}")
        self.p = problem
        self.nw = walkers.n_walkers
        problem.reserve(self.nw)
        dev = walkers.spins.device
        self.spins = walkers.spins.clone()
        self.log_w = walkers.log_weight_dev.clone()
        self.e_loc = walkers.e_loc_dev.clone()
        self.spins_b = torch.empty_like(self.spins)
        self.e_loc_b = torch.empty_like(self.e_loc)
        self.parents = torch.empty(self.nw, dtype=torch.int32, device=dev)
        self.work = torch.empty(_lib.work_doubles(self.nw), dtype=torch.float64, device=dev)
        self.evals = torch.zeros(self.nw, dtype=torch.int32, device=dev)
        # carried slot weights of the record path (gfmc_carry_bytes; none for
        # the other paths): zero = nothing carried yet
        nb = int(_lib.load().gfmc_carry_bytes(problem.handle, self.nw))
        self._carry = torch.zeros(nb + 256, dtype=torch.uint8, device=dev) if nb else None
        self.carry = None
        if nb:
            off = (-self._carry.data_ptr()) % 256
            self.carry = self._carry[off:off + nb]

    def drop_carry(self):
        """Forget the carried weights (the walkers changed outside the step
        loops); the next step re-evaluates every slot."""
        if self.carry is not None:
            self.carry[:256].zero_()

    def steps(self, n, beta, alw, seed, step0, warmup=False, series=None, evals=False,
              diag=None, dseries=None):
        """Enqueue ``n`` branching steps; returns the (n, 2) device series.
        With ``diag`` (a :class:`~pygfmc_amd.diagfw.DiagRecorder`) every step
        also appends its forward-walking row to ``dseries`` (n, n_cols,
        n_p)."""
        if series is None:
            series = torch.empty((n, 2), dtype=torch.float64, device=self.spins.device)
        tp = _lib.TAG_WARMUP_PROPAGATE if warmup else _lib.TAG_PROPAGATE
        tr = _lib.TAG_WARMUP_RECONFIG if warmup else _lib.TAG_RECONFIG
        common = (self.p.handle, self.nw, _lib.ptr(self.spins), _lib.ptr(self.log_w),
                  _lib.ptr(self.e_loc), _lib.ptr(self.spins_b), _lib.ptr(self.e_loc_b),
                  _lib.ptr(self.parents), _lib.ptr(self.work), int(n), float(beta), float(alw),
                  int(seed) & 0xFFFFFFFFFFFFFFFF, int(step0) & 0xFFFFFFFF, tp, tr,
                  _lib.ptr(series), _lib.ptr(self.evals) if evals else None)
        if diag is not None and dseries is None:
            dseries = diag.series(n, self.spins.device)
        dargs = ((diag.handle, _lib.ptr(diag.ring), _lib.ptr(dseries), _lib.ptr(diag.work))
                 if diag is not None else (None, None, None, None))
        _lib.check(_lib.load().gfmc_run_steps_carry(*common, *dargs, _lib.ptr(self.carry),
                                                    _lib.stream()), "gfmc_run_steps_carry")
        return series

    def step(self, beta, alw, seed, step, out2, events=None, evals=True):
        """One production step as three separate launches (propagate with
        the chain's carried weights, reconfigure, gather) on torch's current
        stream; ``events`` (start, end) bracket the propagation kernel.
        ``out2``: 2-double device row."""
        lib = _lib.load()
        st = _lib.stream()
        if events is not None:
            events[0].record()
        _lib.check(lib.gfmc_propagate_carry(
            self.p.handle, self.nw, _lib.ptr(self.spins), _lib.ptr(self.log_w),
            _lib.ptr(self.e_loc), float(beta), float(alw), int(seed) & 0xFFFFFFFFFFFFFFFF,
            int(step) & 0xFFFFFFFF, _lib.TAG_PROPAGATE, _lib.ptr(self.evals) if evals else None,
            _lib.ptr(self.parents), _lib.ptr(self.carry), st), "gfmc_propagate_carry")
        if events is not None:
            events[1].record()
        _lib.check(lib.gfmc_reconfigure(self.nw, _lib.ptr(self.log_w), _lib.ptr(self.e_loc),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, int(step) & 0xFFFFFFFF,
                                        _lib.TAG_RECONFIG, None, _lib.ptr(self.parents),
                                        _lib.ptr(out2), _lib.ptr(self.work), st),
                   "gfmc_reconfigure")
        _lib.check(lib.gfmc_gather(self.p.n_sites, self.nw, _lib.ptr(self.spins),
                                   _lib.ptr(self.e_loc), _lib.ptr(self.parents),
                                   _lib.ptr(self.spins_b), _lib.ptr(self.e_loc_b),
                                   _lib.ptr(self.log_w), st), "gfmc_gather")
        self.spins, self.spins_b = self.spins_b, self.spins
        self.e_loc, self.e_loc_b = self.e_loc_b, self.e_loc

    def walkers(self):
        return Walker.from_device(self.spins.clone(), self.log_w.clone(), self.e_loc.clone(),
                                  self.p.n_sites)


class StepGraph:
    """``n`` branching steps of a chain (or batch chain) captured once into
    a HIP graph (gfmc_graph_capture_steps) and replayed by :meth:`launch`;
    each launch runs positions ``step0 .. step0 + n - 1`` into ``series``."""

    def __init__(self, chain, n, beta, alw, seed, evals=False, warmup=False, diag=None):
        lib = _lib.load()
        dev = chain.spins.device
        self.n = int(n)
        self.diag = diag
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        batch = hasattr(chain, "b")
        shape = (self.n, chain.b.n_lat, 2) if batch else (self.n, 2)
        self.series = torch.empty(shape, dtype=torch.float64, device=dev)
        tp = _lib.TAG_WARMUP_PROPAGATE if warmup else _lib.TAG_PROPAGATE
        tr = _lib.TAG_WARMUP_RECONFIG if warmup else _lib.TAG_RECONFIG
        seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        ev = _lib.ptr(chain.evals) if evals else None
        h = ctypes.c_void_p()
        common = (_lib.ptr(chain.spins), _lib.ptr(chain.log_w), _lib.ptr(chain.e_loc),
                  _lib.ptr(chain.spins_b), _lib.ptr(chain.e_loc_b), _lib.ptr(chain.parents),
                  _lib.ptr(chain.work), self.n, float(beta))
        if batch:
            if diag is not None:
                raise NotImplementedError("batches measure the energy only")
            self._alw = alw
            rc = lib.gfmc_graph_capture_batch_steps(chain.b.handle, *common, _lib.ptr(alw), seed,
                                                    tp, tr, _lib.ptr(self.series), ev,
                                                    _lib.ptr(self.counter), ctypes.byref(h))
        else:
            dargs = (None, None, None, None)
            if diag is not None:
                self.dseries = diag.series(self.n, dev)
                dargs = (diag.handle, _lib.ptr(diag.ring), _lib.ptr(self.dseries),
                         _lib.ptr(diag.work))
            rc = lib.gfmc_graph_capture_steps_carry(chain.p.handle, chain.nw, *common,
                                                    float(alw), seed, tp, tr,
                                                    _lib.ptr(self.series), ev, *dargs,
                                                    _lib.ptr(chain.carry),
                                                    _lib.ptr(self.counter), ctypes.byref(h))
        _lib.check(rc, "gfmc_graph_capture")
        self.handle = h
        self.chain = chain

    def launch(self, step0, out=None, dout=None):
        """Run the ``n`` steps from position ``step0``; ``out`` (optional)
        receives a copy of the series rows, ``dout`` of the forward-walking
        rows of a graph captured with ``diag``."""
        v = int(step0) & 0xFFFFFFFF
        self.counter.fill_(v - 2 ** 32 if v >= 2 ** 31 else v)
        _lib.check(_lib.load().gfmc_graph_launch(self.handle, _lib.stream()),
                   "gfmc_graph_launch")
        if out is not None:
            out.copy_(self.series)
        if dout is not None:
            dout.copy_(self.dseries)
        return self.series

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().gfmc_graph_destroy(self.handle)
        except Exception:
            pass


# -- single-call API -----------------------------------------------------------

def evolve_walkers(walkers, hamiltonian, step_size, guide, method, rng_key,
                   avg_log_weight, lam=None, precision=None, return_evals=False):
    """One branching step of every walker without reconfiguration
    (evolve_walkers, evolution.py:761-814).  ``rng_key`` (seed, position)
    selects the Philox stream: counter = (hop, walker, position, PROPAGATE)."""
    require_method(method)
    require_step_size(method, step_size)
    key = as_key(rng_key)
    p = get_problem(hamiltonian, guide, precision)
    set_method(p, method, lam)
    w = walkers.clone()
    evals = torch.zeros(w.n_walkers, dtype=torch.int32, device=w.spins.device)
    p.propagate(w.spins, w.log_weight_dev, w.e_loc_dev, float(step_size),
                float(avg_log_weight), key.seed, key.position, _lib.TAG_PROPAGATE, evals)
    return (w, evals) if return_evals else w


def reconfiguration(walkers, rng_key, psi=None):
    """Systematic resampling (reconfiguration.py:27-62).  Returns
    ``(walkers, log_G_n - log n, E_n, indices)``; ``psi`` overrides the
    Philox offset (counter (0, 0, position, RECONFIG))."""
    key = as_key(rng_key)
    lib = _lib.load()
    nw = walkers.n_walkers
    dev = walkers.spins.device
    parents = torch.empty(nw, dtype=torch.int32, device=dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    work = torch.empty(_lib.work_doubles(nw), dtype=torch.float64, device=dev)
    psi_c = ctypes.c_double(float(psi)) if psi is not None else None
    _lib.check(lib.gfmc_reconfigure(
        nw, _lib.ptr(walkers.log_weight_dev), _lib.ptr(walkers.e_loc_dev), key.seed,
        key.position & 0xFFFFFFFF, _lib.TAG_RECONFIG,
        ctypes.byref(psi_c) if psi_c is not None else None,
        _lib.ptr(parents), _lib.ptr(out2), _lib.ptr(work), _lib.stream()), "gfmc_reconfigure")
    spins = torch.empty_like(walkers.spins)
    e = torch.empty_like(walkers.e_loc_dev)
    lw = torch.empty_like(walkers.log_weight_dev)
    _lib.check(lib.gfmc_gather(walkers.n_sites, nw, _lib.ptr(walkers.spins),
                               _lib.ptr(walkers.e_loc_dev), _lib.ptr(parents), _lib.ptr(spins),
                               _lib.ptr(e), _lib.ptr(lw), _lib.stream()), "gfmc_gather")
    o = out2.cpu().numpy()
    return (Walker.from_device(spins, lw, e, walkers.n_sites), float(o[0]), float(o[1]),
            parents.cpu().numpy())


def systematic_parents(w, psi):
    """Parents of systematic resampling for normalised weights ``w`` and
    offset ``psi`` (reconfiguration.py:9-24), on the device."""
    w = np.asarray(w, dtype=np.float64)
    with np.errstate(divide="ignore"):
        lw = np.log(w)
    dev = _lib.device()
    lwt = torch.as_tensor(lw, device=dev)
    e = torch.zeros_like(lwt)
    parents = torch.empty(len(w), dtype=torch.int32, device=dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    work = torch.empty(_lib.work_doubles(len(w)), dtype=torch.float64, device=dev)
    psi_c = ctypes.c_double(float(psi))
    _lib.check(_lib.load().gfmc_reconfigure(len(w), _lib.ptr(lwt), _lib.ptr(e), 0, 0,
                                            _lib.TAG_RECONFIG, ctypes.byref(psi_c),
                                            _lib.ptr(parents), _lib.ptr(out2), _lib.ptr(work),
                                            _lib.stream()), "gfmc_reconfigure")
    return parents.cpu().numpy()


def sample_index(weights, uniforms, precision="fp64"):
    """Slots drawn from non-negative ``weights`` at each uniform in [0, 1)
    with the propagator's sampler (sample_index, evolution.py:30-59)."""
    w = torch.as_tensor(np.asarray(weights, dtype=np.float64), device=_lib.device())
    u = torch.as_tensor(np.atleast_1d(np.asarray(uniforms, dtype=np.float64)), device=w.device)
    out = torch.empty(u.shape[0], dtype=torch.int32, device=w.device)
    prec = _lib.GFMC_FP64 if precision in ("fp64", _lib.GFMC_FP64) else _lib.GFMC_FP32
    _lib.check(_lib.load().gfmc_sample_index(prec, w.shape[0], _lib.ptr(w), u.shape[0],
                                             _lib.ptr(u), _lib.ptr(out), _lib.stream()),
               "gfmc_sample_index")
    return out.cpu().numpy()


def estimate_average_log_branch_weight(walkers, hamiltonian, guide, method="cont_time",
                                       lam=None, precision=None):
    """mean(-sum rates) for cont_time, mean(log(lam - sum rates)) for the
    discrete methods (estimators.py:722-751)."""
    require_method(method)
    lam = require_lam(method, lam)
    p = get_problem(hamiltonian, guide, precision)
    _, e, _ = p.rates(walkers.spins, want_rates=False)
    if lam is None:
        return float(torch.mean(-e).item())
    return float(torch.mean(torch.log(lam - e)).item())


# -- the driver ----------------------------------------------------------------

def _warmup(chain, key, step_f, alw, warmup_steps, verbosity, method="cont_time", lam=None,
            trace=None):
    """Four-stage ramp; the shift is updated per stage unless pinned
    (driver.py:713-788): alw <- alw + (LSE(log G_stage) - log n) / step.
    ``trace`` (a list) receives (step size, steps, measured shift) per
    stage."""
    pinned = alw is not None
    if not pinned:
        _, e, _ = chain.p.rates(chain.spins, want_rates=False)
        alw = float(torch.mean(-e).item() if lam is None else
                    torch.mean(torch.log(lam - e)).item())
    counter = key.position
    stages = warmup_schedule(step_f, warmup_steps, method)
    for i, (step_, n_w) in enumerate(stages):
        t0 = time.time()
        with span(f"gfmc warmup {i + 1}/{len(stages)} step={step_:g} n={n_w}"):
            series = chain.steps(n_w, step_, alw, key.seed, counter, warmup=True)
        counter += n_w
        lg = series[:, 0].cpu().numpy()
        check_population(lg, 0)
        m = np.max(lg)
        lse = m + np.log(np.sum(np.exp(lg - m)))
        measured = alw + (lse - np.log(n_w)) / step_
        if trace is not None:
            trace.append((step_, n_w, float(measured)))
        if verbosity >= 8:
            label = f"{i + 1}/{len(stages)}"
            rate = n_w / max(time.time() - t0, 1e-9)
            if method == "cont_time":
                print(f"[gfmc] - [run_warmup {label}] - step={step_}, {n_w} steps @ "
                      f"{rate:.1f} it/s, avg_log_weight = {measured:.6f}", flush=True)
            else:
                print(f"[gfmc] - [run_warmup {label}] - step={step_}, {n_w} steps @ "
                      f"{rate:.1f} it/s, average_b_x = {np.exp(measured):.6f}", flush=True)
            if pinned:
                print(f"[gfmc] - [run_warmup {label}] - alw pinned at {alw:.6f}; this stage "
                      f"would have measured {measured:.6f}", flush=True)
        if not pinned:
            alw = float(measured)
    if not stages and verbosity >= 8:
        print(f"[gfmc] - [run_gfmc] - no warmup (warmup_steps=0), avg_log_weight={alw:.6f}",
              flush=True)
    return alw


def run_gfmc(walkers, n_steps, hamiltonian, step_size, guide, p_array, *,
             method="cont_time", rng_key=None, lam=None, diag_operators=None,
             off_diag_operators=None, return_raw=True, avg_log_weight=None,
             warmup_steps=None, verbosity=7, number_of_outputs=100, off_diag_stride=1,
             off_diag_batch_size=None, precision=None, use_graph=True, lanes="auto",
             _step_offset=0):
    """Single-GPU GFMC run (run_gfmc, driver.py:1148-1547).

    Returns ``(walkers, energy, obs, stats, raw)`` (or the first four with
    ``return_raw=False``): ``energy`` {p: E_p}, ``obs`` {name: {p: value}}
    of the diagonal operators (``diag_operators``: {name: SpinHalf,
    SzSzStructureFactor or Identity}; a scalar per p, or one value per
    wavevector) and of the off-diagonal operators (``off_diag_operators``:
    {name: SpinHalf with rates >= 0, or a diagonal SpinHalf / Identity},
    forward walking with one shadow chain per operator at every measured
    step: every ``off_diag_stride``-th step, with ``return_raw=False`` only
    those every window can use),
    ``stats`` with ``avg_log_weight``, ``it_per_sec``, ``runtime``,
    ``total_wall_time``, ``number_of_devices`` and ``parameters/*``, ``raw``
    with the per-step ``log_weights`` and ``energy`` series, ``diag`` {name:
    (n_steps, [n_vals,] n_p)}, the lead-in ``additional_diag`` (the m_max
    steps after the run, traced with its ancestry), ``off_diag`` and
    ``off_diag_norm`` {name: (n_meas, n_p)} (the walker mean of the operator
    values after m shadow steps and the shadow's log weight product) and
    ``off_diag_steps`` (n_meas,) (the 0-based step of each row).
    ``off_diag_batch_size`` is accepted for compatibility (it must divide
    the walker count); the GPU applies an operator to all walkers at once.  ``rng_key``:
    an int seed or (seed, position); ``stats['parameters/rng_key_next']``
    continues the chain.  ``precision``: 'fp64' / 'fp32' (default from the
    Hamiltonian's dtype).  ``use_graph``: production chunks replay one
    captured HIP graph (identical results to eager launches).  ``lanes``:
    lanes per walker of the propagation kernel (64, 32, 16) or "auto" for
    the autotuner (pygfmc_amd.autotune); pin it for bit-reproducible runs.
    ``_step_offset`` (the reference's continuation offset) must be 0: a
    continuation resumes from the position in ``rng_key`` instead.  As in
    the reference there is no catch-all: an unknown keyword is a
    ``TypeError``.
    """
    wall_start = time.time()
    if _step_offset != 0:
        raise ValueError("_step_offset is not used: continue a chain by passing "
                         "stats['parameters/rng_key_next'] as rng_key")
    diag_ops = dict(diag_operators) if diag_operators else {}
    off_ops = dict(off_diag_operators) if off_diag_operators else {}
    if rng_key is None:
        import secrets
        rng_key = secrets.randbits(32)
        if verbosity >= 1:
            print(f"[gfmc] - [run_gfmc] - rng_key not given; using seed {rng_key}", flush=True)
    key = as_key(rng_key)
    if walkers.n_walkers == 0:
        raise ValueError("At least one walker is required.")
    n_steps = validated_int("n_steps", n_steps, 1)
    require_method(method)
    require_step_size(method, step_size)
    if warmup_steps is None:
        warmup_steps = DEFAULT_WARMUP_STEPS if avg_log_weight is None else 0
    warmup_steps = validated_int("warmup_steps", warmup_steps, 0)
    number_of_outputs = validated_int("number_of_outputs", number_of_outputs, 1)
    off_diag_stride = validated_int("off_diag_stride", off_diag_stride, 1)
    if off_diag_batch_size is not None:
        off_diag_batch_size = validated_int("off_diag_batch_size", off_diag_batch_size, 1)
        if walkers.n_walkers % off_diag_batch_size:
            raise ValueError(f"off_diag_batch_size={off_diag_batch_size} must divide "
                             f"N_walkers={walkers.n_walkers}")
    p_sorted = validated_p_array(p_array)
    collisions = set(diag_ops) & set(off_ops)
    if collisions:
        raise ValueError(f"diag_operators and off_diag_operators share name(s) "
                         f"{sorted(collisions)}; the returned obs dict is keyed by name, so "
                         f"colliding entries would silently overwrite each other.")
    from .offdiagfw import check_even_p, measured_steps, measurement_plan
    if off_ops:
        check_even_p(off_ops, p_sorted)
    plan = measurement_plan(n_steps, p_sorted, len(off_ops), off_diag_stride, return_raw)
    measured = measured_steps(n_steps, plan) if off_ops else None
    if avg_log_weight is None and warmup_steps == 0:
        raise ValueError("warmup_steps=0 needs an avg_log_weight: without a warm-up nothing "
                         "measures it.")
    step_f = float(step_size)
    p = get_problem(hamiltonian, guide, precision)
    lam_v = set_method(p, method, lam)
    chain = Chain(p, walkers)
    tune_times = None
    if lanes == "auto":
        from .autotune import tune_problem
        _, tune_times = tune_problem(p, chain.spins, step_f, 0.0)
    else:
        p.set_lanes(int(lanes))
    alw_given = avg_log_weight
    alw = _warmup(chain, key, step_f, avg_log_weight, warmup_steps, verbosity, method, lam_v)

    print_every = max(1, n_steps // max(1, number_of_outputs))
    dev = chain.spins.device
    series = torch.empty((n_steps, 2), dtype=torch.float64, device=dev)
    m_array = [q // 2 for q in p_sorted]
    # a captured MLP-guided step runs a fixed number of hop iterations: twice
    # the most any eager step needed, plus 16; the count of steps it left
    # unfinished is checked after every chunk
    graphable = True
    unfinished0, max_hops = p.mlp_step_stats()
    if p.guide_kind == _lib.GUIDE_MLP and max_hops > 0:
        p.set_mlp_hop_cap(2 * max_hops + 16)
    diag = None
    dseries = None
    if diag_ops:
        from .diagfw import DiagRecorder
        diag = DiagRecorder(diag_ops, walkers.n_sites, m_array, walkers.n_walkers, dev)
        dseries = diag.series(n_steps, dev)
    od = od_fw = od_norm = None
    od_row = None
    if off_ops:
        from .offdiagfw import OffDiagMeasurer
        od = OffDiagMeasurer(off_ops, p, guide, m_array, walkers.n_walkers, dev)
        od.check_positivity(chain.spins, verbose=verbosity >= 1)
        n_meas = int(measured.sum())
        od_fw, od_norm = od.rows(n_meas, dev)
        od_row = np.cumsum(measured) - 1
    dchunks = []
    evals_sum, evals_n = 0.0, 0
    torch.cuda.synchronize()
    t_start = time.time()
    pending = None
    step0 = 0
    host_chunks = []
    graph = None
    from .estimators import DeviceEnergyEstimator
    fw = DeviceEnergyEstimator(p_sorted, dev)
    while step0 < n_steps or pending is not None:
        if step0 < n_steps:
            n = min(print_every, n_steps - step0)
            with span(f"gfmc production steps {step0}-{step0 + n - 1}"):
                chain.evals.zero_()
                drows = dseries[step0:step0 + n] if diag is not None else None
                if od is not None:
                    # one step at a time: a measured step launches the shadow
                    # chains right after its resampling (driver.py:290-331)
                    for k in range(n):
                        t = step0 + k
                        chain.steps(1, step_f, alw, key.seed, key.position + t,
                                    series=series[t:t + 1], evals=True, diag=diag,
                                    dseries=drows[k:k + 1] if diag is not None else None)
                        if measured[t]:
                            r = int(od_row[t])
                            od.measure(chain.spins, step_f, alw, key.seed, key.position + t,
                                       od_fw[r], od_norm[r])
                elif (use_graph and graphable and n == print_every
                        and n_steps >= 2 * print_every):
                    if graph is None:
                        graph = StepGraph(chain, n, step_f, alw, key.seed, evals=True, diag=diag)
                    graph.launch(key.position + step0, out=series[step0:step0 + n], dout=drows)
                else:
                    chain.steps(n, step_f, alw, key.seed, key.position + step0,
                                series=series[step0:step0 + n], evals=True, diag=diag,
                                dseries=drows)
                fw.update(series[step0:step0 + n])
                ev = chain.evals.to(torch.float64).mean() / n
            nxt = (step0, n, ev)
            step0 += n
        else:
            nxt = None
        if pending is not None:
            s0, n, ev = pending
            chunk = series[s0:s0 + n].cpu().numpy()
            if p.guide_kind == _lib.GUIDE_MLP and p.mlp_step_stats()[0] > unfinished0:
                raise RuntimeError(
                    f"MLP guide: a branching step captured into a graph needed more hop "
                    f"iterations than its cap (steps {s0}..{s0 + n - 1}); raise it with "
                    f"Problem.set_mlp_hop_cap or run with use_graph=False")
            check_population(chunk[:, 0], s0, chunk[:, 1])
            host_chunks.append(chunk)
            if diag is not None:
                dchunks.append(dseries[s0:s0 + n].cpu().numpy())
            evals_sum += float(ev.item())
            evals_n += 1
            if verbosity >= 3 and len(host_chunks) % max(1, number_of_outputs // 10) == 0:
                print(f"[gfmc] - [run_gfmc] - step {s0 + n}/{n_steps}", flush=True)
        pending = nxt
    torch.cuda.synchronize()
    runtime = time.time() - t_start
    arr = np.concatenate(host_chunks, axis=0)
    log_weights, E = arr[:, 0].copy(), arr[:, 1].copy()
    energy = estimate_energy(p_sorted, log_weights, E)
    energy_device = fw.result()
    out_walkers = chain.walkers()
    diag_np = additional_diag = None
    off_np = off_norm_np = off_steps = None
    obs = {}
    m_common = max(m_array) if (diag is not None and od is not None) else 0
    if diag is not None:
        diag_np = diag.split(np.concatenate(dchunks, axis=0))
        obs = estimate_diagonal(p_sorted, log_weights, diag_np, m_max_common=m_common)
    if od is not None:
        fw_h, nrm_h = od_fw.cpu().numpy(), od_norm.cpu().numpy()
        off_np = {name: np.ascontiguousarray(fw_h[:, i]) for i, name in enumerate(od.names)}
        off_norm_np = {name: np.ascontiguousarray(nrm_h[:, i])
                       for i, name in enumerate(od.names)}
        off_steps = np.flatnonzero(measured).astype(np.int64)
        obs.update(estimate_off_diagonal(p_sorted, log_weights, off_np, off_norm_np, off_steps,
                                         m_array, m_common, od.names))
    if diag is not None:
        if diag.m_max > 0:
            # the lead-in a continuation substitutes for its first rows
            # (driver.py:1100-1127): m_max further steps traced with this
            # run's ancestry; the returned walkers stay the production-final
            # ones
            extra = diag.series(diag.m_max, dev)
            chain.steps(diag.m_max, step_f, alw, key.seed, key.position + n_steps,
                        diag=diag, dseries=extra)
            additional_diag = diag.split(extra.cpu().numpy())
    key_next = key.advance(n_steps)
    stats = {
        "avg_log_weight": float(alw if alw_given is None else alw_given),
        "it_per_sec": n_steps / runtime,
        "runtime": runtime,
        "total_wall_time": time.time() - wall_start,
        "number_of_devices": 1,
        "hip_graph": graph is not None,
        "energy_device": energy_device,
        "autotune_seconds": tune_times,
        "evals_per_step_mean": evals_sum / max(1, evals_n),
        "parameters/n_steps": n_steps,
        "parameters/n_walkers": walkers.n_walkers,
        "parameters/n_sites": walkers.n_sites,
        "parameters/step_size": step_f,
        "parameters/lam": float("nan") if lam is None else float(lam),
        "parameters/p_array": np.asarray(p_sorted),
        "parameters/m_array": np.asarray([q // 2 for q in p_sorted]),
        "parameters/warmup_steps": warmup_steps,
        "parameters/method": method,
        "parameters/hamiltonian": type(hamiltonian).__name__,
        "parameters/guide": type(guide).__name__ if guide is not None else "None",
        "parameters/has_guide": int(guide is not None),
        "parameters/precision": "fp64" if p.precision == _lib.GFMC_FP64 else "fp32",
        "parameters/lanes_per_walker": p.lanes,
        "parameters/waves_per_workgroup": p.launch[0],
        "parameters/xcd_contiguous_walkers": bool(p.launch[1]),
        "parameters/rng_key": np.asarray([key.seed, key.position], dtype=np.uint64),
        "parameters/rng_key_next": np.asarray([key_next.seed, key_next.position],
                                              dtype=np.uint64),
        "parameters/n_diag_operators": len(diag_ops),
        "parameters/diag_operators": np.asarray(list(diag_ops), dtype="S64"),
        "parameters/diag_operator_hashes": np.asarray(
            [hashlib.sha1(op.fingerprint()).hexdigest()[:32] for op in diag_ops.values()],
            dtype="S32"),
        "parameters/off_diag_stride": int(plan.off_diag_stride),
        "parameters/off_diag_offset": int(plan.off_diag_offset),
        "parameters/off_diag_equal_windows": int(bool(plan.equal_windows)),
        "parameters/n_off_diag_operators": len(off_ops),
        "parameters/off_diag_operators": np.asarray(list(off_ops), dtype="S64"),
        "parameters/off_diag_operator_hashes": np.asarray(
            [hashlib.sha1(op.fingerprint()).hexdigest()[:32] for op in off_ops.values()],
            dtype="S32"),
    }
    if verbosity >= 1:
        print(f"[gfmc] - [run_gfmc] - {n_steps} steps @ {stats['it_per_sec']:.1f} it/s; "
              f"E = {energy}", flush=True)
    if return_raw:
        raw = dict(log_weights=log_weights, energy=E, diag=diag_np, off_diag=off_np,
                   off_diag_norm=off_norm_np, off_diag_steps=off_steps,
                   additional_diag=additional_diag, p_array=list(p_sorted), m_array=m_array,
                   diag_operators=list(diag_ops) or None,
                   off_diag_operators=list(off_ops) or None)
        return out_walkers, energy, obs, stats, raw
    return out_walkers, energy, obs, stats


def run(walkers, n_steps, hamiltonian, step_size, guide, p_array, *, method="cont_time",
        rng_key=None, lam=None, n_devices=None, **kw):
    """Entry point (multi_gpu.run, multi_gpu.py:822-868): the keywords of
    :func:`run_gfmc` plus ``n_devices`` (validated as the reference does,
    an integer >= 1).  This engine runs one GPU: ``n_devices`` above 1
    raises ``ValueError`` instead of sharding the walkers."""
    if n_devices is not None:
        n_devices = validated_int("n_devices", n_devices, 1)
        if n_devices > 1:
            raise ValueError(f"n_devices={n_devices}: this engine runs on one GPU; sharding "
                             "walkers over several GPUs is not implemented")
    return run_gfmc(walkers, n_steps, hamiltonian, step_size, guide, p_array, method=method,
                    rng_key=rng_key, lam=lam, **kw)
