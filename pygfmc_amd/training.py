"""Training of the network guides (RBM, two-layer MLP) on configurations
GFMC generates.

The loss is the variance of the local energy over the current walker
population, L = (1/n) sum_i (E_i - mean E)^2; it vanishes for an exact
eigenstate on the walkers' support.  The reference has no counterpart: its
guides are trained by NetKet VMC energy minimisation (train_with_netket,
generic_netket_training.py:331-473).  Every step runs on the GPU through
the C-ABI: GFMC branching steps with the current guide (gfmc_run_steps), the
gradient (gfmc_rbm_vmin_grad: rates, theta by BF16 MFMA, the per-move
backward pass, the B^T X product by MFMA; gfmc_mlp_vmin_grad: the rows of
every candidate move through the network and back on BF16 MFMA), Adam
(gfmc_adam_step) and the re-packing of the weights into the problem
(gfmc_rbm_set_params / gfmc_mlp_set_params).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .guides import MLPSpinHalf, RBMSpinHalf
from .problem import Problem
from .walkers import as_key

__all__ = ["RBMTrainer", "MLPTrainer", "vmin_gradient", "train_variance"]


def orbit_tables(orbits, dev):
    """Device (offsets, members, n_orbits) of an orbit index per flat
    parameter: the members of orbit o in ascending flat order; parameters of
    orbit -1 (frozen) belong to none."""
    o = np.asarray(orbits, dtype=np.int64)
    live = np.flatnonzero(o >= 0)
    uniq, inv = np.unique(o[live], return_inverse=True)
    members = live[np.argsort(inv, kind="stable")].astype(np.int32)
    counts = np.bincount(inv, minlength=len(uniq))
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return (torch.as_tensor(offsets, device=dev), torch.as_tensor(members, device=dev),
            len(uniq))


def flat_params(rbm):
    """[W (hidden x N row-major), a, b] as one float32 vector."""
    return np.concatenate([rbm.W.reshape(-1), rbm.a, rbm.b]).astype(np.float32)


def rbm_from_flat(rbm, flat):
    flat = np.asarray(flat, dtype=np.float32)
    H, N = rbm.hidden, rbm.N
    return rbm.with_parameters(flat[:H * N].reshape(H, N), flat[H * N:H * N + N],
                               flat[H * N + N:])


class RBMTrainer:
    """Device state of variance minimisation for one RBM problem: FP32 master
    parameters, Adam moments, the gradient and its workspace."""

    def __init__(self, problem, rbm, lr=1e-2, betas=(0.9, 0.999), eps=1e-8, clip=0.0,
                 optimizer="adam", sr_shift=1e-2, sr_abs=1e-4, sr_iter=100, sr_max_norm=0.02):
        if not isinstance(rbm, RBMSpinHalf):
            raise TypeError(f"RBMTrainer trains an RBMSpinHalf, got {type(rbm).__name__}")
        if problem.info.guide != _lib.GUIDE_RBM:
            raise ValueError("the problem's guide is not an RBM")
        self.p = problem
        self.rbm = rbm
        lib = _lib.load()
        self.n_params = int(lib.gfmc_rbm_n_params(problem.handle))
        dev = _lib.device()
        self.params = torch.as_tensor(flat_params(rbm), device=dev)
        if self.params.numel() != self.n_params:
            raise ValueError("parameter count differs from the problem's")
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0
        self.lr, self.betas, self.eps, self.clip = float(lr), betas, float(eps), float(clip)
        self._work = None
        self._work_n = -1
        self._orbits = None
        self._frozen = None
        if getattr(rbm, "orbits", None) is not None:
            self._orbits = orbit_tables(rbm.orbits, dev)
            fz = np.flatnonzero(np.asarray(rbm.orbits) < 0)
            if fz.size:
                self._frozen = torch.as_tensor(fz, device=dev)
        if optimizer not in ("adam", "sr"):
            raise ValueError(f"optimizer must be 'adam' or 'sr', got {optimizer!r}")
        self.optimizer = optimizer
        self.sr_shift, self.sr_abs, self.sr_iter = float(sr_shift), float(sr_abs), int(sr_iter)
        self.sr_max_norm = float(sr_max_norm)
        if optimizer == "sr":
            if self._orbits is None:
                raise ValueError("the natural-gradient (SR) update needs a tied guide "
                                 "(RBMSpinHalf.symmetric / from_jastrow)")
            m = self._orbits[2]
            self._g_t = torch.zeros(m, dtype=torch.float32, device=dev)
            self._delta = torch.zeros(m, dtype=torch.float32, device=dev)
            self._sr_work = torch.empty(int(lib.gfmc_sr_work_doubles(m)), dtype=torch.float64,
                                        device=dev)
            self._O = None
            self._theta = None
            self._sr_warm = False
        self._last_spins = None

    def tie(self, v):
        """Orbit means of ``v`` in place, frozen entries zeroed (no-op for an
        untied guide)."""
        if self._orbits is not None:
            off, mem, n_orb = self._orbits
            _lib.check(_lib.load().gfmc_orbit_mean(v.numel(), _lib.ptr(v), n_orb, _lib.ptr(off),
                                                   _lib.ptr(mem), _lib.stream()),
                       "gfmc_orbit_mean")
            if self._frozen is not None:
                v[self._frozen] = 0.0

    def _workspace(self, n):
        if n != self._work_n:
            nb = int(_lib.load().gfmc_rbm_vmin_work_bytes(self.p.handle, n))
            if nb < 0:
                raise _lib.GfmcError(_lib.load().gfmc_last_error().decode())
            self._work = torch.empty(nb, dtype=torch.uint8, device=self.params.device)
            self._work_n = n
        return self._work

    def gradient(self, spins, out2):
        """Enqueue the variance gradient over ``spins`` into ``self.grad``;
        ``out2`` (2 device doubles) receives (mean E, variance)."""
        n = spins.shape[0]
        _lib.check(_lib.load().gfmc_rbm_vmin_grad(
            self.p.handle, n, _lib.ptr(spins), self.clip, _lib.ptr(self.grad), _lib.ptr(out2),
            _lib.ptr(self._workspace(n)), _lib.stream()), "gfmc_rbm_vmin_grad")
        if self.optimizer == "sr":
            self._last_spins = spins
        else:
            self.tie(self.grad)

    def _sr_direction(self):
        """delta = (S + shift diag S + abs) ^ -1 g over the orbits of the
        last gradient's samples."""
        lib = _lib.load()
        off, mem, m = self._orbits
        spins = self._last_spins
        n = spins.shape[0]
        if self._O is None or self._O.shape[0] != n:
            self._O = torch.empty((n, m), dtype=torch.float32, device=spins.device)
            self._theta = torch.empty((n, self.rbm.hidden), dtype=torch.float32,
                                      device=spins.device)
        _lib.check(lib.gfmc_orbit_sum(self.n_params, _lib.ptr(self.grad), m, _lib.ptr(off),
                                      _lib.ptr(mem), _lib.ptr(self._g_t), _lib.stream()),
                   "gfmc_orbit_sum")
        _lib.check(lib.gfmc_rbm_orbit_O(self.p.handle, n, _lib.ptr(spins), m, _lib.ptr(off),
                                        _lib.ptr(mem), _lib.ptr(self._O), _lib.ptr(self._theta),
                                        _lib.stream()), "gfmc_rbm_orbit_O")
        # successive steps warm-start from the previous step's solution
        _lib.check(lib.gfmc_sr_solve(n, m, _lib.ptr(self._O), _lib.ptr(self._g_t),
                                     self.sr_shift, self.sr_abs, self.sr_iter, 1e-5,
                                     self.sr_max_norm, int(self._sr_warm),
                                     _lib.ptr(self._delta), _lib.ptr(self._sr_work),
                                     _lib.stream()), "gfmc_sr_solve")
        self._sr_warm = True
        return self._delta

    def update(self):
        """One Adam step (or, with ``optimizer="sr"``, one natural-gradient
        step params -= lr delta on the tied parameters) on the master
        parameters, then the problem's BF16 weights from them."""
        self.t += 1
        lib = _lib.load()
        if self.optimizer == "sr":
            off, mem, m = self._orbits
            delta = self._sr_direction()
            _lib.check(lib.gfmc_orbit_axpy(self.n_params, _lib.ptr(self.params), m,
                                           _lib.ptr(off), _lib.ptr(mem), _lib.ptr(delta),
                                           -self.lr, _lib.stream()), "gfmc_orbit_axpy")
            self.push()
            return
        _lib.check(lib.gfmc_adam_step(self.n_params, _lib.ptr(self.params), _lib.ptr(self.grad),
                                      _lib.ptr(self.m), _lib.ptr(self.v), self.t, self.lr,
                                      float(self.betas[0]), float(self.betas[1]), self.eps,
                                      _lib.stream()), "gfmc_adam_step")
        self.push()

    def push(self):
        _lib.check(_lib.load().gfmc_rbm_set_params(self.p.handle, _lib.ptr(self.params),
                                                   _lib.stream()), "gfmc_rbm_set_params")

    def guide(self):
        """The trained guide (host copy; W rounded to BF16)."""
        return rbm_from_flat(self.rbm, self.params.cpu().numpy())


def mlp_flat_params(g):
    """[W1 (h1 x N), b1, W2 (h2 x h1), b2, w3, b3] as one float32 vector."""
    return np.concatenate([g.W1.reshape(-1), g.b1, g.W2.reshape(-1), g.b2, g.w3,
                           [g.b3]]).astype(np.float32)


def mlp_from_flat(g, flat):
    flat = np.asarray(flat, dtype=np.float32)
    N, (h1, h2) = g.N, g.hidden
    o = np.cumsum([0, h1 * N, h1, h2 * h1, h2, h2, 1])
    return g.with_parameters(flat[o[0]:o[1]].reshape(h1, N), flat[o[1]:o[2]],
                             flat[o[2]:o[3]].reshape(h2, h1), flat[o[3]:o[4]],
                             flat[o[4]:o[5]], float(flat[o[5]]))


class MLPTrainer(RBMTrainer):
    """Device state of variance minimisation for one MLP problem (the same
    interface as :class:`RBMTrainer`; the gradient workspace is the
    problem's own)."""

    def __init__(self, problem, mlp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, clip=0.0):
        if not isinstance(mlp, MLPSpinHalf):
            raise TypeError(f"MLPTrainer trains an MLPSpinHalf, got {type(mlp).__name__}")
        if problem.info.guide != _lib.GUIDE_MLP:
            raise ValueError("the problem's guide is not an MLP")
        self.p = problem
        self.rbm = mlp
        self.n_params = int(_lib.load().gfmc_mlp_n_params(problem.handle))
        dev = _lib.device()
        self.params = torch.as_tensor(mlp_flat_params(mlp), device=dev)
        if self.params.numel() != self.n_params:
            raise ValueError("parameter count differs from the problem's")
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0
        self.lr, self.betas, self.eps, self.clip = float(lr), betas, float(eps), float(clip)
        self._orbits = None
        self.optimizer = "adam"

    def gradient(self, spins, out2):
        self.p.reserve(spins.shape[0])
        _lib.check(_lib.load().gfmc_mlp_vmin_grad(
            self.p.handle, spins.shape[0], _lib.ptr(spins), self.clip, _lib.ptr(self.grad),
            _lib.ptr(out2), _lib.stream()), "gfmc_mlp_vmin_grad")

    def push(self):
        _lib.check(_lib.load().gfmc_mlp_set_params(self.p.handle, _lib.ptr(self.params),
                                                   _lib.stream()), "gfmc_mlp_set_params")

    def guide(self):
        return mlp_from_flat(self.rbm, self.params.cpu().numpy())


def _trainer(problem, guide, **kw):
    if isinstance(guide, MLPSpinHalf):
        if kw.pop("optimizer", "adam") != "adam":
            raise ValueError("the MLP guide trains with Adam")
        for k in ("sr_shift", "sr_abs", "sr_iter", "sr_max_norm"):
            kw.pop(k, None)
        return MLPTrainer(problem, guide, **kw)
    return RBMTrainer(problem, guide, **kw)


def vmin_gradient(hamiltonian, rbm, spins, clip=0.0):
    """(flat gradient, mean E, variance) of the local-energy variance of the
    network guide ``rbm`` (RBM or MLP) over the packed device configurations
    ``spins``."""
    p = Problem(hamiltonian, rbm, _lib.GFMC_FP32)
    tr = _trainer(p, rbm, clip=clip)
    out2 = torch.empty(2, dtype=torch.float64, device=spins.device)
    tr.gradient(spins, out2)
    o = out2.cpu().numpy()
    return tr.grad.cpu().numpy(), float(o[0]), float(o[1])


def train_variance(hamiltonian, rbm, walkers, *, n_iter=200, steps_per_iter=4, step_size=0.2,
                   lr=1e-2, betas=(0.9, 0.999), eps=1e-8, clip=0.0, rng_key=0,
                   warmup_steps=40, sampler="self", verbosity=0, trainer=None,
                   optimizer="adam", sr_shift=1e-2, sr_abs=1e-4, sr_iter=100,
                   sr_max_norm=0.02):
    """Variance minimisation of the network guide ``rbm`` (RBMSpinHalf or
    MLPSpinHalf) on GFMC samples.

    The walkers evolve under GFMC guided by ``sampler``: ``"self"`` (the RBM
    being trained), ``None`` (no guide) or any guide ``run`` accepts, e.g. a
    Jastrow guide.  A fixed sampler draws from psi_S phi_0, which does not
    move with the RBM; sampling with the RBM itself draws from
    psi_RBM phi_0 but can freeze when the RBM suppresses every move.
    ``warmup_steps`` branching steps, then per iteration ``steps_per_iter``
    steps followed by one gradient + Adam update on the resampled
    population.  ``optimizer="sr"`` replaces Adam by natural-gradient steps
    (stochastic reconfiguration: the covariance S of the samples' tied
    log-derivatives, (S + sr_shift diag S + sr_abs) delta = g by conjugate
    gradients, scaled so that the rms change of log psi over the samples is at
    most ``sr_max_norm``; params -= lr delta) for a tied guide.  ``trainer`` continues
    an earlier run (its parameters and optimizer state).  Returns ``(trained guide, history)``: arrays
    ``mean_e`` and ``variance`` (of the RBM's local energies, before each
    update), ``seconds`` (GPU wall time), ``walker_steps`` and ``trainer``.
    """
    key = as_key(rng_key)
    from .engine import Chain, set_method
    from .problem import get_problem
    if trainer is None:
        p = Problem(hamiltonian, rbm, _lib.GFMC_FP32)
        trainer = _trainer(p, rbm, lr=lr, betas=betas, eps=eps, clip=clip, optimizer=optimizer,
                           sr_shift=sr_shift, sr_abs=sr_abs, sr_iter=sr_iter,
                           sr_max_norm=sr_max_norm)
    else:
        p = trainer.p
        trainer.lr, trainer.betas, trainer.eps, trainer.clip = float(lr), betas, float(eps), \
            float(clip)
    tr = trainer
    ps = p if (isinstance(sampler, str) and sampler == "self") else get_problem(hamiltonian,
                                                                                sampler)
    set_method(ps, "cont_time", None)
    chain = Chain(ps, walkers)
    dev = chain.spins.device
    hist = torch.zeros((max(n_iter, 1), 2), dtype=torch.float64, device=dev)
    _, e, _ = ps.rates(chain.spins, want_rates=False)
    alw = float(torch.mean(-e).item())
    torch.cuda.synchronize()
    t0 = time.time()
    counter = key.position
    if warmup_steps:
        chain.steps(warmup_steps, step_size, alw, key.seed, counter, warmup=True)
        counter += warmup_steps
    for it in range(n_iter):
        chain.steps(steps_per_iter, step_size, alw, key.seed, counter)
        counter += steps_per_iter
        tr.gradient(chain.spins, hist[it])
        tr.update()
        if verbosity and (it + 1) % max(1, n_iter // 10) == 0:
            h = hist[it].cpu().numpy()
            print(f"[train] iter {it + 1}/{n_iter}: E = {h[0]:.6f}, var = {h[1]:.6f}",
                  flush=True)
    torch.cuda.synchronize()
    secs = time.time() - t0
    h = hist[:n_iter].cpu().numpy()
    bad = np.flatnonzero(~np.isfinite(h).all(axis=1))
    if bad.size:
        raise ValueError(
            f"[train] the local-energy statistics are not finite from iteration {int(bad[0])} "
            "on: the guide diverged (a ratio overflowed the rate precision, or its samples "
            "died out); lower the learning rate or the trust region (sr_max_norm)")
    history = dict(mean_e=h[:, 0].copy(), variance=h[:, 1].copy(), seconds=secs,
                   walker_steps=walkers.n_walkers * (warmup_steps + n_iter * steps_per_iter),
                   trainer=tr)
    return tr.guide(), history
