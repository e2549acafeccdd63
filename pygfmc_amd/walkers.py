"""The walker population and its initialisers (walkers/walker.py:50-342).

A :class:`Walker` keeps its state on the GPU: bit-packed spins
(``n_words = ceil(N/32)`` uint32 words per walker, stored in an int32
tensor), FP64 log-weights and FP64 local energies.  ``config``,
``log_weight`` and ``e_loc`` read back host copies, so code written against
the reference's ``Walker(config, log_weight, e_loc)`` NamedTuple keeps
working.
"""
from __future__ import annotations

from collections import namedtuple

import numpy as np
import torch

from . import _lib

__all__ = ["Walker", "PhiloxKey", "prepare_walkers_sz", "prepare_walkers_random",
           "prepare_walkers_from_configs", "pack_configs", "unpack_spins", "as_key"]


class PhiloxKey(namedtuple("PhiloxKey", "seed position")):
    """The RNG state of a chain: a 64-bit Philox key and the chain position
    (the step counter of the next production step).  Plays the role of the
    reference's JAX key, which is its chain position (driver.py:213-217)."""

    def advance(self, n):
        return PhiloxKey(self.seed, self.position + int(n))


def as_key(rng_key):
    """An int seed, a PhiloxKey, or a (seed, position) pair -> PhiloxKey."""
    if isinstance(rng_key, PhiloxKey):
        return rng_key
    if isinstance(rng_key, (tuple, list, np.ndarray)) and len(rng_key) == 2:
        return PhiloxKey(int(rng_key[0]) & 0xFFFFFFFFFFFFFFFF, int(rng_key[1]))
    if isinstance(rng_key, (int, np.integer)) and not isinstance(rng_key, bool):
        return PhiloxKey(int(rng_key) & 0xFFFFFFFFFFFFFFFF, 0)
    raise TypeError(f"rng_key must be an int seed, a PhiloxKey or (seed, position); "
                    f"got {type(rng_key).__name__}")


def n_words_of(n_sites):
    return (int(n_sites) + 31) // 32


def pack_configs(config):
    """(B, N) +-1 float64 device tensor -> (B, n_words) packed spins."""
    lib = _lib.load()
    c = config.to(dtype=torch.float64).contiguous()
    B, N = c.shape
    spins = torch.empty((B, n_words_of(N)), dtype=torch.int32, device=c.device)
    _lib.check(lib.gfmc_pack(N, B, _lib.ptr(c), _lib.ptr(spins), _lib.stream()), "gfmc_pack")
    return spins


def unpack_spins(spins, n_sites):
    """(B, n_words) packed spins -> (B, N) +-1 float64 device tensor."""
    lib = _lib.load()
    B = spins.shape[0]
    out = torch.empty((B, int(n_sites)), dtype=torch.float64, device=spins.device)
    _lib.check(lib.gfmc_unpack(int(n_sites), B, _lib.ptr(spins), _lib.ptr(out), _lib.stream()),
               "gfmc_unpack")
    return out


class Walker:
    """Walker population on the GPU (mirrors ``Walker``, walker.py:50-64).

    ``Walker(config, log_weight, e_loc)`` packs host or device arrays;
    :meth:`from_device` wraps existing device tensors.
    """

    def __init__(self, config, log_weight=None, e_loc=None):
        dev = _lib.device()
        c = torch.as_tensor(np.asarray(config, dtype=np.float64) if not torch.is_tensor(config)
                            else config, dtype=torch.float64, device=dev)
        if c.ndim != 2:
            raise ValueError(f"config must have shape (N_walkers, N_sites); got {tuple(c.shape)}")
        if not bool(torch.all((c == 1) | (c == -1))):
            raise ValueError("spin-1/2 configurations must be +-1")
        nw = c.shape[0]
        self.n_sites = int(c.shape[1])
        self.spins = pack_configs(c)
        self.log_weight_dev = self._vec(log_weight, nw, dev)
        self.e_loc_dev = self._vec(e_loc, nw, dev)

    @staticmethod
    def _vec(v, nw, dev):
        if v is None:
            return torch.zeros(nw, dtype=torch.float64, device=dev)
        t = torch.as_tensor(np.asarray(v) if not torch.is_tensor(v) else v,
                            dtype=torch.float64, device=dev).reshape(-1).clone()
        if t.shape[0] != nw:
            raise ValueError(f"expected {nw} values, got {t.shape[0]}")
        return t

    @classmethod
    def from_device(cls, spins, log_weight, e_loc, n_sites):
        obj = object.__new__(cls)
        obj.n_sites = int(n_sites)
        obj.spins = spins
        obj.log_weight_dev = log_weight
        obj.e_loc_dev = e_loc
        return obj

    @property
    def n_walkers(self):
        return int(self.spins.shape[0])

    @property
    def config(self):
        return unpack_spins(self.spins, self.n_sites).cpu().numpy()

    @property
    def log_weight(self):
        return self.log_weight_dev.cpu().numpy()

    @property
    def e_loc(self):
        return self.e_loc_dev.cpu().numpy()

    def clone(self):
        return Walker.from_device(self.spins.clone(), self.log_weight_dev.clone(),
                                  self.e_loc_dev.clone(), self.n_sites)

    def __len__(self):
        return 3

    def __iter__(self):
        return iter((self.config, self.log_weight, self.e_loc))


def _check_states(states):
    vals = sorted(float(v) for v in np.asarray(states, dtype=float).ravel())
    if vals != [-1.0, 1.0]:
        raise ValueError(f"pygfmc_amd walkers are spin-1/2 with states (-1, 1); got {states!r}")


def prepare_walkers_sz(N_walkers, N_sites, rng_key, sz=0, states=(-1.0, 1.0)):
    """Walkers in the sector sum_i sigma_i = sz (prepare_walkers_sz,
    walker.py:257-306): every walker a random permutation of one base
    multiset, drawn by a Philox-driven Fisher-Yates shuffle on the GPU."""
    _check_states(states)
    N_walkers, N_sites = int(N_walkers), int(N_sites)
    if N_walkers < 1:
        raise ValueError("At least one walker is required.")
    szf = float(sz)
    if szf != np.floor(szf) or abs(szf) > N_sites or (N_sites + int(szf)) % 2:
        raise ValueError(f"sz={sz} is not reachable with N_sites={N_sites} sites drawn from "
                         f"states={list(states)}; N_sites + sz must be even and |sz| <= N.")
    n_up = (N_sites + int(szf)) // 2
    key = as_key(rng_key)
    dev = _lib.device()
    spins = torch.empty((N_walkers, n_words_of(N_sites)), dtype=torch.int32, device=dev)
    _lib.check(_lib.load().gfmc_init_walkers_sz(N_sites, N_walkers, n_up, key.seed,
                                                _lib.ptr(spins), _lib.stream()),
               "gfmc_init_walkers_sz")
    z = torch.zeros(N_walkers, dtype=torch.float64, device=dev)
    return Walker.from_device(spins, z, z.clone(), N_sites)


def prepare_walkers_random(N_walkers, N_sites, rng_key, states=(-1.0, 1.0)):
    """Every site an independent fair +-1 (prepare_walkers_random,
    walker.py:207-254), Philox bits on the GPU."""
    _check_states(states)
    N_walkers, N_sites = int(N_walkers), int(N_sites)
    if N_walkers < 1:
        raise ValueError("At least one walker is required.")
    key = as_key(rng_key)
    dev = _lib.device()
    spins = torch.empty((N_walkers, n_words_of(N_sites)), dtype=torch.int32, device=dev)
    _lib.check(_lib.load().gfmc_init_walkers_random(N_sites, N_walkers, key.seed,
                                                    _lib.ptr(spins), _lib.stream()),
               "gfmc_init_walkers_random")
    z = torch.zeros(N_walkers, dtype=torch.float64, device=dev)
    return Walker.from_device(spins, z, z.clone(), N_sites)


def prepare_walkers_from_configs(configs):
    """Walkers at given +-1 configurations (walker.py:309-342)."""
    c = np.asarray(configs) if not torch.is_tensor(configs) else configs
    if c.ndim != 2:
        raise ValueError(f"configs must have shape (N_walkers, N_sites); got shape "
                         f"{tuple(c.shape)}.")
    return Walker(c)
