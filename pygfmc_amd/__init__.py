"""pygfmc_amd -- continuous-time Green's-function Monte Carlo on AMD MI355X.

The user API follows PyGFMC (``run``/``run_gfmc``, ``SpinHalf.create``,
``JastrowSpinHalf``, ``prepare_walkers_sz``, ``Walker``); the hot path runs
in hand-written HIP kernels (``libgfmc_hip.so``, built by
``python -m pygfmc_amd.build``) with PyTorch tensors as device buffers.
"""
from .engine import (DEFAULT_WARMUP_STEPS, evolve_walkers, estimate_average_log_branch_weight,
                     reconfiguration, run, run_gfmc, sample_index, systematic_parents)
from .estimators import (estimate_diagonal, estimate_energy, estimate_off_diagonal,
                         evaluate_observables)
from .guides import JastrowSpinHalf, MLPSpinHalf, RBMSpinHalf
from .batch import Batch, run_batch
from .checkpoint import load_raw, read_checkpoint, run_from_checkpoint, write_checkpoint
from .hamiltonian import SpinHalf, tfim_chain_ground_energy, transverse_field_ising
from .lattice import Lattice, chain, kagome, square
from .observables import Identity, SzSzStructureFactor, is_valid_k
from .training import train_variance, vmin_gradient
from .walkers import (PhiloxKey, Walker, prepare_walkers_from_configs, prepare_walkers_random,
                      prepare_walkers_sz)

__version__ = "0.1.0"

__all__ = ["SzSzStructureFactor", "Identity", "is_valid_k", "estimate_diagonal",
           "estimate_off_diagonal", "run", "run_gfmc", "evolve_walkers", "reconfiguration",
           "systematic_parents", "sample_index", "estimate_average_log_branch_weight",
           "estimate_energy", "evaluate_observables", "SpinHalf", "JastrowSpinHalf",
           "RBMSpinHalf", "MLPSpinHalf", "train_variance", "vmin_gradient", "run_batch", "Batch",
           "write_checkpoint", "read_checkpoint", "load_raw", "run_from_checkpoint",
           "transverse_field_ising", "tfim_chain_ground_energy", "Lattice", "chain", "square",
           "kagome", "Walker", "PhiloxKey", "prepare_walkers_sz", "prepare_walkers_random",
           "prepare_walkers_from_configs", "DEFAULT_WARMUP_STEPS"]
