"""End-to-end runs on the GPU against exact energies (ED, Bethe ansatz)."""
import sys
import warnings

import numpy as np
import pytest

from helpers import golden_system, ground_energy
from pygfmc_amd import (JastrowSpinHalf, SpinHalf, chain, evaluate_observables,
                        prepare_walkers_random, prepare_walkers_sz, run, square)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def energy_and_error(raw, p):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    return E[p], err[p]


@pytest.mark.parametrize("guided", [False, True])
def test_golden_system_energy_matches_ed(guided):
    """chain(8) XXZ with h_x = 0.5 (no conserved S^z): GFMC at p = 16 within
    4 sigma + 2e-3 of exact diagonalisation."""
    say("ED chain8", guided)
    H, G = golden_system(guide=guided)
    e0 = ground_energy(H)
    w = prepare_walkers_random(1024, 8, 3)
    _, E, _, stats, raw = run(w, 3000, H, 0.2, G, [4, 16], rng_key=5, warmup_steps=400,
                              verbosity=0)
    e, err = energy_and_error(raw, 16)
    say("  E", e, "+-", err, "exact", e0)
    assert abs(e - e0) < 4 * err + 2e-3


def test_square4_heisenberg_energy_matches_ed():
    """4x4 periodic Heisenberg in the sign-free frame, Jastrow guide
    (beta = 0.15), FP32 ratios: within 4 sigma + 1e-3 of ED (the CPU oracle
    gives -11.2269 +- 0.0008 against -11.22848 on this setup)."""
    say("ED square4")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    e0 = ground_energy(H)
    G = JastrowSpinHalf(16, L.edges, beta=0.15)
    w = prepare_walkers_sz(1024, 16, 1)
    _, E, _, stats, raw = run(w, 3000, H, 0.2, G, [4, 32], rng_key=2, warmup_steps=400,
                              verbosity=0)
    e, err = energy_and_error(raw, 32)
    say("  E", e, "+-", err, "exact", e0)
    assert abs(e - e0) < 4 * err + 1e-3
    assert err < 0.01


def test_heisenberg_ring_100_bethe_ansatz():
    """The quick-start configuration (README.md:36-53): L = 100 Heisenberg
    ring, Jastrow beta = 0.3, 1024 walkers with Sz = 0, step 0.2,
    p = [4, 16, 64, 128], 2000 warm-up + 3000 steps, FP64.  E[128]/L within
    3 sigma + 2e-5 of the Bethe-ansatz value -0.4432296."""
    say("L=100 ring")
    L = chain(100)
    H = SpinHalf.create(L.edges, J_x=-1.0, J_y=-1.0, J_z=1.0, dtype=np.float64)
    G = JastrowSpinHalf(100, L.edges, beta=0.3, dtype=np.float64)
    w = prepare_walkers_sz(1024, 100, 0)
    _, E, _, stats, raw = run(w, 3000, H, 0.2, G, [4, 16, 64, 128], rng_key=0,
                              warmup_steps=2000, verbosity=0)
    e, err = energy_and_error(raw, 128)
    say("  E/L", e / 100, "+-", err / 100, "it/s", stats["it_per_sec"])
    assert abs(e / 100 - (-0.4432296)) < 3 * err / 100 + 2e-5
    assert err / 100 < 5e-4
