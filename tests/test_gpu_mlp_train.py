"""MLP guide training on the GPU: the variance-loss gradient (the rows of
every candidate move through the network and back on BF16 MFMA) against the
FP64 oracle, parameter upload, determinism, and a training run that cuts
the local-energy variance and keeps the GFMC energy at the exact value."""
import sys

import numpy as np
import pytest
import torch

from helpers import ground_energy
from pygfmc_amd import (MLPSpinHalf, SpinHalf, chain, evaluate_observables, kagome,
                        prepare_walkers_random, prepare_walkers_sz, run, train_variance)
from pygfmc_amd import _lib
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.training import MLPTrainer, mlp_flat_params, mlp_from_flat, vmin_gradient

pytestmark = pytest.mark.gpu

# FP32 rates and coefficients, hi + lo BF16 operands (2^-16 relative) and
# FP32 accumulation over rows: compared relative to the largest entry.
GRAD_TOL = 2e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def systems():
    out = []
    L = kagome(2, 2)
    out.append(("kagome2x2_64x32", SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32),
                MLPSpinHalf(12, (64, 32), seed=1, scale=1.5, out_scale=1.0), 80, True))
    out.append(("chain8_hx_64x64", SpinHalf.create(chain(8).edges, -1.0, -0.4, 0.7, h_x=0.6,
                                                   dtype=np.float32),
                MLPSpinHalf(8, (64, 64), seed=2, scale=1.5, out_scale=1.0), 60, False))
    L = kagome(3, 3)
    out.append(("kagome3x3_512x256", SpinHalf.create(L.edges, -1.0, -1.0, 1.0,
                                                     dtype=np.float32),
                MLPSpinHalf(27, (512, 256), seed=3, scale=1.0, out_scale=0.3), 40, True))
    return out


SYSTEMS = systems()
IDS = [s[0] for s in SYSTEMS]


def samples(orc, H, n, sz):
    return orc.init_sz(H.N, n, H.N // 2 + (H.N % 2), 11) if sz else orc.init_random(H.N, n, 11)


@pytest.mark.parametrize("name,H,G,n,sz", SYSTEMS, ids=IDS)
@pytest.mark.parametrize("clip", [0.0, 1.5])
def test_variance_gradient_matches_oracle(orc, name, H, G, n, sz, clip):
    say("mlp vmin grad", name, clip)
    spins = samples(orc, H, n, sz)
    want, m_ref, v_ref = orc.Problem(lower(H, G)).mlp_vmin_grad(spins, clip)
    got, m, v = vmin_gradient(H, G, to_dev(spins), clip)
    scale = np.max(np.abs(want))
    err = np.max(np.abs(got - want))
    say(f"  max |d grad| / max |grad| = {err / scale:.2e}")
    assert err <= GRAD_TOL * scale
    assert abs(m - m_ref) <= 1e-5 * (1 + abs(m_ref))
    assert abs(v - v_ref) <= 2e-5 * (1 + abs(v_ref))


def test_gradient_is_deterministic():
    name, H, G, n, sz = SYSTEMS[2]
    w = prepare_walkers_sz(300, H.N, 4, sz=1)
    g1, _, _ = vmin_gradient(H, G, w.spins)
    g2, _, _ = vmin_gradient(H, G, w.spins)
    np.testing.assert_array_equal(g1, g2)


def test_set_params_equals_a_fresh_problem(orc):
    name, H, G, n, sz = SYSTEMS[0]
    other = MLPSpinHalf(12, (64, 32), seed=9, scale=1.2, out_scale=0.8, b3=0.3)
    p = Problem(H, G, _lib.GFMC_FP32)
    tr = MLPTrainer(p, G)
    tr.params.copy_(torch.as_tensor(mlp_flat_params(other), device="cuda"))
    tr.push()
    spins = to_dev(samples(orc, H, 200, sz))
    a = p.log_psi(spins).cpu().numpy()
    b = Problem(H, other, _lib.GFMC_FP32).log_psi(spins).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    back = mlp_from_flat(G, mlp_flat_params(other))
    np.testing.assert_array_equal(back.W1, other.W1)
    assert back.b3 == np.float32(0.3)


def test_training_cuts_variance_and_keeps_the_exact_energy():
    """kagome 2x2 XXZ: variance minimisation on GFMC samples from a nearly
    flat MLP cuts the local-energy variance several-fold; GFMC with the
    trained guide stays at the exact energy with a smaller error bar."""
    say("mlp training kagome2x2")
    L = kagome(2, 2)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    e0 = ground_energy(H)
    g0 = MLPSpinHalf(12, (64, 32), seed=0, scale=1.0, out_scale=0.01)
    w = prepare_walkers_sz(1024, 12, 3)
    trained, hist = train_variance(H, g0, w, n_iter=300, steps_per_iter=2, lr=3e-3, rng_key=5)
    v0, v1 = hist["variance"][:5].mean(), hist["variance"][-20:].mean()
    say(f"  variance {v0:.4f} -> {v1:.4f} in {hist['seconds']:.2f} s")
    assert v1 < 0.3 * v0
    res = {}
    for nm, G in (("untrained", g0), ("trained", trained)):
        _, E, _, _, raw = run(prepare_walkers_sz(1024, 12, 7), 1500, H, 0.2, G, [8, 32],
                              rng_key=7, warmup_steps=200, verbosity=0)
        _, _, err, _ = evaluate_observables(raw, n_blocks=20)
        res[nm] = (E[32], err[32])
        say(f"  {nm}: E = {E[32]:.5f} +- {err[32]:.5f} (exact {e0:.5f})")
    for nm, (E, err) in res.items():
        assert abs(E - e0) < 4 * err + 2e-3, nm
    assert res["trained"][1] < 0.7 * res["untrained"][1]
