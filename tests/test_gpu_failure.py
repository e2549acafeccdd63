"""Failure handling forced from the host: an extinct population is an error
that names the step (the reference's
tests/core/test_positivity_and_zero_weight.py:101), a discrete sub-step
whose total weight lam - H_xx + R is zero keeps the walker (its weight
becomes zero) instead of drawing a slot that has no weight, and a total
rate that overflows (a guide ratio beyond the rate precision) ends the
walker's step with a non-finite weight, which the run reports, instead of
hopping with zero hold time forever."""
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import SpinHalf, chain, prepare_walkers_from_configs, run_gfmc
from pygfmc_amd import _lib
from pygfmc_amd.engine import set_method
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("method", ["multi_step", "single_step"])
def test_population_extinction_is_an_error_not_nan(method):
    """lam equals the diagonal element of every walker: every branch weight
    is zero and the driver names the step instead of returning NaN."""
    say("extinction", method)
    n = 4
    H = SpinHalf.create(chain(n).edges, J_z=1.0)          # diagonal only
    up = prepare_walkers_from_configs(np.ones((16, n)))
    lam = float(len(chain(n).edges)) * 0.25                # == H_xx of all-up
    with pytest.raises(ValueError, match="population extinct at step 0"):
        run_gfmc(up, 10, H, 2, None, [2], method=method, lam=lam, rng_key=0, verbosity=0)


def zero_total_case(orc, H, configs, lam):
    """One single_step propagation (3 sub-steps) on the GPU and in the
    oracle, FP64, same Philox draws."""
    p = Problem(H, None, _lib.GFMC_FP64)
    set_method(p, "single_step", lam)
    spins = orc.pack(configs)
    nw = spins.shape[0]
    s = torch.as_tensor(spins.view(np.int32), device="cuda")
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    p.propagate(s, lw, e, 3, 0.0, 31, 2, 0, ev)
    torch.cuda.synchronize()
    ref = orc.Problem(lower(H, None)).propagate(spins, np.zeros(nw), np.zeros(nw), 3, 0.0, 31, 2,
                                                0, method=1, lam=lam)
    return spins, (s.cpu().numpy().view(np.uint32), lw.cpu().numpy(), e.cpu().numpy(),
                   ev.cpu().numpy()), ref


def test_single_step_without_off_diagonal_slots_never_moves(orc):
    """A diagonal-only Hamiltonian (no off-diagonal slot at all) with lam ==
    max H_xx: the walkers at the maximum have zero total weight, the others
    stay with weight lam - H_xx; nobody moves."""
    say("zero total, no slots")
    H = SpinHalf.create(chain(6).edges, J_z=1.0)
    rng = np.random.default_rng(1)
    configs = rng.choice([-1.0, 1.0], size=(40, 6))
    configs[:8] = 1.0                                      # H_xx = 1.5 = lam
    spins, got, ref = zero_total_case(orc, H, configs, lam=1.5)
    np.testing.assert_array_equal(got[0], spins)
    np.testing.assert_array_equal(ref[0], spins)
    np.testing.assert_array_equal(got[3], ref[3])
    aligned = np.all(configs == configs[:, :1], axis=1)    # H_xx == lam exactly
    assert aligned[:8].all()
    np.testing.assert_array_equal(got[1] == -np.inf, aligned)
    np.testing.assert_array_equal(ref[1] == -np.inf, aligned)
    np.testing.assert_allclose(got[1][~aligned], ref[1][~aligned], rtol=1e-12, atol=1e-12)


def test_single_step_zero_total_with_slots_stays(orc):
    """XXZ chain: on the all-up configuration every hop has zero rate (no
    anti-parallel bond) and lam == H_xx makes the stay weight zero as well;
    such walkers stay (the oracle too) and get zero weight, the others move
    as the oracle does."""
    say("zero total, slots")
    H = SpinHalf.create(chain(6).edges, J_x=-1.0, J_y=-1.0, J_z=1.0)
    rng = np.random.default_rng(2)
    configs = rng.choice([-1.0, 1.0], size=(40, 6))
    configs[:8] = 1.0
    lam = 1.5
    spins, got, ref = zero_total_case(orc, H, configs, lam)
    np.testing.assert_array_equal(got[0][:8], spins[:8])
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[3], ref[3])
    assert np.all(got[1][:8] == -np.inf) and np.all(ref[1][:8] == -np.inf)
    fin = np.isfinite(ref[1])
    np.testing.assert_array_equal(np.isfinite(got[1]), fin)
    np.testing.assert_allclose(got[1][fin], ref[1][fin], rtol=1e-12, atol=1e-12)


def test_overflowing_total_rate_ends_the_step_like_the_oracle(orc):
    """Jastrow beta = 200 on a ring (FP64): walkers with a move of ratio
    e^800 have total rate inf, take no hop and end the step with log weight
    +inf, as in the oracle; every walker's spins, log weight and evaluation
    count equal the oracle's."""
    from pygfmc_amd import JastrowSpinHalf
    say("overflowing total rate, fp64")
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(8, L.edges, beta=200.0, dtype=np.float64)
    p = Problem(H, G, _lib.GFMC_FP64)
    spins = orc.init_sz(8, 64, 4, 7)
    nw = spins.shape[0]
    s = torch.as_tensor(spins.view(np.int32), device="cuda")
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    p.propagate(s, lw, e, 0.2, 0.0, 5, 1, 0, ev)
    torch.cuda.synchronize()
    ref = orc.Problem(lower(H, G)).propagate(spins, np.zeros(nw), np.zeros(nw), 0.2, 0.0, 5, 1, 0)
    np.testing.assert_array_equal(s.cpu().numpy().view(np.uint32), ref[0])
    np.testing.assert_array_equal(ev.cpu().numpy(), ref[3])
    got = lw.cpu().numpy()
    np.testing.assert_array_equal(np.isinf(got), np.isinf(ref[1]))
    assert np.isinf(got).any() and np.isfinite(got).any()
    fin = np.isfinite(ref[1])
    np.testing.assert_allclose(got[fin], ref[1][fin], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision,beta", [(np.float64, 200.0), (np.float32, 30.0)],
                         ids=["fp64", "fp32"])
def test_guide_ratio_overflow_is_an_error_not_a_hang(precision, beta):
    """A guide whose ratios overflow the rate precision: the run raises
    (the population check) instead of stepping forever."""
    from pygfmc_amd import JastrowSpinHalf, prepare_walkers_sz, run
    say("ratio overflow", np.dtype(precision).name)
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=precision)
    G = JastrowSpinHalf(8, L.edges, beta=beta, dtype=precision)
    w = prepare_walkers_sz(256, 8, 1)
    with pytest.raises(ValueError, match="population extinct|non-finite energy"):
        run(w, 20, H, 0.2, G, [2], rng_key=0, warmup_steps=0, avg_log_weight=0.0,
            verbosity=0)
