"""Failure handling forced from the host: an extinct population is an error
that names the step (the reference's
tests/core/test_positivity_and_zero_weight.py:101), and a discrete sub-step
whose total weight lam - H_xx + R is zero keeps the walker (its weight
becomes zero) instead of drawing a slot that has no weight."""
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
