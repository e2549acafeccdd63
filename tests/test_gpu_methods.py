"""The discrete propagators (single_step, multi_step) on the GPU: bit-exact
moves against the oracle, batches, and energies against exact
diagonalisation."""
import sys
import warnings

import numpy as np
import pytest
import torch

from helpers import golden_system, ground_energy
from pygfmc_amd import (JastrowSpinHalf, SpinHalf, chain, evaluate_observables,
                        prepare_walkers_from_configs, prepare_walkers_sz, run, run_batch, square,
                        transverse_field_ising)
from pygfmc_amd import _lib
from pygfmc_amd.batch import Batch, BatchChain
from pygfmc_amd.engine import METHODS, Chain, set_method
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def systems():
    out = []
    H, G = golden_system()
    out.append(("golden_jastrow", H, G, 3.0))
    L = chain(100)
    out.append(("heis_chain100_noguide", SpinHalf.create(L.edges, -1.0, -1.0, 1.0,
                                                         dtype=np.float64), None, 25.0))
    L = square(4, 4)
    out.append(("xxz_square4", SpinHalf.create(L.edges, -1.0, -1.0, 1.5, dtype=np.float64),
                JastrowSpinHalf(16, L.edges, beta=0.2, dtype=np.float64), 12.0))
    L = chain(30)
    out.append(("tfim_chain30", transverse_field_ising(L.edges, 1.0, 0.8, dtype=np.float64),
                JastrowSpinHalf(30, L.edges, beta=-0.2, dtype=np.float64), 31.0))
    return out


SYSTEMS = systems()
CASES = [(s, m, g) for s in SYSTEMS for m in ("single_step", "multi_step") for g in (64, 16)]


@pytest.mark.parametrize("name,H,G,lam,method,lanes",
                         [(*s, m, g) for s, m, g in CASES],
                         ids=[f"{s[0]}-{m}-{g}" for s, m, g in CASES])
def test_discrete_propagation_bit_exact_vs_oracle(orc, name, H, G, lam, method, lanes):
    say("discrete propagate", name, method, lanes)
    prob = Problem(H, G, _lib.GFMC_FP64)
    if lanes not in prob.lanes_ok:
        pytest.skip(f"{name}: {lanes} lanes per walker not usable for this model")
    prob.set_lanes(lanes)
    set_method(prob, method, lam)
    nw = 203
    spins = orc.init_random(H.N, nw, 4) if H.n_hx else orc.init_sz(H.N, nw, H.N // 2, 4)
    lw0 = np.random.default_rng(0).normal(scale=0.1, size=nw)
    n_bra, alw, seed, step = 6, 0.3, 0xDEADBEEF12345, 7
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, lw0, np.zeros(nw), n_bra, alw, seed, step, 0, lanes=lanes,
        method=METHODS[method], lam=lam)
    s = to_dev(spins)
    lw = torch.as_tensor(lw0, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    prob.propagate(s, lw, e, n_bra, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ev.cpu().numpy(), ev_ref)
    np.testing.assert_array_equal(s.cpu().numpy().view(np.uint32), s_ref)
    np.testing.assert_allclose(lw.cpu().numpy(), lw_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    moved = np.any(s.cpu().numpy().view(np.uint32) != spins, axis=1)
    assert moved.mean() > 0.3


def test_discrete_step_loop_matches_oracle_chain(orc):
    name, H, G, lam = SYSTEMS[0]
    say("discrete step loop", name)
    prob = Problem(H, G, _lib.GFMC_FP64)
    set_method(prob, "multi_step", lam)
    nw = 256
    spins = orc.init_random(H.N, nw, 8)
    n, alw, seed, step0 = 5, 0.5, 4242, 100
    s_ref, e_ref, ser_ref = orc.run_chain(orc.Problem(lower(H, G)), spins, n, 6, alw, seed,
                                          step0, method=2, lam=lam)
    from pygfmc_amd.walkers import Walker
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ch = Chain(prob, Walker.from_device(to_dev(spins), z, z.clone(), H.N))
    ser = ch.steps(n, 6, alw, seed, step0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ch.spins.cpu().numpy().view(np.uint32), s_ref)
    np.testing.assert_allclose(ser.cpu().numpy(), ser_ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("method", ["single_step", "multi_step"])
def test_discrete_batch_equals_single_runs(method):
    say("discrete batch", method)
    Ls = [16, 24, 40]
    Hs = [transverse_field_ising(chain(L).edges, 1.0, 0.7 + 0.3 * i, dtype=np.float32)
          for i, L in enumerate(Ls)]
    Gs = [JastrowSpinHalf(L, chain(L).edges, beta=-0.2, dtype=np.float32) for L in Ls]
    lams = [float(L) + 2.0 for L in Ls]
    nws = [64, 132, 1028]
    from pygfmc_amd import prepare_walkers_random
    ws = [prepare_walkers_random(nw, L, i) for i, (L, nw) in enumerate(zip(Ls, nws))]
    b = Batch(Hs, Gs, nws, method=method, lam=lams)
    ch = BatchChain(b, ws)
    alw = torch.tensor([0.1, 0.2, 0.3], dtype=torch.float64, device="cuda")
    series = ch.steps(3, 5, alw, 31, 9).cpu().numpy()
    for l in range(len(Ls)):
        p = Problem(Hs[l], Gs[l], _lib.GFMC_FP32)
        set_method(p, method, lams[l])
        c = Chain(p, ws[l])
        s1 = c.steps(3, 5, float(alw[l].item()), 31 + l, 9).cpu().numpy()
        np.testing.assert_array_equal(series[:, l], s1)


@pytest.mark.parametrize("method", ["single_step", "multi_step"])
def test_discrete_gfmc_energy_matches_ed(method):
    say("discrete energy", method)
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(16, L.edges, beta=0.15, dtype=np.float64)
    e0 = ground_energy(H)
    w = prepare_walkers_sz(2048, 16, 1)
    _, _, _, st, raw = run(w, 2000, H, 6, G, [4, 32], method=method, lam=12.0, rng_key=4,
                           warmup_steps=400, verbosity=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    say(f"  E {E[32]:.5f} +- {err[32]:.5f} exact {e0:.5f}; {st['it_per_sec']:.0f} it/s")
    assert abs(E[32] - e0) < 4 * err[32] + 1e-3


def test_published_throughput_configuration_runs():
    """The reference's throughput benchmark configuration (Heisenberg ring of
    100 sites, no guide, Neel start, single_step, n_bra = 6, lam = 25): the
    run completes with a finite energy near the Bethe-ansatz value."""
    say("fig3 config")
    L = chain(100)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    neel = np.tile([1.0, -1.0], 50)
    w = prepare_walkers_from_configs(np.repeat(neel[None], 4096, axis=0))
    _, E, _, st, raw = run(w, 500, H, 6, None, [4, 16], method="single_step", lam=25.0,
                           rng_key=1, warmup_steps=500, verbosity=0)
    say(f"  E/L {E[16] / 100:.5f}, {st['it_per_sec'] * 4096:.3g} walker-steps/s")
    assert abs(E[16] / 100 + 0.4432296) < 0.01
