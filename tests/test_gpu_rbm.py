"""RBM guide on the GPU: FP32 kernels on BF16 weights against the FP64
oracle, ratio consistency, and a GFMC run against exact diagonalisation."""
import sys
import warnings

import numpy as np
import pytest
import torch

from helpers import ground_energy, random_rbm
from pygfmc_amd import SpinHalf, chain, evaluate_observables, prepare_walkers_sz, run, square
from pygfmc_amd import _lib
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu

# FP32 arithmetic over H hidden units: errors of a few 1e-7 per unit add up
# to ~1e-5 relative in a rate; the log-ratios are compared absolutely.
RTOL = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(spins.view(np.int32), device="cuda")


def rbm_systems():
    out = []
    L = chain(12)
    H = SpinHalf.create(L.edges, -1.0, -0.6, 0.7, h_x=0.4, dtype=np.float64)
    out.append(("chain12_hx", H, random_rbm(12, alpha=2, seed=1)))
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square4_heis", H, random_rbm(16, alpha=4, seed=2, scale=0.2)))
    L = square(10, 10)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square10_alpha4", H, random_rbm(100, alpha=4, seed=3, scale=0.05)))
    # wider than the lanes-over-hidden-units kernels take (H > 512): the
    # lanes-over-slots kernels
    L = chain(12)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("chain12_wide", H, random_rbm(12, alpha=48, seed=4, scale=0.1)))
    return out


SYSTEMS = rbm_systems()
IDS = [s[0] for s in SYSTEMS]


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_rbm_rates_match_oracle(orc, name, H, G):
    say("rbm rates", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    spins = orc.init_sz(H.N, 96, H.N // 2, 7) if H.n_hx == 0 else orc.init_random(H.N, 96, 7)
    r_ref, e_ref, ld_ref = orc.Problem(lower(H, G)).rates(spins)
    r, e, ld = prob.rates(to_dev(spins), want_logdiff=True)
    torch.cuda.synchronize()
    r, e, ld = r.cpu().numpy(), e.cpu().numpy(), ld.cpu().numpy()
    np.testing.assert_allclose(ld, ld_ref, atol=3e-5)
    np.testing.assert_allclose(r, r_ref, rtol=RTOL, atol=RTOL * np.abs(r_ref).max())
    np.testing.assert_allclose(e, e_ref, rtol=RTOL, atol=RTOL * np.abs(e_ref).max())


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_rbm_log_psi_matches_oracle_and_ratios(orc, name, H, G):
    say("rbm log psi", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    spins = orc.init_random(H.N, 64, 9)
    ref = orc.Problem(lower(H, G)).log_psi(spins)
    lp = prob.log_psi(to_dev(spins)).cpu().numpy()
    np.testing.assert_allclose(lp, ref, rtol=1e-5, atol=1e-4)
    # the kernel's ratios against its own log psi of the successors
    x = orc.unpack(spins[:1], H.N)[0]
    succ = H.successors(x)
    lp_s = prob.log_psi(to_dev(orc.pack(succ))).cpu().numpy()
    _, _, ld = prob.rates(to_dev(spins[:1]), want_rates=False, want_eloc=False,
                          want_logdiff=True)
    np.testing.assert_allclose(ld.cpu().numpy()[0, 1:], lp_s[1:] - lp_s[0], atol=2e-4)


def test_rbm_propagation_statistics_match_fp64_oracle(orc):
    """One step from the same walkers: the FP32 RBM kernel and the FP64
    oracle take the same path for most walkers (same uniforms), with the
    same local energies to FP32 accuracy where they do."""
    name, H, G = SYSTEMS[1]
    prob = Problem(H, G, _lib.GFMC_FP32)
    nw = 512
    spins = orc.init_sz(16, nw, 8, 11)
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, np.zeros(nw), np.zeros(nw), 0.2, 0.3, 77, 3, 0)
    s = to_dev(spins)
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    prob.propagate(s, lw, e, 0.2, 0.3, 77, 3, 0, ev)
    torch.cuda.synchronize()
    got = s.cpu().numpy().view(np.uint32)
    same = np.all(got == s_ref, axis=1)
    say("  same path:", same.mean(), "evals", ev_ref.mean())
    assert same.mean() > 0.95
    np.testing.assert_array_equal(ev.cpu().numpy()[same], ev_ref[same])
    np.testing.assert_allclose(e.cpu().numpy()[same], e_ref[same], rtol=RTOL,
                               atol=RTOL * np.abs(e_ref).max())


def test_rbm_guided_gfmc_energy_matches_ed():
    """Any positive guide gives the exact ground energy: 4x4 Heisenberg with
    a random small RBM guide, within 4 sigma + 2e-3 of ED."""
    say("rbm ED square4")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    e0 = ground_energy(H)
    G = random_rbm(16, alpha=4, seed=5, scale=0.05)
    w = prepare_walkers_sz(1024, 16, 1)
    _, E, _, stats, raw = run(w, 2000, H, 0.2, G, [4, 32], rng_key=4, warmup_steps=400,
                              verbosity=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    say("  E", E[32], "+-", err[32], "exact", e0, "it/s", stats["it_per_sec"])
    assert abs(E[32] - e0) < 4 * err[32] + 2e-3


@pytest.mark.parametrize("which", [2, 3], ids=["square10_alpha4", "chain12_wide"])
def test_rbm_propagation_matches_fp64_oracle_at_size(orc, which):
    """The 10 x 10 alpha = 4 configuration's shape (and the wide fallback):
    one propagation from the same walkers with the same uniforms, most
    walkers on the oracle's path with its evaluation counts and local
    energies."""
    name, H, G = SYSTEMS[which]
    say("rbm propagate", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    nw = 192
    spins = orc.init_sz(H.N, nw, H.N // 2, 13)
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, np.zeros(nw), np.zeros(nw), 0.2, -50.0, 91, 5, 0)
    s = to_dev(spins)
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    prob.propagate(s, lw, e, 0.2, -50.0, 91, 5, 0, ev)
    torch.cuda.synchronize()
    got = s.cpu().numpy().view(np.uint32)
    same = np.all(got == s_ref, axis=1) & (ev.cpu().numpy() == ev_ref)
    say("  same path:", same.mean(), "evals", ev_ref.mean())
    assert same.mean() > 0.9 and ev_ref.mean() > 2
    np.testing.assert_allclose(e.cpu().numpy()[same], e_ref[same], rtol=RTOL,
                               atol=RTOL * np.abs(e_ref).max())
    scale = np.abs(e_ref).max() * 0.2
    np.testing.assert_allclose(lw.cpu().numpy()[same], lw_ref[same], rtol=RTOL, atol=RTOL * scale)


# ---- translation-tied RBM on a torus: the table-factor kernel ---------------

def tied_systems():
    from pygfmc_amd import RBMSpinHalf
    rng = np.random.default_rng(11)
    out = []
    L = square(10, 10)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square10_tied", H,
                RBMSpinHalf.symmetric(L.extras["translations"], 4, b=rng.normal(scale=0.2, size=4),
                                      a=0.05, seed=3, scale=0.05)))
    L = square(6, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 0.7, dtype=np.float64)
    out.append(("square6x4_xxz_tied", H,
                RBMSpinHalf.symmetric(L.extras["translations"], 3, b=rng.normal(scale=0.2, size=3),
                                      a=-0.1, seed=5, scale=0.2)))
    L = square(10, 10)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("square10_from_jastrow", H,
                RBMSpinHalf.from_jastrow(L.edges, L.extras["translations"], 0.15, 4, seed=1,
                                         scale=0.01)))
    return out


TIED = tied_systems()
TIED_IDS = [s[0] for s in TIED]


def propagate_once(prob, spins, beta, alw, seed, step):
    nw = spins.shape[0]
    s = to_dev(spins)
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    prob.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    return (s.cpu().numpy().view(np.uint32), lw.cpu().numpy(), e.cpu().numpy(),
            ev.cpu().numpy())


@pytest.mark.parametrize("name,H,G", TIED, ids=TIED_IDS)
def test_tied_rbm_propagation_matches_fp64_oracle(orc, name, H, G):
    """The table-factor kernel against the FP64 oracle: one propagation from
    the same walkers with the same uniforms, most walkers on the oracle's
    path with its evaluation counts, local energies and log weights."""
    say("tied rbm propagate", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    assert prob.tied is not None
    nw = 192
    spins = orc.init_sz(H.N, nw, H.N // 2, 13)
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, np.zeros(nw), np.zeros(nw), 0.2, -50.0, 91, 5, 0)
    s, lw, e, ev = propagate_once(prob, spins, 0.2, -50.0, 91, 5)
    same = np.all(s == s_ref, axis=1) & (ev == ev_ref)
    say("  same path:", same.mean(), "evals", ev_ref.mean())
    assert same.mean() > 0.9 and ev_ref.mean() > 2
    np.testing.assert_allclose(e[same], e_ref[same], rtol=RTOL, atol=RTOL * np.abs(e_ref).max())
    scale = np.abs(e_ref).max() * 0.2
    np.testing.assert_allclose(lw[same], lw_ref[same], rtol=RTOL, atol=RTOL * scale)


@pytest.mark.parametrize("name,H,G", TIED, ids=TIED_IDS)
def test_tied_rbm_matches_untied_kernel(name, H, G):
    """Tied and untied kernels evaluate the same factors (bit-identical lane
    products) and differ only in the order of one FP32 sum over lanes: the
    same moves for nearly every walker, energies to FP32 rounding."""
    say("tied vs untied", name)
    pt = Problem(H, G, _lib.GFMC_FP32)
    pu = Problem(H, G, _lib.GFMC_FP32, tied=False)
    assert pt.tied is not None and pu.tied is None
    from oracle import init_sz
    spins = init_sz(H.N, 1024, H.N // 2, 21)
    a = propagate_once(pt, spins, 0.2, -50.0, 7, 3)
    b = propagate_once(pu, spins, 0.2, -50.0, 7, 3)
    same = np.all(a[0] == b[0], axis=1) & (a[3] == b[3])
    say("  same path:", same.mean())
    assert same.mean() > 0.99
    np.testing.assert_allclose(a[2][same], b[2][same], rtol=1e-5,
                               atol=1e-5 * np.abs(b[2]).max())
    np.testing.assert_allclose(a[1][same], b[1][same], rtol=1e-5,
                               atol=1e-5 * np.abs(b[2]).max())


def test_tied_tables_follow_set_params():
    """gfmc_rbm_set_params on a tied problem rebuilds the factor tables: the
    propagation then equals an untied problem built with the new
    parameters."""
    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.training import flat_params
    say("tied set_params")
    L = square(6, 6)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    T = L.extras["translations"]
    G0 = RBMSpinHalf.symmetric(T, 2, seed=1, scale=0.1)
    G1 = RBMSpinHalf.symmetric(T, 2, b=[0.3, -0.2], a=0.1, seed=2, scale=0.25)
    pt = Problem(H, G0, _lib.GFMC_FP32)
    assert pt.tied is not None
    flat = torch.as_tensor(flat_params(G1), device="cuda")
    _lib.check(_lib.load().gfmc_rbm_set_params(pt.handle, _lib.ptr(flat), _lib.stream()),
               "gfmc_rbm_set_params")
    pu = Problem(H, G1, _lib.GFMC_FP32, tied=False)
    from oracle import init_sz
    spins = init_sz(36, 512, 18, 4)
    a = propagate_once(pt, spins, 0.2, -20.0, 5, 1)
    b = propagate_once(pu, spins, 0.2, -20.0, 5, 1)
    same = np.all(a[0] == b[0], axis=1) & (a[3] == b[3])
    say("  same path:", same.mean())
    assert same.mean() > 0.99
    np.testing.assert_allclose(a[2][same], b[2][same], rtol=1e-5,
                               atol=1e-5 * np.abs(b[2]).max())


def test_tied_rbm_gfmc_energy_matches_ed():
    """Any positive guide gives the exact ground energy: 4 x 4 Heisenberg
    with a tied RBM guide on the table-factor kernel, within 4 sigma + 2e-3
    of ED."""
    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.problem import get_problem
    say("tied rbm ED square4")
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    e0 = ground_energy(H)
    G = RBMSpinHalf.from_jastrow(L.edges, L.extras["translations"], 0.2, 3, seed=2, scale=0.05)
    assert get_problem(H, G).tied is not None
    w = prepare_walkers_sz(1024, 16, 1)
    _, E, _, stats, raw = run(w, 2000, H, 0.2, G, [4, 32], rng_key=4, warmup_steps=400,
                              verbosity=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    say("  E", E[32], "+-", err[32], "exact", e0, "it/s", stats["it_per_sec"])
    assert abs(E[32] - e0) < 4 * err[32] + 2e-3
