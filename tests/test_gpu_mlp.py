"""The two-layer MLP guide on the GPU: log psi, rates and local energies
against the oracle (the layer-1 activations are reproduced bit for bit, the
layer-2 sums differ by the FP32 accumulation of the MFMA), the property that
a configuration's value does not depend on the move that reached it,
propagation against the oracle, and GFMC energies against exact
diagonalisation."""
import sys

import numpy as np
import pytest
import torch

from helpers import ground_energy
from pygfmc_amd import (MLPSpinHalf, SpinHalf, chain, kagome, prepare_walkers_random,
                        prepare_walkers_sz, run)
from pygfmc_amd import _lib
from pygfmc_amd.engine import set_method
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def systems():
    out = []
    L = kagome(2, 2)
    out.append(("kagome2x2_64x32", SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32),
                MLPSpinHalf(12, (64, 32), seed=1, scale=1.5, out_scale=1.0)))
    L = chain(8)
    out.append(("chain8_hx_64x64", SpinHalf.create(L.edges, -1.0, -0.4, 0.7, h_x=0.6,
                                                   dtype=np.float32),
                MLPSpinHalf(8, (64, 64), seed=2, scale=1.5, out_scale=1.0)))
    L = chain(40)
    out.append(("chain40_128x128", SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32),
                MLPSpinHalf(40, (128, 128), seed=3, scale=1.5, out_scale=1.0)))
    L = kagome(4, 4)
    out.append(("kagome4x4_512x256", SpinHalf.create(L.edges, -1.0, -1.0, 1.0,
                                                     dtype=np.float32),
                MLPSpinHalf(48, (512, 256), seed=4, scale=1.5, out_scale=1.0)))
    return out


SYSTEMS = systems()
IDS = [s[0] for s in SYSTEMS]


def spins_for(orc, H, nw, seed):
    return orc.init_random(H.N, nw, seed) if H.n_hx else orc.init_sz(H.N, nw, H.N // 2, seed)


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_log_psi_matches_oracle(orc, name, H, G):
    say("mlp log psi", name)
    spins = spins_for(orc, H, 300, 1)
    want = orc.Problem(lower(H, G)).log_psi(spins)
    p = Problem(H, G, _lib.GFMC_FP32)
    got = p.log_psi(to_dev(spins)).cpu().numpy()
    err = np.abs(got - want)
    say("  max |d log psi|", err.max(), "exact", np.mean(err == 0))
    assert err.max() < 1e-5
    assert np.std(want) > 0.02


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_rates_and_local_energies_match_oracle(orc, name, H, G):
    say("mlp rates", name)
    spins = spins_for(orc, H, 200, 2)
    r_ref, e_ref, ld_ref = orc.Problem(lower(H, G)).rates(spins)
    p = Problem(H, G, _lib.GFMC_FP32)
    r, e, ld = p.rates(to_dev(spins), want_logdiff=True)
    r, e, ld = r.cpu().numpy(), e.cpu().numpy(), ld.cpu().numpy()
    nz = r_ref != 0
    np.testing.assert_array_equal(r != 0, nz)
    assert np.all(np.abs(r - r_ref) <= 2e-5 * np.abs(r_ref) + 1e-12)
    np.testing.assert_allclose(r[:, 0], r_ref[:, 0], rtol=1e-12, atol=1e-12)
    assert np.all(np.abs(e - e_ref) <= 2e-5 * (1 + np.abs(e_ref)))
    assert np.all(np.abs(ld[nz] - ld_ref[nz]) <= 1e-5)


def test_candidate_values_equal_direct_evaluations(orc):
    """log psi(x') - log psi(x) from the move rows equals the difference of
    the direct evaluations of x' and x, bit for bit (exact layer-1 sums,
    the same MFMA and reduction order for every row)."""
    name, H, G = SYSTEMS[3]
    say("mlp path independence", name)
    spins = spins_for(orc, H, 64, 3)
    p = Problem(H, G, _lib.GFMC_FP32)
    _, _, ld = p.rates(to_dev(spins), want_logdiff=True)
    ld = ld.cpu().numpy()
    lp = p.log_psi(to_dev(spins)).cpu().numpy().astype(np.float32)
    cfg = orc.unpack(spins, H.N)
    checked = 0
    for w in range(0, 64, 7):
        x = cfg[w]
        for e in range(0, len(H.edge_i), 5):
            i, j = int(H.edge_i[e]), int(H.edge_j[e])
            if x[i] == x[j]:
                continue
            y = x.copy()
            y[i], y[j] = -y[i], -y[j]
            lpy = p.log_psi(to_dev(orc.pack(y[None]))).cpu().numpy().astype(np.float32)[0]
            assert ld[w, 1 + e] == float(np.float32(lpy - lp[w]))
            checked += 1
    assert checked > 50


@pytest.mark.parametrize("name,H,G", SYSTEMS[:2], ids=IDS[:2])
def test_propagation_matches_oracle(orc, name, H, G):
    """Same Philox draws: the walkers whose FP32 rates drew the same moves
    as the FP64 oracle agree in configuration, evaluations, weight and local
    energy; nearly all do."""
    say("mlp propagate", name)
    nw = 300
    spins = spins_for(orc, H, nw, 4)
    lw0 = np.random.default_rng(0).normal(scale=0.1, size=nw)
    beta, alw, seed, step = 0.3, 0.2, 777, 5
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, lw0, np.zeros(nw), beta, alw, seed, step, 0)
    p = Problem(H, G, _lib.GFMC_FP32)
    s = to_dev(spins)
    lw = torch.as_tensor(lw0, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    p.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    same = np.all(s.cpu().numpy().view(np.uint32) == s_ref, axis=1) & (ev.cpu().numpy() ==
                                                                        ev_ref)
    say("  identical paths", same.mean(), "mean evals", ev_ref.mean())
    assert same.mean() > 0.97 and ev_ref.mean() > 2.0
    np.testing.assert_allclose(lw.cpu().numpy()[same], lw_ref[same], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(e.cpu().numpy()[same], e_ref[same], rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize("case", ["kagome2x2", "chain8_hx"])
def test_gfmc_energy_with_mlp_guide_matches_ed(case):
    say("mlp physics", case)
    if case == "kagome2x2":
        H, G = SYSTEMS[0][1], SYSTEMS[0][2]
        w = prepare_walkers_sz(2048, 12, 1)
    else:
        H, G = SYSTEMS[1][1], SYSTEMS[1][2]
        w = prepare_walkers_random(2048, 8, 1)
    E0 = ground_energy(H)
    _, E, _, st, raw = run(w, 1500, H, 0.2, G, [8, 32], rng_key=5, warmup_steps=200,
                           verbosity=0)
    from pygfmc_amd import evaluate_observables
    _, _, err, _ = evaluate_observables(raw, n_blocks=20)
    say("  E", E[32], "+-", err[32], "ED", E0, "hip_graph", st["hip_graph"])
    assert abs(E[32] - E0) < 4 * err[32] + 2e-3
    assert st["hip_graph"]


def test_captured_mlp_steps_equal_eager_steps():
    """The MLP-guided step loop captured into HIP graphs (a fixed number
    of hop iterations per step) gives bitwise the run of eager steps."""
    say("mlp graph vs eager")
    H, G = SYSTEMS[0][1], SYSTEMS[0][2]
    out = []
    for use_graph in (True, False):
        w = prepare_walkers_sz(512, 12, 3)
        wf, E, _, st, raw = run(w, 120, H, 0.2, G, [8], rng_key=11, warmup_steps=40,
                                verbosity=0, use_graph=use_graph, number_of_outputs=6)
        assert st["hip_graph"] == use_graph
        out.append((np.asarray(raw["log_weights"]), np.asarray(raw["energy"]),
                    wf.spins.cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_captured_mlp_step_over_its_hop_cap_is_an_error():
    """A captured step that needs more hop iterations than the cap raises
    after the replay (a graph captured with the cap set to 1 from the host;
    every kernel still ends normally)."""
    H, G = SYSTEMS[0][1], SYSTEMS[0][2]
    from pygfmc_amd.engine import Chain, StepGraph
    from pygfmc_amd.problem import get_problem
    p = get_problem(H, G)
    ch = Chain(p, prepare_walkers_sz(512, 12, 3))
    ch.steps(5, 0.2, 0.0, 11, 0)
    u0, mx = p.mlp_step_stats()
    assert mx >= 2
    p.set_mlp_hop_cap(1)
    try:
        g = StepGraph(ch, 3, 0.2, 0.0, 11)
        g.launch(5)
        torch.cuda.synchronize()
        assert p.mlp_step_stats()[0] == u0 + 3
        p.set_mlp_hop_cap(2 * mx + 16)
        g2 = StepGraph(ch, 3, 0.2, 0.0, 11)
        g2.launch(8)
        torch.cuda.synchronize()
        assert p.mlp_step_stats()[0] == u0 + 3
    finally:
        p.set_mlp_hop_cap(512)


def test_mlp_problem_rejects_what_it_does_not_run():
    name, H, G = SYSTEMS[0]
    p = Problem(H, G, _lib.GFMC_FP32)
    assert p.lanes_ok == [64]
    with pytest.raises(_lib.GfmcError):
        p.set_lanes(32)
    with pytest.raises(_lib.GfmcError):
        set_method(p, "single_step", 5.0)
    with pytest.raises(_lib.GfmcError):
        Problem(H, MLPSpinHalf(12, (64, 32), seed=1), _lib.GFMC_FP64)
