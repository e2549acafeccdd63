"""Two and four walkers per wavefront (32 / 16 lanes per walker): the
propagation kernel against the oracle run with the same blocked order, the
step loop, batches, and the autotuner that picks the lane count."""
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import (JastrowSpinHalf, SpinHalf, chain, prepare_walkers_random, run, square,
                        transverse_field_ising)
from pygfmc_amd import _lib
from pygfmc_amd.autotune import tune_problem
from pygfmc_amd.batch import Batch, BatchChain
from pygfmc_amd.engine import Chain
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.walkers import Walker

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def systems():
    out = []
    L = chain(100)
    out.append(("heis_chain100", SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64),
                JastrowSpinHalf(100, L.edges, beta=0.3, dtype=np.float64)))
    L = chain(40)
    rng = np.random.default_rng(3)
    H = SpinHalf.create(L.edges, -1.0, -0.6, rng.uniform(-1.0, 2.0, 40),
                        h_z=rng.uniform(-0.5, 0.5, 40), h_x=0.4, dtype=np.float64)
    out.append(("xyz_chain40_fields", H, JastrowSpinHalf(40, L.edges, beta=0.2, h=0.05,
                                                         dtype=np.float64)))
    L = square(6, 6)
    out.append(("xxz_square6", SpinHalf.create(L.edges, -1.0, -1.0, 1.5, dtype=np.float64),
                JastrowSpinHalf(36, L.edges, beta=0.25, dtype=np.float64)))
    L = chain(24)
    out.append(("tfim_chain24", transverse_field_ising(L.edges, 1.0, 0.9, dtype=np.float64),
                JastrowSpinHalf(24, L.edges, beta=-0.2, dtype=np.float64)))
    return out


SYSTEMS = systems()
IDS = [s[0] for s in SYSTEMS]


# a move on the square lattice with a nearest-neighbour Jastrow guide changes
# more than 16 slots, so four walkers per wavefront do not fit there
CASES = [(s, g) for s in SYSTEMS for g in (32, 16) if not (s[0] == "xxz_square6" and g == 16)]


def test_lane_counts_offered_per_system():
    expect = {"heis_chain100": [64, 32, 16], "xyz_chain40_fields": [64, 32, 16],
              "xxz_square6": [64, 32], "tfim_chain24": [64, 32, 16]}
    for name, H, G in SYSTEMS:
        assert Problem(H, G, _lib.GFMC_FP64).lanes_ok == expect[name], name


@pytest.mark.parametrize("name,H,G,lanes", [(*s, g) for s, g in CASES],
                         ids=[f"{s[0]}-{g}" for s, g in CASES])
def test_grouped_propagation_bit_exact_vs_oracle(orc, name, H, G, lanes):
    say("grouped propagate", name, lanes)
    prob = Problem(H, G, _lib.GFMC_FP64)
    prob.set_lanes(lanes)
    nw = 203    # not a multiple of the walkers per wavefront
    spins = orc.init_random(H.N, nw, 4) if H.n_hx else orc.init_sz(H.N, nw, H.N // 2, 4)
    lw0 = np.random.default_rng(0).normal(scale=0.1, size=nw)
    beta, alw, seed, step = 0.2, 0.37, 0xDEADBEEF12345, 7
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, lw0, np.zeros(nw), beta, alw, seed, step, 0, lanes=lanes)
    s = to_dev(spins)
    lw = torch.as_tensor(lw0, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    prob.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ev.cpu().numpy(), ev_ref)
    np.testing.assert_array_equal(s.cpu().numpy().view(np.uint32), s_ref)
    np.testing.assert_allclose(lw.cpu().numpy(), lw_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    assert ev_ref.mean() > 1.0


def test_grouped_step_loop_matches_oracle_chain(orc):
    name, H, G = SYSTEMS[1]
    say("grouped step loop", name)
    prob = Problem(H, G, _lib.GFMC_FP64)
    prob.set_lanes(16)
    nw = 256
    spins = orc.init_random(H.N, nw, 8)
    n, beta, alw, seed, step0 = 5, 0.2, 0.1, 4242, 100
    s_ref, e_ref, ser_ref = orc.run_chain(orc.Problem(lower(H, G)), spins, n, beta, alw, seed,
                                          step0, lanes=16)
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ch = Chain(prob, Walker.from_device(to_dev(spins), z, z.clone(), H.N))
    ser = ch.steps(n, beta, alw, seed, step0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ch.spins.cpu().numpy().view(np.uint32), s_ref)
    np.testing.assert_allclose(ser.cpu().numpy(), ser_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ch.e_loc.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lanes", [32, 16])
def test_grouped_batch_equals_grouped_single_runs(lanes):
    say("grouped batch", lanes)
    Ls = [16, 22, 40, 60]
    Hs = [transverse_field_ising(chain(L).edges, 1.0, 0.7 + 0.2 * i, dtype=np.float32)
          for i, L in enumerate(Ls)]
    Gs = [JastrowSpinHalf(L, chain(L).edges, beta=-0.2, dtype=np.float32) for L in Ls]
    nws = [64, 132, 1024, 8]
    ws = [prepare_walkers_random(nw, L, i) for i, (L, nw) in enumerate(zip(Ls, nws))]
    b = Batch(Hs, Gs, nws)
    assert lanes in b.lanes_ok
    b.set_lanes(lanes)
    ch = BatchChain(b, ws)
    alw = torch.tensor([0.1, 0.2, 0.3, 0.4], dtype=torch.float64, device="cuda")
    series = ch.steps(4, 0.2, alw, 31, 9).cpu().numpy()
    for l in range(len(Ls)):
        p = Problem(Hs[l], Gs[l], _lib.GFMC_FP32)
        p.set_lanes(lanes)
        c = Chain(p, ws[l])
        s1 = c.steps(4, 0.2, float(alw[l].item()), 31 + l, 9).cpu().numpy()
        np.testing.assert_array_equal(series[:, l], s1)
        np.testing.assert_array_equal(ch.walkers(l).spins.cpu().numpy(), c.spins.cpu().numpy())


def test_batch_rejects_walker_counts_the_groups_cannot_tile():
    Hs = [transverse_field_ising(chain(16).edges, 1.0, 0.8, dtype=np.float32)]
    Gs = [JastrowSpinHalf(16, chain(16).edges, beta=-0.2, dtype=np.float32)]
    b = Batch(Hs, Gs, [66])
    assert 32 in b.lanes_ok and 16 not in b.lanes_ok
    with pytest.raises(_lib.GfmcError):
        b.set_lanes(16)


def test_generic_path_runs_one_walker_per_wavefront():
    L = chain(12)
    H = SpinHalf.create(L.edges, -1.0 - 0.1 * np.arange(12), -0.5, 1.0, dtype=np.float64)
    p = Problem(H, JastrowSpinHalf(12, L.edges, beta=0.2), _lib.GFMC_FP64)
    assert p.lanes_ok == [64]
    with pytest.raises(_lib.GfmcError):
        p.set_lanes(16)


def test_autotuner_picks_a_usable_lane_count_and_run_reports_it():
    say("autotune")
    L = chain(16)
    H = transverse_field_ising(L.edges, 1.0, 1.1, dtype=np.float32)
    G = JastrowSpinHalf(16, L.edges, beta=-0.2, dtype=np.float32)
    p = Problem(H, G, _lib.GFMC_FP32)
    w = prepare_walkers_random(4096, 16, 1)
    best, times = tune_problem(p, w.spins, 0.2, 0.0)
    say("  times", times, "->", best)
    assert best[0] in p.lanes_ok and p.lanes == best[0]
    assert best[1] in p.wpb_ok and p.launch == tuple(best[1:])
    assert len(times) == len(p.lanes_ok) * len(p.wpb_ok) * 2
    _, E, _, st, _ = run(w, 200, H, 0.2, G, [4, 16], rng_key=3, warmup_steps=40, verbosity=0)
    assert st["parameters/lanes_per_walker"] in (64, 32, 16)


@pytest.mark.parametrize("lanes", [64, 32, 16])
def test_launch_shape_does_not_change_results(lanes):
    """Wavefronts per workgroup and the XCD-contiguous walker mapping change
    only where walkers run: carried production steps on the 32 x 32 XXZ
    workload (10 KiB of LDS per walker, so four 32- or 16-lane wavefronts
    exceed 64 KiB per workgroup and the launch takes fewer) are bitwise
    those of one wavefront per workgroup, and every walker is propagated."""
    say("launch shapes, lanes", lanes)
    L = square(32, 32)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 2.0, dtype=np.float32)
    G = JastrowSpinHalf(1024, L.edges, beta=0.3, dtype=np.float32)
    p = Problem(H, G, _lib.GFMC_FP32)
    if lanes not in p.lanes_ok:
        pytest.skip(f"{lanes} lanes per walker not usable here")
    p.set_lanes(lanes)
    assert p.wpb_ok == (1, 2, 4)
    from oracle import init_sz
    nw = 1000      # not a multiple of the walkers per workgroup
    spins = init_sz(1024, nw, 512, 31)
    res = {}
    for wpb in (1, 2, 4):
        for xcd in (0, 1):
            p.set_launch(wpb, xcd)
            z = torch.zeros(nw, dtype=torch.float64, device="cuda")
            ch = Chain(p, Walker.from_device(to_dev(spins), z, z.clone(), 1024))
            out = torch.empty((3, 2), dtype=torch.float64, device="cuda")
            ch.steps(3, 0.2, -700.0, 41, 0, series=out, evals=True)
            torch.cuda.synchronize()
            ev = ch.evals.cpu().numpy()
            assert ev.min() > 0, (wpb, xcd)
            res[(wpb, xcd)] = (ch.spins.cpu().numpy(), ch.log_w.cpu().numpy(),
                               ch.e_loc.cpu().numpy(), ev, out.cpu().numpy())
    p.set_launch(1, 0)
    ref = res[(1, 0)]
    for k, v in res.items():
        for a, b in zip(v, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"launch {k}")
