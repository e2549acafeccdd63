"""Oracle parity at the benchmark workloads' own shapes.

* XXZ on the 32 x 32 periodic square lattice (1024 sites, 2048 edges, K = 32
  slots per lane at 64 lanes), Delta in {0.5, 1, 2}, Jastrow beta 0.3: rates,
  one propagation and carried step loops against the oracle (FP64 bit-exact
  moves, 1e-12 weights; FP32 rates within 1e-5), at 64 and 32 lanes per
  walker, carried against re-evaluated bit for bit; at the full 65 536
  walkers a sample of walkers spread over the launch against the oracle with
  their own Philox counters.
* The uneven Ising-chain batch (64 chains, L ~ U{16..256}, h ~ U[0.5, 1.5]
  from seed 0, Jastrow beta -0.25): the batch kernels against per-lattice
  oracle chains (key seed + l) directly, at 64 and 16 lanes, and at 2048
  walkers per lattice a sample of every lattice.
* Kagome 12 x 12 (432 sites) with the 512 -> 256 MLP guide: log psi, rates
  and one propagation against the oracle, and at the full 262 144 walkers a
  sample spread over the launch.
"""
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import (JastrowSpinHalf, MLPSpinHalf, SpinHalf, chain, kagome,
                        prepare_walkers_random, prepare_walkers_sz, square,
                        transverse_field_ising)
from pygfmc_amd import _lib
from pygfmc_amd.batch import Batch, BatchChain
from pygfmc_amd.engine import Chain
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.walkers import Walker

pytestmark = pytest.mark.gpu

DELTAS = [0.5, 1.0, 2.0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def from_dev(t):
    return t.cpu().numpy().view(np.uint32)


def xxz32(delta, dtype):
    lat = square(32, 32)
    H = SpinHalf.create(lat.edges, -1.0, -1.0, delta, dtype=dtype)
    G = JastrowSpinHalf(1024, lat.edges, beta=0.3, dtype=dtype)
    return H, G


def sample_ids(n, k):
    """k walker indices spread over [0, n), the last walkers included."""
    ids = np.unique(np.concatenate([np.linspace(0, n - 1, k - 4).astype(np.int64),
                                    np.arange(n - 4, n)]))
    return ids


# -- XXZ 32 x 32 ----------------------------------------------------------------

@pytest.mark.parametrize("delta", DELTAS)
def test_xxz32_rates_match_oracle(orc, delta):
    say("xxz32 rates", delta)
    H, G = xxz32(delta, np.float64)
    spins = orc.init_sz(1024, 48, 512, 1)
    r_ref, e_ref, ld_ref = orc.Problem(lower(H, G)).rates(spins)
    p = Problem(H, G, _lib.GFMC_FP64)
    r, e, ld = p.rates(to_dev(spins), want_logdiff=True)
    np.testing.assert_allclose(r.cpu().numpy(), r_ref, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ld.cpu().numpy(), ld_ref, rtol=1e-12, atol=1e-13)
    p32 = Problem(H, G, _lib.GFMC_FP32)
    r, e, _ = p32.rates(to_dev(spins))
    scale = np.abs(r_ref).max()
    np.testing.assert_allclose(r.cpu().numpy(), r_ref, rtol=1e-5, atol=1e-5 * scale)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-5)


@pytest.mark.parametrize("lanes", [64, 32])
@pytest.mark.parametrize("delta", DELTAS)
def test_xxz32_propagation_fp64_bit_exact(orc, delta, lanes):
    say("xxz32 propagate fp64", delta, lanes)
    H, G = xxz32(delta, np.float64)
    p = Problem(H, G, _lib.GFMC_FP64)
    assert lanes in p.lanes_ok
    p.set_lanes(lanes)
    nw = 96
    spins = orc.init_sz(1024, nw, 512, 2)
    lw0 = np.random.default_rng(1).normal(scale=0.1, size=nw)
    beta, alw, seed, step = 0.2, -650.0 * delta, 0xC2C2C2C2, 11
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, lw0, np.zeros(nw), beta, alw, seed, step, 0, lanes=lanes)
    s = to_dev(spins)
    lw = torch.as_tensor(lw0, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    p.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ev.cpu().numpy(), ev_ref)
    np.testing.assert_array_equal(from_dev(s), s_ref)
    np.testing.assert_allclose(lw.cpu().numpy(), lw_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    assert ev_ref.mean() > 20


@pytest.mark.parametrize("prec", [_lib.GFMC_FP64, _lib.GFMC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("delta", DELTAS)
def test_xxz32_carried_step_loops(orc, delta, prec):
    """Step loops with carried slot weights equal loops that re-evaluate
    every slot (bit for bit, lane counts switched between calls); in FP64
    both equal the oracle chain (identical spins, series and local energies
    to 1e-12)."""
    say("xxz32 carry", delta, prec)
    dt = np.float64 if prec == _lib.GFMC_FP64 else np.float32
    H, G = xxz32(delta, dt)
    p = Problem(H, G, prec)
    nw = 128
    spins = orc.init_sz(1024, nw, 512, 3)
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    w = Walker.from_device(to_dev(spins), z, z.clone(), 1024)
    a, b = Chain(p, w), Chain(p, w)
    assert a.carry is not None
    b.carry = None
    alw, seed = -650.0 * delta, 321
    s_ref = spins
    prob = orc.Problem(lower(H, G))
    pos = 40
    for lanes, n in ((64, 3), (32, 2), (64, 2)):
        p.set_lanes(lanes)
        sa = a.steps(n, 0.2, alw, seed, pos, evals=True).cpu().numpy()
        sb = b.steps(n, 0.2, alw, seed, pos, evals=True).cpu().numpy()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(from_dev(a.spins), from_dev(b.spins))
        np.testing.assert_array_equal(a.e_loc.cpu().numpy(), b.e_loc.cpu().numpy())
        if prec == _lib.GFMC_FP64:
            s_ref, e_ref, ser_ref = orc.run_chain(prob, s_ref, n, 0.2, alw, seed, pos,
                                                  lanes=lanes)
            np.testing.assert_array_equal(from_dev(a.spins), s_ref)
            np.testing.assert_allclose(sa, ser_ref, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(a.e_loc.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
        pos += n
    np.testing.assert_array_equal(a.evals.cpu().numpy(), b.evals.cpu().numpy())


@pytest.mark.parametrize("delta", DELTAS)
def test_xxz32_full_population_sample_matches_oracle(orc, delta):
    """65 536 walkers in one launch: FP64 — the sampled walkers' moves,
    evaluation counts and weights equal the oracle's (same Philox counters);
    FP32 (the workload's precision) — nearly all sampled walkers draw the
    same moves as the FP64 oracle, and those agree in weight and local
    energy to 1e-5."""
    nw = 65536
    beta, alw, seed, step = 0.2, -650.0 * delta, 4711, 3
    ids = sample_ids(nw, 36)
    w = prepare_walkers_sz(nw, 1024, 17)
    allsp = from_dev(w.spins)
    for prec in (_lib.GFMC_FP64, _lib.GFMC_FP32):
        say("xxz32 full population", delta, prec)
        dt = np.float64 if prec == _lib.GFMC_FP64 else np.float32
        H, G = xxz32(delta, dt)
        p = Problem(H, G, prec)
        p.set_lanes(64)
        s = w.spins.clone()
        lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
        e = torch.zeros(nw, dtype=torch.float64, device="cuda")
        ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
        p.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
        torch.cuda.synchronize()
        H64, G64 = xxz32(delta, np.float64)
        s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H64, G64)).propagate_ids(
            ids, allsp[ids], np.zeros(len(ids)), np.zeros(len(ids)), beta, alw, seed, step, 0)
        got_s, got_ev = from_dev(s)[ids], ev.cpu().numpy()[ids]
        got_lw, got_e = lw.cpu().numpy()[ids], e.cpu().numpy()[ids]
        if prec == _lib.GFMC_FP64:
            np.testing.assert_array_equal(got_ev, ev_ref)
            np.testing.assert_array_equal(got_s, s_ref)
            np.testing.assert_allclose(got_lw, lw_ref, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(got_e, e_ref, rtol=1e-12, atol=1e-12)
        else:
            same = np.all(got_s == s_ref, axis=1) & (got_ev == ev_ref)
            say("  identical paths", same.mean())
            assert same.mean() >= 0.9
            scale = np.abs(e_ref).max() * beta
            np.testing.assert_allclose(got_lw[same], lw_ref[same], rtol=1e-5, atol=1e-5 * scale)
            np.testing.assert_allclose(got_e[same], e_ref[same], rtol=1e-5)
        assert ev_ref.mean() > 20


# -- the uneven Ising-chain batch --------------------------------------------------

def tfim_draw(dtype, n_lat=64):
    """The 64-chain draw: L ~ U{16..256}, h ~ U[0.5, 1.5] from seed 0,
    H = -J sum s^z s^z - h sum s^x (J = 1), Jastrow beta -0.25."""
    rng = np.random.default_rng(0)
    Ls = rng.integers(16, 257, 64)[:n_lat]
    hs = rng.uniform(0.5, 1.5, 64)[:n_lat]
    Hs = [transverse_field_ising(chain(int(L)).edges, 1.0, h, dtype=dtype) for L, h in zip(Ls, hs)]
    Gs = [JastrowSpinHalf(int(L), chain(int(L)).edges, beta=-0.25, dtype=dtype) for L in Ls]
    return [int(L) for L in Ls], hs, Hs, Gs


def test_tfim_draw_covers_the_range():
    Ls, hs, _, _ = tfim_draw(np.float64)
    assert len(Ls) == 64 and min(Ls) >= 16 and max(Ls) <= 256
    assert max(Ls) >= 250 and min(Ls) <= 24       # long and short chains both present
    assert np.all((hs >= 0.5) & (hs <= 1.5))


@pytest.mark.parametrize("lanes", [64, 16])
def test_tfim_batch_step_loop_matches_oracle_chains(orc, lanes):
    """All 64 chains in one batch, FP64: after each call of the batch step
    loop every lattice's spins equal its own oracle chain with key seed + l
    bit for bit, its series and local energies to 1e-12."""
    say("tfim batch vs oracle", lanes)
    Ls, hs, Hs, Gs = tfim_draw(np.float64)
    nw = 96
    walkers = [prepare_walkers_random(nw, L, 100 + l) for l, L in enumerate(Ls)]
    b = Batch(Hs, Gs, [nw] * len(Ls))
    assert lanes in b.lanes_ok
    b.set_lanes(lanes)
    ch = BatchChain(b, walkers)
    alw_h = np.array([-1.1 * L for L in Ls])
    alw = torch.as_tensor(alw_h, device="cuda")
    seed, pos = 202, 9
    probs = [orc.Problem(lower(H, G)) for H, G in zip(Hs, Gs)]
    refs = [from_dev(w.spins) for w in walkers]
    for n in (2, 2):
        ser = ch.steps(n, 0.2, alw, seed, pos).cpu().numpy()
        torch.cuda.synchronize()
        for l in range(len(Ls)):
            s_ref, e_ref, ser_ref = orc.run_chain(probs[l], refs[l], n, 0.2, alw_h[l], seed + l,
                                                  pos, lanes=lanes)
            refs[l] = s_ref
            wl = ch.walkers(l)
            np.testing.assert_array_equal(from_dev(wl.spins), s_ref, err_msg=f"lattice {l}")
            np.testing.assert_allclose(ser[:, l], ser_ref, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(wl.e_loc, e_ref, rtol=1e-12, atol=1e-12)
        pos += n


@pytest.mark.parametrize("prec", [_lib.GFMC_FP64, _lib.GFMC_FP32], ids=["fp64", "fp32"])
def test_tfim_batch_full_size_sample_matches_oracle(orc, prec):
    """The workload's size: 64 chains x 2048 walkers in one propagation
    launch (16 lanes per walker).  For every lattice, sampled walkers against
    the oracle with their own counters: FP64 identical moves and counts,
    weights to 1e-12; FP32 nearly all identical, those within 1e-5."""
    say("tfim batch full size", prec)
    dt = np.float64 if prec == _lib.GFMC_FP64 else np.float32
    Ls, hs, Hs, Gs = tfim_draw(dt)
    _, _, H64, G64 = tfim_draw(np.float64)
    nw = 2048
    walkers = [prepare_walkers_random(nw, L, 300 + l) for l, L in enumerate(Ls)]
    b = Batch(Hs, Gs, [nw] * len(Ls))
    b.set_lanes(16)
    ch = BatchChain(b, walkers)
    alw_h = np.array([-1.1 * L for L in Ls])
    alw = torch.as_tensor(alw_h, device="cuda")
    beta, seed, step = 0.2, 55, 4
    _lib.check(_lib.load().gfmc_batch_propagate(
        b.handle, _lib.ptr(ch.spins), _lib.ptr(ch.log_w), _lib.ptr(ch.e_loc), beta,
        _lib.ptr(alw), seed, step, _lib.TAG_PROPAGATE, _lib.ptr(ch.evals), _lib.stream()),
        "gfmc_batch_propagate")
    torch.cuda.synchronize()
    ids = sample_ids(nw, 10)
    n_same = n_all = 0
    for l in range(len(Ls)):
        a = b.offsets[l]
        wl = ch.walkers(l)
        got_s = from_dev(wl.spins)[ids]
        got_lw, got_e = wl.log_weight[ids], wl.e_loc[ids]
        got_ev = ch.evals[a:a + nw].cpu().numpy()[ids]
        s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H64[l], G64[l])).propagate_ids(
            ids, from_dev(walkers[l].spins)[ids], np.zeros(len(ids)), np.zeros(len(ids)), beta,
            alw_h[l], seed + l, step, 0, lanes=16)
        if prec == _lib.GFMC_FP64:
            np.testing.assert_array_equal(got_ev, ev_ref, err_msg=f"lattice {l}")
            np.testing.assert_array_equal(got_s, s_ref, err_msg=f"lattice {l}")
            np.testing.assert_allclose(got_lw, lw_ref, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(got_e, e_ref, rtol=1e-12, atol=1e-12)
        else:
            same = np.all(got_s == s_ref, axis=1) & (got_ev == ev_ref)
            n_same += int(same.sum())
            n_all += len(ids)
            scale = np.abs(e_ref).max() * beta
            np.testing.assert_allclose(got_lw[same], lw_ref[same], rtol=1e-5, atol=1e-5 * scale)
            np.testing.assert_allclose(got_e[same], e_ref[same], rtol=1e-5)
    if prec == _lib.GFMC_FP32:
        say("  identical paths", n_same / n_all)
        assert n_same / n_all >= 0.9


# -- kagome 12 x 12 with the two-layer MLP guide ------------------------------------

def kagome12_mlp():
    lat = kagome(12, 12)
    H = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    G = MLPSpinHalf(lat.N, (512, 256), seed=0, scale=1.5, out_scale=1.0)
    return H, G


def test_kagome12_mlp_log_psi_and_rates_match_oracle(orc):
    say("kagome12 mlp log psi / rates")
    H, G = kagome12_mlp()
    assert H.N == 432 and len(H.edge_i) == 864
    prob = orc.Problem(lower(H, G))
    p = Problem(H, G, _lib.GFMC_FP32)
    spins = orc.init_sz(432, 16, 216, 5)
    want = prob.log_psi(spins)
    got = p.log_psi(to_dev(spins)).cpu().numpy()
    say("  max |d log psi|", np.abs(got - want).max())
    assert np.abs(got - want).max() < 2e-5 and np.std(want) > 0.05
    r_ref, e_ref, ld_ref = prob.rates(spins[:8])
    r, e, ld = p.rates(to_dev(spins[:8]), want_logdiff=True)
    r, e, ld = r.cpu().numpy(), e.cpu().numpy(), ld.cpu().numpy()
    nz = r_ref != 0
    np.testing.assert_array_equal(r != 0, nz)
    assert np.all(np.abs(r - r_ref) <= 2e-5 * np.abs(r_ref) + 1e-12)
    np.testing.assert_allclose(r[:, 0], r_ref[:, 0], rtol=1e-12, atol=1e-12)
    assert np.all(np.abs(e - e_ref) <= 2e-5 * (1 + np.abs(e_ref)))
    assert np.all(np.abs(ld[nz] - ld_ref[nz]) <= 1e-5)


def test_kagome12_mlp_full_population_sample_matches_oracle(orc):
    """262 144 walkers in one MLP propagation: sampled walkers spread over
    the population (the last ones included) draw the same moves as the FP64
    oracle with their own counters, and agree in weight and local energy."""
    say("kagome12 mlp full population")
    H, G = kagome12_mlp()
    nw = 262144
    beta, alw, seed, step = 0.04, -200.0, 606, 2
    w = prepare_walkers_sz(nw, 432, 21)
    allsp = from_dev(w.spins)
    p = Problem(H, G, _lib.GFMC_FP32)
    s = w.spins.clone()
    lw = torch.zeros(nw, dtype=torch.float64, device="cuda")
    e = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ev = torch.zeros(nw, dtype=torch.int32, device="cuda")
    p.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    ids = sample_ids(nw, 16)
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate_ids(
        ids, allsp[ids], np.zeros(len(ids)), np.zeros(len(ids)), beta, alw, seed, step, 0)
    got_s, got_ev = from_dev(s)[ids], ev.cpu().numpy()[ids]
    same = np.all(got_s == s_ref, axis=1) & (got_ev == ev_ref)
    say("  identical paths", same.mean(), "mean evals", ev_ref.mean())
    assert same.mean() >= 0.85 and ev_ref.mean() > 2
    scale = np.abs(e_ref).max() * beta
    np.testing.assert_allclose(lw.cpu().numpy()[ids][same], lw_ref[same], rtol=1e-5,
                               atol=2e-5 * scale)
    np.testing.assert_allclose(e.cpu().numpy()[ids][same], e_ref[same], rtol=2e-5)
