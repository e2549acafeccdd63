"""GPU kernels against the CPU oracle on the same inputs (bit-exact for
integers, 1e-12 relative in FP64, 1e-5 in FP32)."""
import sys

import numpy as np
import pytest
import torch

from helpers import golden_system, random_model
from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain, kagome, square
from pygfmc_amd import _lib
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                    "golden_spinhalf.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev():
    return torch.device("cuda")


def to_dev(spins):
    return torch.as_tensor(spins.view(np.int32), device=dev())


def from_dev(t):
    return t.cpu().numpy().view(np.uint32)


def systems():
    """(name, H, G) of the parity cases."""
    out = []
    H, G = golden_system()
    out.append(("golden_jastrow", H, G))
    out.append(("golden_none", H, None))
    for seed, lat, pairs in ((0, "chain", "edges"), (1, "chain", "nnn"), (2, "kagome", "edges")):
        H, G = random_model(N=8 if lat == "chain" else 12, seed=seed, lattice=lat,
                            guide_pairs=pairs)
        out.append((f"random_{lat}_{pairs}_{seed}", H, G))
    L = chain(100)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("heisenberg_chain100", H, JastrowSpinHalf(100, L.edges, beta=0.3,
                                                          dtype=np.float64)))
    L = square(8, 8)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 2.0, dtype=np.float64)
    out.append(("xxz_square8", H, JastrowSpinHalf(64, L.edges, beta=0.3)))
    L = chain(70)
    rng = np.random.default_rng(5)
    H = SpinHalf.create(L.edges, J_z=-4.0, h_x=2 * rng.uniform(0.5, 1.5, 70), dtype=np.float64)
    out.append(("tfim_chain70", H, JastrowSpinHalf(70, L.edges, beta=0.1)))
    L = kagome(3, 3)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    out.append(("kagome3", H, JastrowSpinHalf(27, L.edges, beta=0.25)))
    return out


SYSTEMS = systems()
IDS = [s[0] for s in SYSTEMS]


def random_spins(orc, N, nw, seed, sz=None):
    if sz is None:
        return orc.init_random(N, nw, seed)
    return orc.init_sz(N, nw, (N + sz) // 2, seed)


def wave_tree_inputs(seed):
    """64 values of mixed magnitude: reassociating their sum changes the
    rounded result, so only the exact tree reproduces the oracle bitwise."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, 64) * 10.0 ** rng.integers(-8, 8, 64)


def test_wave_tree_inputs_distinguish_summation_orders(orc):
    v = wave_tree_inputs(0)
    tree = orc.wave_tree(v)
    seq = np.cumsum(v)
    assert np.any(tree != seq)                    # a sequential scan would be caught


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_wave_scan_matches_oracle_tree_bitwise(orc, prec):
    say("wave tree", prec)
    L = _lib.load()
    for seed in range(5):
        v = wave_tree_inputs(seed)
        if prec == "fp32":
            v = v.astype(np.float32).astype(np.float64)
        src = torch.as_tensor(v, device=dev())
        out = torch.empty(64, dtype=torch.float64, device=dev())
        _lib.check(L.gfmc_wave_scan(0 if prec == "fp64" else 1, _lib.ptr(src), _lib.ptr(out),
                                    _lib.stream()), "gfmc_wave_scan")
        got = out.cpu().numpy()
        if prec == "fp64":
            np.testing.assert_array_equal(got, orc.wave_tree(v))
        else:
            ref = wave_tree_f32(v)
            np.testing.assert_array_equal(got.astype(np.float32), ref)


def wave_tree_f32(v):
    """The wave tree evaluated in float32 (numpy), for the FP32 kernels."""
    s = np.asarray(v, dtype=np.float32).copy()
    for d in (1, 2, 4, 8):
        t = s.copy()
        for l in range(64):
            if (l & 15) >= d:
                s[l] = np.float32(t[l] + t[l - d])
    t = s.copy()
    for l in range(64):
        if (l >> 4) in (1, 3):
            s[l] = np.float32(t[l] + t[(l >> 4) * 16 - 1])
    t = s.copy()
    for l in range(32, 64):
        s[l] = np.float32(t[l] + t[31])
    return s


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_rates_fp64_match_oracle(orc, name, H, G):
    say("rates fp64", name)
    prob = Problem(H, G, _lib.GFMC_FP64)
    spins = random_spins(orc, H.N, 96, seed=1)
    r_ref, e_ref, ld_ref = orc.Problem(lower(H, G)).rates(spins)
    r, e, ld = prob.rates(to_dev(spins), want_logdiff=True)
    torch.cuda.synchronize()
    np.testing.assert_allclose(r.cpu().numpy(), r_ref, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ld.cpu().numpy(), ld_ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_rates_fp32_within_1e5(orc, name, H, G):
    say("rates fp32", name)
    prob = Problem(H, G, _lib.GFMC_FP32)
    spins = random_spins(orc, H.N, 96, seed=2)
    r_ref, e_ref, _ = orc.Problem(lower(H, G)).rates(spins)
    r, e, _ = prob.rates(to_dev(spins))
    torch.cuda.synchronize()
    scale = np.abs(r_ref).max()
    np.testing.assert_allclose(r.cpu().numpy(), r_ref, rtol=1e-5, atol=1e-5 * scale)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-5, atol=1e-5 * np.abs(e_ref).max())


def test_golden_local_energies_on_gpu(orc):
    g = np.load(GOLDEN, allow_pickle=False)
    for guide in ("none", "jastrow"):
        H, G = golden_system(guide=guide == "jastrow")
        prob = Problem(H, G, _lib.GFMC_FP64)
        for method in ("cont_time", "multi_step", "single_step"):
            say("golden", guide, method)
            c = g[f"spinhalf_{guide}_{method}/config"]
            _, e, _ = prob.rates(to_dev(orc.pack(c)), want_rates=False)
            np.testing.assert_allclose(e.cpu().numpy(), g[f"spinhalf_{guide}_{method}/e_loc"],
                                       rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name,H,G", SYSTEMS[:4], ids=IDS[:4])
def test_log_psi_matches_oracle(orc, name, H, G):
    if G is None:
        pytest.skip("no guide")
    say("log_psi", name)
    prob = Problem(H, G, _lib.GFMC_FP64)
    spins = random_spins(orc, H.N, 64, seed=3)
    ref = orc.Problem(lower(H, G)).log_psi(spins)
    out = prob.log_psi(to_dev(spins)).cpu().numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("name,H,G", SYSTEMS, ids=IDS)
def test_propagate_fp64_bit_exact_moves(orc, name, H, G, full):
    """Same Philox counters -> identical configurations and evaluation counts;
    weights and local energies to 1e-12."""
    say("propagate fp64", name, "full" if full else "incremental")
    prob = Problem(H, G, _lib.GFMC_FP64, full=full)
    nw = 128
    spins = random_spins(orc, H.N, nw, seed=4)
    rng = np.random.default_rng(0)
    lw0 = rng.normal(scale=0.1, size=nw)
    beta, alw, seed, step = 0.2, 0.37, 0xDEADBEEF12345, 7
    s_ref, lw_ref, e_ref, ev_ref = orc.Problem(lower(H, G)).propagate(
        spins, lw0, np.zeros(nw), beta, alw, seed, step, 0)
    s = to_dev(spins)
    lw = torch.as_tensor(lw0, device=dev())
    e = torch.zeros(nw, dtype=torch.float64, device=dev())
    ev = torch.zeros(nw, dtype=torch.int32, device=dev())
    prob.propagate(s, lw, e, beta, alw, seed, step, 0, ev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ev.cpu().numpy(), ev_ref)
    np.testing.assert_array_equal(from_dev(s), s_ref)
    np.testing.assert_allclose(lw.cpu().numpy(), lw_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
    assert ev_ref.mean() > 1.0          # moves really happened


def test_propagate_fp32_statistics_match_fp64(orc):
    """FP32 ratios: the same step's mean local energy and evaluation count
    agree with FP64 within sampling error."""
    L = chain(100)
    H64 = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(100, L.edges, beta=0.3)
    nw = 4096
    spins = random_spins(orc, 100, nw, seed=6, sz=0)
    out = {}
    for prec in (_lib.GFMC_FP64, _lib.GFMC_FP32):
        say("propagate stats", prec)
        prob = Problem(H64, G, prec)
        s = to_dev(spins)
        lw = torch.zeros(nw, dtype=torch.float64, device=dev())
        e = torch.zeros(nw, dtype=torch.float64, device=dev())
        ev = torch.zeros(nw, dtype=torch.int32, device=dev())
        prob.propagate(s, lw, e, 0.2, 0.0, 99, 0, 0, ev)
        torch.cuda.synchronize()
        out[prec] = (e.cpu().numpy(), ev.cpu().numpy(), from_dev(s))
    e64, ev64, s64 = out[_lib.GFMC_FP64]
    e32, ev32, s32 = out[_lib.GFMC_FP32]
    # the same uniforms: most walkers take the identical path
    same = np.all(s64 == s32, axis=1)
    assert same.mean() > 0.95
    np.testing.assert_allclose(e32[same], e64[same], rtol=1e-5, atol=1e-5 * 50)
    assert np.array_equal(ev32[same], ev64[same])


@pytest.mark.parametrize("n", [1, 3, 64, 1000, 1023, 1025, 70000])
def test_reconfigure_matches_oracle(orc, n):
    say("reconfigure", n)
    from pygfmc_amd.engine import reconfiguration
    from pygfmc_amd.walkers import Walker
    rng = np.random.default_rng(n)
    lw = rng.normal(scale=1.5, size=n)
    lw[rng.uniform(size=n) < 0.05] = -np.inf
    if not np.isfinite(lw).any():
        lw[0] = 0.0
    e = rng.normal(size=n)
    N = 40
    spins = orc.init_random(N, n, seed=n)
    w = Walker.from_device(to_dev(spins), torch.as_tensor(lw, device=dev()),
                           torch.as_tensor(e, device=dev()), N)
    seed, step = 1234, 17
    psi = orc.reconfig_psi(seed, step, 1)
    par_ref, lg_ref, E_ref, _ = orc.reconfigure(lw, e, psi)
    w2, lg, En, par = reconfiguration(w, (seed, step))
    np.testing.assert_array_equal(par, par_ref)
    assert lg == pytest.approx(lg_ref, rel=1e-13, abs=1e-13)
    assert En == pytest.approx(E_ref, rel=1e-12, abs=1e-13)
    s_ref, e_ref = orc.gather(N, spins, e, par_ref)
    np.testing.assert_array_equal(from_dev(w2.spins), s_ref)
    np.testing.assert_array_equal(w2.e_loc, e_ref)
    assert np.all(w2.log_weight == 0)


def test_systematic_parents_known_answers_on_gpu():
    from pygfmc_amd import systematic_parents
    assert systematic_parents([0.0, 1.0], 0.0).tolist() == [1, 1]
    assert systematic_parents([0.0, 0.5, 0.5, 0.0], 0.0).tolist() == [1, 1, 2, 2]
    short = [0.25, 0.25, 0.4999999, 0.0]
    par = systematic_parents(short, 0.9999999)
    assert par[-1] == 2 and all(short[k] > 0 for k in par)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sampler_on_gpu_matches_oracle(orc, prec):
    say("sampler", prec)
    from pygfmc_amd import sample_index
    rng = np.random.default_rng(7)
    for n in (1, 3, 64, 65, 300, 2049):
        w = rng.uniform(size=n)
        w[rng.uniform(size=n) < 0.3] = 0
        if w.sum() == 0:
            w[0] = 1
        u = rng.uniform(size=3000)
        got = sample_index(w, u, precision=prec)
        assert np.all(w[got] > 0)
        if prec == "fp64":
            np.testing.assert_array_equal(got, orc.sample_many(w, u))
    got = sample_index([1.0, 0.0, 1.0], np.linspace(0, 1, 1001)[:-1], precision=prec)
    assert set(np.unique(got)) <= {0, 2}


def test_initialisers_match_oracle(orc):
    from pygfmc_amd import prepare_walkers_random, prepare_walkers_sz
    say("init")
    w = prepare_walkers_sz(300, 101, 77, sz=1)
    np.testing.assert_array_equal(from_dev(w.spins), orc.init_sz(101, 300, 51, 77))
    assert np.all(w.config.sum(axis=1) == 1)
    w = prepare_walkers_random(300, 70, 78)
    np.testing.assert_array_equal(from_dev(w.spins), orc.init_random(70, 300, 78))


def test_pack_unpack_roundtrip_on_gpu(orc):
    from pygfmc_amd.walkers import Walker
    rng = np.random.default_rng(0)
    c = rng.choice([-1.0, 1.0], size=(33, 77))
    w = Walker(c)
    np.testing.assert_array_equal(w.config, c)
    np.testing.assert_array_equal(from_dev(w.spins), orc.pack(c))


@pytest.mark.parametrize("name,H,G", [SYSTEMS[0], SYSTEMS[5], SYSTEMS[7]],
                         ids=[IDS[0], IDS[5], IDS[7]])
def test_step_loop_matches_oracle_chain(orc, name, H, G):
    """Several full branching steps (propagate, reconfigure, gather) through
    the C step loop reproduce the oracle's chain: identical spins, series to
    1e-12."""
    say("step loop", name)
    from pygfmc_amd.engine import Chain
    from pygfmc_amd.walkers import Walker
    prob = Problem(H, G, _lib.GFMC_FP64)
    nw = 256
    spins = random_spins(orc, H.N, nw, seed=8, sz=0 if H.n_hx == 0 else None)
    n, beta, alw, seed, step0 = 6, 0.2, 0.1, 4242, 100
    s_ref, e_ref, ser_ref = orc.run_chain(orc.Problem(lower(H, G)), spins, n, beta, alw, seed,
                                          step0)
    z = torch.zeros(nw, dtype=torch.float64, device=dev())
    ch = Chain(prob, Walker.from_device(to_dev(spins), z, z.clone(), H.N))
    ser = ch.steps(n, beta, alw, seed, step0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(from_dev(ch.spins), s_ref)
    np.testing.assert_allclose(ser.cpu().numpy(), ser_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ch.e_loc.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12)
