"""Pin the CPU oracle: the reference's golden vectors, its known-answer
tests, brute force on tiny systems, and closed forms.  CPU only."""
import numpy as np
import pytest

from helpers import (basis, golden_system, index_of, jastrow_log_psi, random_model,
                     spinhalf_matrix)
from pygfmc_amd.problem import dependency_lists, lower

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                    "golden_spinhalf.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


# -- golden local energies (test_golden_outputs.py, tolerance rtol 1e-10) --------

@pytest.mark.parametrize("guide", ["none", "jastrow"])
@pytest.mark.parametrize("method", ["cont_time", "multi_step", "single_step"])
def test_oracle_local_energy_matches_golden(orc, golden, guide, method):
    H, G = golden_system(guide=guide == "jastrow")
    prob = orc.Problem(lower(H, G))
    c = golden[f"spinhalf_{guide}_{method}/config"]
    e_ref = golden[f"spinhalf_{guide}_{method}/e_loc"]
    rates, e_loc, _ = prob.rates(orc.pack(c))
    np.testing.assert_allclose(e_loc, e_ref, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(rates.sum(axis=1), e_ref, rtol=1e-12, atol=1e-13)


# -- brute force: Kronecker-product matrix and explicit log psi ------------------

@pytest.mark.parametrize("seed,lattice,pairs", [(0, "chain", "edges"), (1, "chain", "nnn"),
                                                (2, "kagome", "edges")])
def test_oracle_rates_equal_matrix_elements_times_guide_ratios(orc, seed, lattice, pairs):
    N = 8 if lattice == "chain" else 12
    H, G = random_model(N=N, seed=seed, lattice=lattice, guide_pairs=pairs)
    M = spinhalf_matrix(H).toarray()
    rng = np.random.default_rng(seed)
    configs = basis(H.N)[rng.choice(2 ** H.N, size=64, replace=False)]
    prob = orc.Problem(lower(H, G))
    rates, e_loc, log_diff = prob.rates(orc.pack(configs))
    lp = jastrow_log_psi(G, configs)
    for w, x in enumerate(configs):
        col = index_of(x)[0]
        succ = H.successors(x)
        rows = index_of(succ)
        ld_ref = jastrow_log_psi(G, succ) - lp[w]
        # diagonal and off-diagonal matrix elements (the matrix is symmetric)
        assert rates[w, 0] == pytest.approx(M[col, col], abs=1e-12)
        bare = np.array([M[r, col] for r in rows[1:]])
        np.testing.assert_allclose(log_diff[w, 1:], ld_ref[1:], atol=1e-12)
        np.testing.assert_allclose(rates[w, 1:], bare * np.exp(ld_ref[1:]), rtol=1e-12,
                                   atol=1e-14)
        # local energy = (H psi)(x) / psi(x) over the full column
        psi_all = np.exp(jastrow_log_psi(G, basis(H.N)) - lp[w])
        assert e_loc[w] == pytest.approx(M[:, col] @ psi_all, rel=1e-11, abs=1e-11)


def test_oracle_log_psi_is_the_pair_sum(orc):
    H, G = random_model(N=8, seed=4, guide_pairs="nnn")
    configs = basis(8)[::7]
    prob = orc.Problem(lower(H, G))
    np.testing.assert_allclose(prob.log_psi(orc.pack(configs)), jastrow_log_psi(G, configs),
                               atol=1e-12)


def test_dependency_lists_cover_every_changed_slot(orc):
    """After move s only the slots listed for s change value, and only the
    listed edges change their parallel state."""
    for seed, pairs in ((0, "edges"), (3, "nnn")):
        H, G = random_model(N=8, seed=seed, guide_pairs=pairs)
        dep_off, dep_slots, edge_off, dep_edges = dependency_lists(H, G, G.build_cache(H))
        prob = orc.Problem(lower(H, G))
        rng = np.random.default_rng(seed)
        for x in basis(8)[rng.choice(256, 12, replace=False)]:
            r0, _, _ = prob.rates(orc.pack(x[None]))
            anti0 = x[H.edge_i] != x[H.edge_j]
            for s in range(H.n_off):
                y = H.generate_state(x, s + 1)
                r1, _, _ = prob.rates(orc.pack(y[None]))
                changed = np.flatnonzero(r1[0, 1:] != r0[0, 1:])
                listed = set(dep_slots[dep_off[s]:dep_off[s + 1]].tolist())
                assert set(changed.tolist()) <= listed
                anti1 = y[H.edge_i] != y[H.edge_j]
                listed_e = set(dep_edges[edge_off[s]:edge_off[s + 1]].tolist())
                assert set(np.flatnonzero(anti0 != anti1).tolist()) <= listed_e


# -- the windowed estimator (estimators.py:268-305) -------------------------------

def test_energy_estimator_reproduces_golden(golden):
    from pygfmc_amd.estimators import estimate_energy
    cases = sorted({k.split("/")[0] for k in golden.files if "/raw/" in k})
    assert len(cases) == 45
    exact = 0
    for c in cases:
        E = estimate_energy([2, 4], golden[f"{c}/raw/log_weights"], golden[f"{c}/raw/energy"])
        for p in (2, 4):
            ref = float(golden[f"{c}/energy/{p}"])
            assert E[p] == pytest.approx(ref, rel=1e-14, abs=1e-14)
            exact += E[p] == ref
    assert exact >= 80        # the rest differ in the last bit (NumPy's sum order)


def test_jackknife_two_blocks_is_half_chain_split(golden):
    """n_blocks = 2: err = |x(first half) - x(second half)| / 2
    (test_io_hdf5.py:724 restated for the energy)."""
    from pygfmc_amd.estimators import evaluate_observables
    raw = dict(log_weights=golden["spinhalf_jastrow_cont_time/raw/log_weights"],
               energy=golden["spinhalf_jastrow_cont_time/raw/energy"], p_array=[2, 4])
    n = len(raw["energy"])
    half = lambda a, b: dict(log_weights=raw["log_weights"][a:b], energy=raw["energy"][a:b],  # noqa
                             p_array=[2, 4])
    x0, _ = evaluate_observables(half(n // 2, n))
    x1, _ = evaluate_observables(half(0, n // 2))
    with pytest.warns(UserWarning):
        _, _, err, _ = evaluate_observables(raw, n_blocks=2)
    for p in (2, 4):
        assert err[p] == pytest.approx(abs(x0[p] - x1[p]) / 2, rel=1e-12)


def test_jackknife_error_of_iid_series_is_the_standard_error():
    """For p = 1 and constant weights the estimator is the mean, and the
    jackknife over single-sample blocks is exactly the standard error."""
    from pygfmc_amd.estimators import evaluate_observables
    rng = np.random.default_rng(0)
    E = rng.normal(size=400)
    raw = dict(log_weights=np.zeros(400), energy=E, p_array=[1])
    est, _, err, _ = evaluate_observables(raw, n_blocks=400)
    assert est[1] == pytest.approx(E[1:].mean(), rel=1e-12)
    # the estimator uses E[1:] (window p = 1); the replicas keep 398 or 399 of them
    assert err[1] == pytest.approx(E[1:].std(ddof=1) / np.sqrt(399), rel=0.02)


# -- the wave-tree summation order ------------------------------------------------

def test_wave_tree_is_an_inclusive_scan(orc):
    """Exact on integers (no rounding); within a few ulps of the exact prefix
    sum otherwise; and really a different association than sequential."""
    rng = np.random.default_rng(0)
    ints = rng.integers(-1000, 1000, 64).astype(float)
    np.testing.assert_array_equal(orc.wave_tree(ints), np.cumsum(ints))
    v = rng.uniform(0, 1, 64) * 10.0 ** rng.integers(-8, 8, 64)
    from fractions import Fraction
    exact = np.array([float(sum(Fraction(x) for x in v[:k + 1])) for k in range(64)])
    np.testing.assert_allclose(orc.wave_tree(v), exact, rtol=4e-16 * 8)
    assert np.any(orc.wave_tree(v) != np.cumsum(v))


# -- Philox4x32-10 known-answer vectors (Random123 kat_vectors) -------------------

PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(orc, ctr, key, want):
    assert tuple(int(v) for v in orc.philox(ctr, key)) == want


def test_product_philox_matches_known_answers():
    """The HIP library's Philox (its host entry; no GPU needed)."""
    import ctypes
    from pygfmc_amd import _lib
    L = _lib.load()
    ctr = np.array([c for c, _, _ in PHILOX_KAT], dtype=np.uint32)
    out = np.zeros((3, 4), dtype=np.uint32)
    for i, (c, k, w) in enumerate(PHILOX_KAT):
        cc = np.array(c, dtype=np.uint32)
        kk = np.array(k, dtype=np.uint32)
        oo = np.zeros(4, dtype=np.uint32)
        P = ctypes.POINTER(ctypes.c_uint32)
        assert L.gfmc_philox_host(1, cc.ctypes.data_as(P), kk.ctypes.data_as(P),
                                  oo.ctypes.data_as(P)) == 0
        out[i] = oo
    assert [tuple(int(v) for v in r) for r in out] == [w for _, _, w in PHILOX_KAT]
    del ctr


# -- sampler KATs (test_sampler.py:20-44) -------------------------------------------

def test_sampler_never_draws_a_zero_weight_slot(orc):
    u = np.linspace(0, 1, 4097)[:-1]
    idx = orc.sample_many([1.0, 0.0, 1.0], u)
    assert set(np.unique(idx)) <= {0, 2}
    idx = orc.sample_many([0.0, 2.0, 0.0, 0.0], u)
    assert set(np.unique(idx)) == {1}
    # leading, interior and trailing zeros across several 64-lane blocks
    w = np.zeros(300)
    w[[5, 130, 131, 299]] = [1.0, 0.5, 0.25, 2.0]
    idx = orc.sample_many(w, u)
    assert set(np.unique(idx)) <= {5, 130, 131, 299}


def test_sampler_draws_proportionally(orc):
    rng = np.random.default_rng(1)
    w = rng.uniform(0, 1, 200)
    w[rng.choice(200, 50, replace=False)] = 0
    u = rng.uniform(size=200000)
    idx = orc.sample_many(w, u)
    freq = np.bincount(idx, minlength=200) / len(u)
    assert np.all(freq[w == 0] == 0)
    np.testing.assert_allclose(freq, w / w.sum(), atol=4e-3)


def test_sampler_matches_exact_searchsorted_on_exact_sums(orc):
    """With dyadic weights every partial sum is exact, so the blocked CDF is
    the exact CDF and the draw is searchsorted(cdf, u, 'right')."""
    rng = np.random.default_rng(2)
    w = rng.integers(0, 8, 257) / 8.0
    u = rng.uniform(size=5000)
    idx = orc.sample_many(w, u)
    cdf = np.cumsum(w)
    ref = np.searchsorted(cdf, u * cdf[-1], side="right")
    np.testing.assert_array_equal(idx, ref)


def test_in_block_tree_order_differs_from_a_sequential_cdf(orc):
    """Inside the drawn lane's block the CDF is the group's tree scan of the
    block, not a sequential running sum: at a point u between the two
    roundings of one slot's CDF value they draw different slots (so the
    kernels' bit-exact move tests tell the orders apart)."""
    import ctypes
    rng = np.random.default_rng(5)
    K = 40
    n = 64 * K
    w = rng.uniform(0.5, 1.5, n) * 10.0 ** rng.integers(-6, 3, n)
    tot, incl = orc.blocked_scan(w)
    L = orc.lib()
    L.orc_sample_blocked.restype = ctypes.c_int
    L.orc_sample_blocked.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.c_double]
    wp = w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = incl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    found = 0
    for j in range(1, 64):
        blk = w[j * K:(j + 1) * K]
        c, seq_exact = incl[j - 1], []
        for v in blk:
            c = c + v
            seq_exact.append(c)
        tree = incl[j - 1] + orc.wave_tree(np.concatenate([blk, np.zeros(64 - K)]))[:K]
        for k in range(K - 1):
            if tree[k] > seq_exact[k]:
                x = seq_exact[k]                     # tree: slot k, sequential: later
                assert L.orc_sample_blocked(wp, n, ip, x) == j * K + k
                found += 1
            elif tree[k] < seq_exact[k]:
                x = tree[k]                          # tree: later, sequential: slot k
                assert L.orc_sample_blocked(wp, n, ip, x) > j * K + k
                found += 1
    assert found > 20


def test_neg_log_unit_is_minus_log_within_one_ulp(orc):
    """The hold's -log(1 - psi) (fdlibm's reduction and polynomial) against
    libm's log on 53-bit uniforms, the ends q = 1 and q = 2^-53 included."""
    rng = np.random.default_rng(3)
    psi = (rng.integers(0, 2 ** 53, 200000, dtype=np.int64)).astype(np.float64) / 2.0 ** 53
    q = 1.0 - psi
    got = orc.neg_log_unit(q)
    want = -np.log(q)
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / np.spacing(want[nz])) <= 1.0
    assert orc.neg_log_unit([1.0])[0] == 0.0
    assert orc.neg_log_unit([2.0 ** -53])[0] == -np.log(2.0 ** -53)


# -- systematic resampling KATs (test_walker_io.py:60-77) --------------------------

def _parents_from_w(orc, w, psi):
    with np.errstate(divide="ignore"):
        lw = np.log(np.asarray(w, dtype=float))
    par, _, _, _ = orc.reconfigure(lw, np.zeros(len(w)), psi)
    return par.tolist()


def test_systematic_parents_known_answers(orc):
    assert _parents_from_w(orc, [0.0, 1.0], 0.0) == [1, 1]
    assert _parents_from_w(orc, [0.0, 0.5, 0.5, 0.0], 0.0) == [1, 1, 2, 2]
    short = [0.25, 0.25, 0.4999999, 0.0]
    par = _parents_from_w(orc, short, 0.9999999)
    assert par[-1] == 2 and all(short[k] > 0 for k in par)
    # dead walkers: -inf log weights are never chosen
    for psi in np.linspace(0, 0.999, 20):
        par, _, _, _ = orc.reconfigure(np.array([-np.inf, 0.0, -np.inf]), np.zeros(3), psi)
        assert par.tolist() == [1, 1, 1]


def test_reconfiguration_matches_the_direct_formulas(orc):
    rng = np.random.default_rng(3)
    n = 5000
    lw = rng.normal(scale=2.0, size=n)
    e = rng.normal(size=n)
    psi = 0.37
    par, lg, En, _ = orc.reconfigure(lw, e, psi)
    m = lw.max()
    logG = m + np.log(np.sum(np.exp(lw - m)))
    w = np.exp(lw - logG)
    assert lg == pytest.approx(logG - np.log(n), rel=1e-13)
    assert En == pytest.approx(np.sum(w * e), rel=1e-12, abs=1e-14)
    ref = np.minimum(np.searchsorted(np.cumsum(w), (np.arange(n) + psi) / n, side="right"),
                     n - 1)
    # rounding of the two prefix sums may move a parent across a boundary only
    # where u lies within a few ulps of a cdf value
    assert np.mean(par == ref) > 0.999
    # offspring counts are floor or ceil of n * w
    cnt = np.bincount(par, minlength=n)
    assert np.all(np.abs(cnt - n * w) < 1.0 + 1e-9)


def test_init_sz_has_exact_magnetisation(orc):
    spins = orc.init_sz(37, 500, 20, seed=11)
    c = orc.unpack(spins, 37)
    assert np.all(c.sum(axis=1) == 2 * 20 - 37)
    # every site is up with probability 20/37
    assert np.abs(((c + 1) / 2).mean(axis=0) - 20 / 37).max() < 0.08
    # walkers differ
    assert len({row.tobytes() for row in spins}) > 490


def test_init_random_is_fair(orc):
    spins = orc.init_random(45, 2000, seed=5)
    c = orc.unpack(spins, 45)
    assert abs(c.mean()) < 0.02
    assert np.all((spins[:, 1] >> 13) == 0)      # bits past N stay clear


def test_pack_unpack_roundtrip(orc):
    rng = np.random.default_rng(0)
    c = rng.choice([-1.0, 1.0], size=(17, 70))
    np.testing.assert_array_equal(orc.unpack(orc.pack(c), 70), c)


# -- RBM guide ---------------------------------------------------------------------

def test_oracle_rbm_ratios_are_log_cosh_differences(orc):
    """The oracle's RBM log-ratios and log psi against the explicit
    a.x + sum_h log cosh(b + W x) on every successor."""
    from helpers import rbm_log_psi, random_rbm
    from pygfmc_amd import SpinHalf, chain
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -0.6, 0.7, h_x=0.4, dtype=np.float64)
    G = random_rbm(8, alpha=2, seed=1)
    prob = orc.Problem(lower(H, G))
    X = basis(8)[::17]
    rates, e_loc, ld = prob.rates(orc.pack(X))
    for w, x in enumerate(X):
        succ = H.successors(x)
        ref = rbm_log_psi(G, succ) - rbm_log_psi(G, x)
        np.testing.assert_allclose(ld[w, 1:], ref[1:], atol=1e-13)
    np.testing.assert_allclose(prob.log_psi(orc.pack(X)), rbm_log_psi(G, X), atol=1e-12)
    M = spinhalf_matrix(H).toarray()
    lp_all = rbm_log_psi(G, basis(8))
    for w, x in enumerate(X):
        col = index_of(x)[0]
        assert e_loc[w] == pytest.approx(M[:, col] @ np.exp(lp_all - lp_all[col]), rel=1e-11)


# -- RBM training: variance-loss gradient -----------------------------------------

def _torch_variance_grad(H, G, X, clip):
    """Gradient of the local-energy variance by torch FP64 autograd, with E_L
    from the Kronecker matrix (E_L(x) = sum_y M_yx psi(y) / psi(x)) and the
    explicit log psi = a.x + sum_h log cosh(b + W x); clip > 0 differentiates
    the clamped surrogate (2/n) sum_i clamp(d_i) E_i."""
    import torch
    M = torch.tensor(spinhalf_matrix(H).toarray())
    B = torch.tensor(basis(H.N))
    W = torch.tensor(G.W.astype(np.float64), requires_grad=True)
    a = torch.tensor(G.a.astype(np.float64), requires_grad=True)
    b = torch.tensor(G.b.astype(np.float64), requires_grad=True)
    lp = B @ a + torch.log(torch.cosh(B @ W.T + b)).sum(1)
    idx = torch.tensor(index_of(X))
    E = (M[:, idx] * torch.exp(lp[:, None] - lp[idx][None, :])).sum(0)
    n = len(idx)
    dev = E - E.mean()
    var = (dev ** 2).mean()
    if clip > 0:
        lim = clip * torch.sqrt(var.detach())
        loss = (2.0 / n) * (torch.clamp(dev.detach(), -lim, lim) * E).sum()
    else:
        loss = var
    loss.backward()
    g = np.concatenate([W.grad.numpy().ravel(), a.grad.numpy(), b.grad.numpy()])
    return g, float(E.detach().mean()), float(var.detach())


@pytest.mark.parametrize("clip", [0.0, 1.0])
@pytest.mark.parametrize("hx", [0.0, 0.4])
def test_oracle_rbm_variance_gradient_matches_autograd(orc, clip, hx):
    from helpers import random_rbm
    from pygfmc_amd import SpinHalf, chain
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -0.6, 0.7, h_x=hx, dtype=np.float64)
    G = random_rbm(8, alpha=2, seed=3, scale=0.3)
    rng = np.random.default_rng(5)
    X = basis(8)[rng.choice(256, 40, replace=False)]
    if hx == 0.0:      # stay in a fixed-magnetisation sector as GFMC does
        X = X[X.sum(1) == 0] if np.any(X.sum(1) == 0) else X
    g, mean, var = orc.Problem(lower(H, G)).rbm_vmin_grad(orc.pack(X), clip)
    g_ref, mean_ref, var_ref = _torch_variance_grad(H, G, X, clip)
    assert mean == pytest.approx(mean_ref, rel=1e-12)
    assert var == pytest.approx(var_ref, rel=1e-10)
    np.testing.assert_allclose(g, g_ref, rtol=1e-9, atol=1e-12 * np.abs(g_ref).max())


def test_oracle_rbm_variance_gradient_matches_finite_differences(orc):
    """Central differences of the oracle's own variance on a few parameters
    (a and b entries, exact in float32 steps)."""
    from helpers import random_rbm
    from pygfmc_amd import SpinHalf, chain
    from pygfmc_amd.guides import RBMSpinHalf
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = random_rbm(8, alpha=2, seed=4, scale=0.3)
    X = basis(8)[[3, 5, 6, 9, 10, 12, 15, 60, 90, 99, 165, 195]]
    X = X[X.sum(1) == 0]
    spins = orc.pack(X)
    g, _, _ = orc.Problem(lower(H, G)).rbm_vmin_grad(spins)
    P = G.parameters()
    nW = G.W.size
    h = 2.0 ** -10
    for which, k, off in (("a", 2, nW + 2), ("b", 5, nW + 8 + 5), ("b", 11, nW + 8 + 11)):
        vals = []
        for sgn in (1, -1):
            q = {key: v.copy() for key, v in P.items()}
            q[which][k] += sgn * h
            Gq = RBMSpinHalf(8, hidden=G.hidden, W=q["W"], a=q["a"], b=q["b"])
            vals.append(orc.Problem(lower(H, Gq)).rbm_vmin_grad(spins)[2])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[off] == pytest.approx(fd, rel=2e-5, abs=1e-9)


def test_oracle_rbm_theta_batch(orc):
    from helpers import random_rbm
    from pygfmc_amd import SpinHalf, chain
    G = random_rbm(10, alpha=3, seed=6)
    H = SpinHalf.create(chain(10).edges, -1.0, -1.0, 1.0, dtype=np.float64)
    X = basis(10)[::37]
    th = orc.Problem(lower(H, G)).rbm_theta(orc.pack(X))
    ref = X @ G.W.astype(np.float64).T + G.b.astype(np.float64)
    np.testing.assert_allclose(th, ref, rtol=0, atol=1e-12)


# -- transverse-field Ising chains: free fermions against ED -------------------------

@pytest.mark.parametrize("L,h", [(6, 0.5), (8, 1.0), (10, 1.37)])
def test_tfim_free_fermion_energy_matches_ed(L, h):
    from helpers import ground_energy
    from pygfmc_amd import chain, tfim_chain_ground_energy, transverse_field_ising
    H = transverse_field_ising(chain(L).edges, 1.0, h)
    assert tfim_chain_ground_energy(L, 1.0, h) == pytest.approx(ground_energy(H), rel=1e-12)


# -- blocked order over 32 / 16 lanes per walker --------------------------------------

@pytest.mark.parametrize("G", [16, 32, 64])
def test_group_scan_is_exact_prefix_sum_on_integers(orc, G):
    """With integer-valued weights every order is exact: the G-lane blocked
    scan returns the prefix sums of the lane blocks and the sampler the
    first slot whose cumulative sum exceeds u."""
    rng = np.random.default_rng(G)
    for n in (1, 7, G, 3 * G + 5, 200):
        v = rng.integers(0, 5, n).astype(float)
        v[-1] += 1.0
        K = (n + G - 1) // G
        with orc.lanes(G):
            tot, incl = orc.blocked_scan(v)
            blocks = [v[l * K:(l + 1) * K].sum() for l in range(G)]
            np.testing.assert_array_equal(incl[:G], np.cumsum(blocks))
            assert tot == v.sum()
            us = rng.uniform(size=50)
            got = orc.sample_many(v, us)
        ref = np.searchsorted(np.cumsum(v), us * v.sum(), side="right")
        np.testing.assert_array_equal(got, ref)


def test_group_tree_differs_from_wave_tree_on_mixed_magnitudes(orc):
    """The 16- and 32-lane trees are different summation orders from the
    64-lane tree (so a kernel using the wrong one fails bit-exact tests)."""
    v = np.random.default_rng(2).uniform(size=200)
    res = {}
    for G in (16, 32, 64):
        with orc.lanes(G):
            res[G] = orc.blocked_scan(v)[0]
    assert len({res[16], res[32], res[64]}) == 3


# -- discrete propagators (single_step / multi_step) ----------------------------------

def _discrete_system():
    from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain
    L = chain(6)
    H = SpinHalf.create(L.edges, -1.0, -0.7, 0.8, h_x=0.3, dtype=np.float64)
    G = JastrowSpinHalf(6, L.edges, beta=0.2, h=0.05, dtype=np.float64)
    return H, G


def test_oracle_single_step_draws_the_column_of_lambda_minus_h(orc):
    """One sub-step from a fixed configuration: the successor frequencies
    match (lam delta_yx - H_yx psi(y)/psi(x)) / b_x from the Kronecker
    matrix, and log_w = log(b_x) exactly, b_x = lam - E_L(x)."""
    H, G = _discrete_system()
    M = spinhalf_matrix(H).toarray()
    B = basis(6)
    lp = jastrow_log_psi(G, B)
    lam = 4.0
    x = np.array([[1, -1, -1, 1, 1, -1]], dtype=float)
    col = index_of(x)[0]
    P = (lam * (np.arange(64) == col) - M[:, col]) * np.exp(lp - lp[col])
    bx = P.sum()
    n = 40000
    spins = np.repeat(orc.pack(x), n, axis=0)
    s, lw, e, ev = orc.Problem(lower(H, G)).propagate(spins, np.zeros(n), np.zeros(n), 1, 0.0,
                                                      77, 3, 0, method=1, lam=lam)
    np.testing.assert_allclose(lw, np.log(bx), rtol=1e-13)
    assert np.all(ev == 2)
    idx = index_of(orc.unpack(s, 6))
    freq = np.bincount(idx, minlength=64) / n
    p = P / bx
    assert np.all(np.abs(freq - p) <= 5 * np.sqrt(p * (1 - p) / n) + 1e-12)
    # e_loc is the local energy of the successor
    e_ref = (M[:, idx] * np.exp(lp[:, None] - lp[idx][None, :])).sum(0)
    np.testing.assert_allclose(e, e_ref, rtol=1e-12)


def test_oracle_multi_step_matches_single_step_in_distribution(orc):
    """Batched geometric holds are the same Markov chain: after n_bra = 4
    sub-steps from a fixed configuration, the successor distribution and the
    weight sum per successor agree between the two discrete methods."""
    H, G = _discrete_system()
    prob = orc.Problem(lower(H, G))
    x = np.array([[1, 1, -1, -1, 1, -1]], dtype=float)
    n, lam = 60000, 3.5
    spins = np.repeat(orc.pack(x), n, axis=0)
    out = {}
    for method in (1, 2):
        s, lw, _, ev = prob.propagate(spins, np.zeros(n), np.zeros(n), 4, 0.0, 5, 1, 0,
                                      method=method, lam=lam)
        idx = index_of(orc.unpack(s, 6))
        out[method] = (np.bincount(idx, minlength=64) / n,
                       np.bincount(idx, weights=np.exp(lw), minlength=64) / n, ev)
    f1, w1, ev1 = out[1]
    f2, w2, ev2 = out[2]
    assert np.all(np.abs(f1 - f2) <= 6 * np.sqrt(f1 * (1 - f1) / n) + 1e-12)
    np.testing.assert_allclose(w1.sum(), w2.sum(), rtol=0.01)
    assert ev2.mean() < ev1.mean()      # the holds skip the stays


@pytest.mark.parametrize("method", [1, 2])
def test_oracle_discrete_gfmc_energy_matches_ed(orc, method):
    """chain(8) Heisenberg without a guide: the oracle's branching chain with
    the discrete propagators reaches the exact ground energy."""
    import warnings
    from helpers import ground_energy
    from pygfmc_amd import SpinHalf, chain
    from pygfmc_amd.estimators import evaluate_observables
    H = SpinHalf.create(chain(8).edges, -1.0, -1.0, 1.0, dtype=np.float64)
    e0 = ground_energy(H)
    prob = orc.Problem(lower(H, None))
    spins = orc.init_sz(8, 256, 4, 3)
    lam = 2.0
    s, _, _ = orc.run_chain(prob, spins, 60, 6, np.log(lam - e0), 9, 0, warmup=True,
                            method=method, lam=lam)
    _, _, ser = orc.run_chain(prob, s, 600, 6, np.log(lam - e0), 9, 100, method=method,
                              lam=lam)
    raw = dict(log_weights=ser[:, 0], energy=ser[:, 1], p_array=[4, 16], m_array=[2, 8])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    assert abs(E[16] - e0) < 4 * err[16] + 1e-3


def test_oracle_propagation_of_a_sample_equals_the_full_population(orc):
    """propagate_ids on selected rows gives exactly those rows of the full
    population's propagation (the walker index is the Philox counter)."""
    from helpers import random_model
    H, G = random_model(N=8, seed=3)
    prob = orc.Problem(lower(H, G))
    nw = 200
    spins = orc.init_random(8, nw, 9)
    lw0 = np.linspace(-0.2, 0.2, nw)
    full = prob.propagate(spins, lw0, np.zeros(nw), 0.3, 0.1, 77, 5, 0)
    ids = np.array([0, 3, 17, 150, 199])
    part = prob.propagate_ids(ids, spins[ids], lw0[ids], np.zeros(len(ids)), 0.3, 0.1, 77, 5, 0)
    for a, b in zip(part, full):
        np.testing.assert_array_equal(a, b[ids])
    # a different index is a different stream
    other = prob.propagate_ids(ids + 1, spins[ids], lw0[ids], np.zeros(len(ids)), 0.3, 0.1, 77,
                               5, 0)
    assert not np.array_equal(other[1], part[1])


@pytest.mark.timeout(60)
def test_overflowing_total_rate_ends_the_step_without_a_hop(orc):
    """A guide whose ratios overflow FP64 (Jastrow beta = 200 on a ring:
    pair moves with both outer bonds aligned have ratio e^800) makes some
    walkers' total rate inf: such a walker takes no hop (a zero hold time
    would never end the step) and its log weight becomes +inf; the others
    step (some end finite, some reach such a configuration by a hop)."""
    from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain
    L = chain(8)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(8, L.edges, beta=200.0, dtype=np.float64)
    spins = orc.init_sz(8, 64, 4, 7)
    p = orc.Problem(lower(H, G))
    out_s, lw, e, ev = p.propagate(spins, np.zeros(64), np.zeros(64), 0.2, 0.0, 5, 1, 0)
    x = orc.unpack(spins, 8)
    # a ring site pair (i, i+1) flipped with both outer bonds aligned changes
    # log psi by 4 beta = 800: the total rate overflows
    big = np.zeros(64, dtype=bool)
    for w in range(64):
        for i in range(8):
            a, b, c, d = x[w, (i - 1) % 8], x[w, i], x[w, (i + 1) % 8], x[w, (i + 2) % 8]
            if b != c and a * b > 0 and c * d > 0:
                big[w] = True
    assert big.any() and (~big).any()
    np.testing.assert_array_equal(lw[big], np.inf)
    np.testing.assert_array_equal(out_s[big], spins[big])      # no hop
    np.testing.assert_array_equal(ev[big], 1)                  # one evaluation
    assert np.isfinite(lw[~big]).any()
