"""Pin the diagonal forward-walking measurement on the CPU: the operator
columns against the Hamiltonian matrix and the structure factor's pair sum,
the oracle's ancestry ring against the explicit lineage, and the host
estimator against the reference's golden estimates."""
import numpy as np
import pytest

from helpers import basis, golden_system, index_of, spinhalf_matrix
from pygfmc_amd import SpinHalf, SzSzStructureFactor, chain, square
from pygfmc_amd.estimators import estimate_diagonal
from pygfmc_amd.observables import is_valid_k, lower_diag_operators
from pygfmc_amd.problem import lower

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                    "golden_spinhalf.npz")
CASES = [f"spinhalf_{g}_{m}" for g in ("none", "jastrow", "cps")
         for m in ("cont_time", "multi_step", "single_step")]


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


# -- the host estimator against the reference's golden estimates ----------------

@pytest.mark.parametrize("case", CASES)
def test_diagonal_estimates_reproduce_the_golden_values(golden, case):
    """The golden runs measured off-diagonal operators too, so their
    diagonal estimates use the common window trim i >= max(m) = 2."""
    lw = golden[f"{case}/raw/log_weights"]
    diag = {op: golden[f"{case}/raw/diag/{op}"] for op in ("zz", "sf")}
    obs = estimate_diagonal([2, 4], lw, diag, m_max_common=2)
    for op in ("zz", "sf"):
        for p in (2, 4):
            np.testing.assert_allclose(obs[op][p], golden[f"{case}/obs/{op}/{p}"], rtol=1e-12,
                                       atol=1e-14)
    untrimmed = estimate_diagonal([2, 4], lw, diag, m_max_common=0)
    assert abs(untrimmed["zz"][4] - obs["zz"][4]) > 1e-6


@pytest.mark.parametrize("case", CASES)
def test_golden_rows_before_the_first_step_of_the_ring_are_zero(golden, case):
    """A row whose lag m reaches before the ring's first step reads the
    initial (zero) values: rows t < m are 0, later rows are not."""
    zz = golden[f"{case}/raw/diag/zz"]
    sf = golden[f"{case}/raw/diag/sf"]
    for j, m in enumerate((1, 2)):
        assert np.all(zz[:m, j] == 0.0) and np.all(sf[:m, :, j] == 0.0)
        assert np.count_nonzero(zz[m:, j]) > 0.9 * (len(zz) - m)


# -- operator columns ----------------------------------------------------------------

def _random_spinhalf_diag(N, seed):
    rng = np.random.default_rng(seed)
    e = chain(N).edges
    return SpinHalf.create(e, J_z=rng.uniform(-1.5, 1.5, len(e)),
                           h_z=rng.uniform(-0.7, 0.7, N), dtype=np.float64)


@pytest.mark.parametrize("make", [
    lambda: SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64),
    lambda: SpinHalf.create(chain(8).edges, J_z=0.7, h_z=0.3, dtype=np.float64),
    lambda: _random_spinhalf_diag(8, 3),
    lambda: SpinHalf.create(square(3, 3).edges, J_z=-1.25, dtype=np.float64),
], ids=["zz_uniform", "zz_uniform_hz", "zz_per_edge_fields", "zz_square"])
def test_pair_column_is_the_hamiltonian_diagonal(orc, make):
    """Each configuration's PAIR column equals the diagonal element of the
    Kronecker-product matrix of the same SpinHalf (and the rate oracle's
    H_xx)."""
    H = make()
    N = H.N
    cfg = basis(N)
    d = orc.Diag(lower_diag_operators({"zz": H}, N))
    vals = d.values(orc.pack(cfg))[:, 0]
    M = spinhalf_matrix(H)
    np.testing.assert_allclose(vals, M.diagonal()[index_of(cfg)], rtol=1e-13, atol=1e-13)
    rates, _, _ = orc.Problem(lower(H, None)).rates(orc.pack(cfg))
    np.testing.assert_allclose(vals, rates[:, 0], rtol=1e-13, atol=1e-13)


def test_structure_factor_column_is_the_pair_sum(orc):
    """(|rho_k|^2 - N)/4 against sum_{i<j} 2 cos(k.(r_i - r_j)) S^z_i S^z_j
    on a 4x4 torus, for every wavevector of its reciprocal lattice."""
    L = 4
    pos = np.array([(x, y) for x in range(L) for y in range(L)], dtype=float)
    ks = 2 * np.pi * np.array([(a, b) for a in range(L) for b in range(L)], dtype=float) / L
    assert np.all(is_valid_k(ks, L * np.eye(2)))
    sf = SzSzStructureFactor(pos, ks, dtype=np.float64)
    rng = np.random.default_rng(5)
    cfg = rng.choice([-1.0, 1.0], size=(40, L * L))
    d = orc.Diag(lower_diag_operators({"sf": sf}, L * L))
    vals = d.values(orc.pack(cfg))
    dr = pos[:, None, :] - pos[None, :, :]
    iu = np.triu_indices(L * L, 1)
    for q, k in enumerate(ks):
        ph = 2 * np.cos(dr @ k)[iu]
        pair = np.array([np.sum(ph * (c[iu[0]] / 2) * (c[iu[1]] / 2)) for c in cfg])
        np.testing.assert_allclose(vals[:, q], pair, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(vals, sf.diagonal_weight(cfg), rtol=1e-13, atol=1e-13)
    assert not is_valid_k([1.0, 0.0], L * np.eye(2))


def test_structure_factor_tables_round_to_the_requested_dtype():
    pos = np.arange(8, dtype=float)[:, None]
    ks = (2 * np.pi * np.arange(3) / 8)[:, None]
    f32 = SzSzStructureFactor(pos, ks)
    f64 = SzSzStructureFactor(pos, ks, dtype=np.float64)
    np.testing.assert_array_equal(f32.cos, np.cos(ks @ pos.T).astype(np.float32))
    assert np.max(np.abs(f32.cos - f64.cos)) > 0


def test_lowering_layout_and_split():
    H = SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64)
    sf = SzSzStructureFactor(np.arange(8.0)[:, None], (2 * np.pi * np.arange(3) / 8)[:, None])
    low = lower_diag_operators({"zz": H, "sf": sf, "zz2": H}, 8)
    assert low.n_cols == 5 and list(low.kind) == [0, 1, 1, 1, 0]
    rows = np.arange(2 * 5 * 2, dtype=float).reshape(2, 5, 2)
    parts = low.split(rows)
    assert parts["zz"].shape == (2, 2) and parts["sf"].shape == (2, 3, 2)
    np.testing.assert_array_equal(parts["sf"], rows[:, 1:4, :])
    np.testing.assert_array_equal(parts["zz2"], rows[:, 4, :])
    with pytest.raises(TypeError):
        lower_diag_operators({"x": object()}, 8)
    with pytest.raises(ValueError):
        lower_diag_operators({"zz": H}, 9)


# -- the oracle's ancestry ring against the explicit lineage ----------------------

def test_ring_walk_equals_the_explicit_lineage(orc):
    """The ring of max(m) + 1 slots, read modulo its size, gives the same rows
    as following the full parent history step by step, with values before
    the first step taken as zero (driver.py:240-288)."""
    H, G = golden_system()
    N, nw = 8, 48
    sf = SzSzStructureFactor(np.arange(8.0)[:, None], (2 * np.pi * np.arange(3) / 8)[:, None],
                             dtype=np.float64)
    low = lower_diag_operators({"zz": SpinHalf.create(chain(8).edges, J_z=1.0,
                                                      dtype=np.float64), "sf": sf}, N)
    d = orc.Diag(low)
    m_list = [0, 1, 1, 3, 4]
    spins = orc.init_random(N, nw, 11)
    n = 11
    _, _, series, rows, pars, vals = orc.diag_chain(orc.Problem(lower(H, G)), d, m_list, spins,
                                                    n, 0.2, 0.3, 77, 5, history=True)
    for t in range(n):
        for j, m in enumerate(m_list):
            anc = np.arange(nw)
            for s in range(t, t - m, -1):
                if s >= 0:
                    anc = pars[s][anc]
            want = vals[t - m][anc].mean(axis=0) if t - m >= 0 else np.zeros(d.V)
            np.testing.assert_allclose(rows[t, :, j], want, rtol=1e-14, atol=1e-15)
    assert np.all(rows[0, :, 1] == 0) and np.any(rows[4, :, 4] != 0)
    assert np.isfinite(series).all()
