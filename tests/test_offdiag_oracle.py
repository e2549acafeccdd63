"""Pin the off-diagonal forward-walking measurement on the CPU: the host
estimator against the reference's golden estimates, the measurement grid
against the reference's rules, the oracle's operator application against the
operator matrix (value = the column sum of the guide-transformed operator,
successors drawn in proportion to its entries), and the oracle's shadow
chain against the main chain it replays."""
import os

import numpy as np
import pytest

from helpers import basis, golden_system, index_of, jastrow_log_psi, spinhalf_matrix
from pygfmc_amd import SpinHalf, chain
from pygfmc_amd.estimators import estimate_off_diagonal, evaluate_observables
from pygfmc_amd.observables import Identity, lower_diag_operators
from pygfmc_amd.offdiagfw import check_even_p, measured_steps, measurement_plan
from pygfmc_amd.problem import lower

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden_spinhalf.npz")
GUIDES = ("none", "jastrow", "cps", "mlp", "exact")
CASES = [f"spinhalf_{g}_{m}" for g in GUIDES for m in ("cont_time", "multi_step", "single_step")]
OFF = ("spsm", "sx", "sfx")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def golden_raw(g, case):
    return dict(log_weights=g[f"{case}/raw/log_weights"], energy=g[f"{case}/raw/energy"],
                diag={op: g[f"{case}/raw/diag/{op}"] for op in ("zz", "sf")},
                off_diag={op: g[f"{case}/raw/off_diag/{op}"] for op in OFF},
                off_diag_norm={op: g[f"{case}/raw/off_diag_norm/{op}"] for op in OFF},
                off_diag_steps=g[f"{case}/raw/off_diag_steps"], p_array=[2, 4], m_array=[1, 2],
                diag_operators=["zz", "sf"], off_diag_operators=list(OFF))


# -- the host estimator against the golden estimates ------------------------------

@pytest.mark.parametrize("case", CASES)
def test_off_diagonal_estimates_reproduce_the_golden_values(golden, case):
    """With diagonal operators in the same run the windows start at
    i >= max(m) = 2 (the common trim)."""
    raw = golden_raw(golden, case)
    obs = estimate_off_diagonal([2, 4], raw["log_weights"], raw["off_diag"],
                                raw["off_diag_norm"], raw["off_diag_steps"], [1, 2],
                                m_max_common=2)
    for op in OFF:
        for p in (2, 4):
            np.testing.assert_allclose(obs[op][p], golden[f"{case}/obs/{op}/{p}"], rtol=1e-12,
                                       atol=1e-14)


@pytest.mark.parametrize("case", CASES[:6])
def test_evaluate_observables_reproduces_every_golden_estimate(golden, case):
    """The re-estimation from raw: energy, diagonal and off-diagonal
    operators together (common trim applied automatically)."""
    raw = golden_raw(golden, case)
    E, obs = evaluate_observables(raw)
    for p in (2, 4):
        np.testing.assert_allclose(E[p], golden[f"{case}/energy/{p}"], rtol=1e-12)
        for op in ("zz", "sf") + OFF:
            np.testing.assert_allclose(obs[op][p], golden[f"{case}/obs/{op}/{p}"],
                                       rtol=1e-12, atol=1e-14)


def test_planted_estimator_mistakes_miss_the_golden_values(golden):
    """Each plausible slip moves an estimate off the golden value: no common
    trim (it moves the windows of p = 2, m = 1), the shadow norm dropped, the
    main-chain weights off by one step."""
    case = "spinhalf_jastrow_cont_time"
    raw = golden_raw(golden, case)
    want = golden[f"{case}/obs/spsm/4"]
    args = ([2, 4], raw["log_weights"], raw["off_diag"], raw["off_diag_norm"],
            raw["off_diag_steps"], [1, 2])
    want2 = golden[f"{case}/obs/spsm/2"]
    assert abs(estimate_off_diagonal(*args, m_max_common=0)["spsm"][2] - want2) > 1e-9
    zero = {k: np.zeros_like(v) for k, v in raw["off_diag_norm"].items()}
    assert abs(estimate_off_diagonal(*args[:3], zero, *args[4:], m_max_common=2)["spsm"][4]
               - want) > 1e-6
    shifted = np.concatenate([[0.0], raw["log_weights"][:-1]])
    assert abs(estimate_off_diagonal([2, 4], shifted, *args[2:], m_max_common=2)["spsm"][4]
               - want) > 1e-9


def test_sfx_without_a_guide_is_exactly_the_bond_count(golden):
    """4 S^x S^x (1 + cos(pi/2)) on the 8 bonds: J_pm = J_pmpm = 1, so every
    configuration has operator value (column sum) 8 whatever its spins;
    without a guide every stored row is 8 exactly.  The estimate is not 8:
    the operator moves the walkers, and the shadow's weights differ from the
    main chain's."""
    for m in ("cont_time", "multi_step", "single_step"):
        case = f"spinhalf_none_{m}"
        assert np.all(golden[f"{case}/raw/off_diag/sfx"] == 8.0)
        assert abs(golden[f"{case}/obs/sfx/2"] - 8.0) > 0.1


def test_jackknife_drops_and_renumbers_the_rows_of_a_deleted_block(golden):
    """A replica of the off-diagonal estimate equals the estimator run on the
    chain with the block removed and the surviving rows renumbered."""
    case = "spinhalf_jastrow_single_step"
    raw = golden_raw(golden, case)
    _, obs, _, err = evaluate_observables(raw, n_blocks=20)
    assert all(np.isfinite(err[op][p]) and err[op][p] > 0 for op in OFF for p in (2, 4))
    lw, steps = raw["log_weights"], raw["off_diag_steps"]
    keep = np.ones(len(lw), dtype=bool)
    keep[30:45] = False
    rows = keep[steps]
    new_steps = (np.cumsum(keep) - 1)[steps[rows]]
    rep = estimate_off_diagonal([2, 4], lw[keep],
                                {k: v[rows] for k, v in raw["off_diag"].items()},
                                {k: v[rows] for k, v in raw["off_diag_norm"].items()},
                                new_steps, [1, 2], m_max_common=2)
    assert np.all(np.diff(new_steps) == 1) and len(new_steps) == 285
    assert abs(rep["spsm"][4] - obs["spsm"][4]) < 0.05


# -- the measurement grid -------------------------------------------------------------

def test_measured_steps_follow_the_stride_and_equal_window_rules():
    plan = measurement_plan(100, [8, 16], 1, 4, return_raw=True)
    t = np.flatnonzero(measured_steps(100, plan))
    np.testing.assert_array_equal(t, np.arange(3, 100, 4))
    assert plan.off_diag_offset == 3 and plan.gate_off_diag
    plan = measurement_plan(100, [8, 16], 1, 4, return_raw=False)
    t = np.flatnonzero(measured_steps(100, plan))
    np.testing.assert_array_equal(t, [k for k in range(16, 92) if (k + 1) % 4 == 0])
    assert plan.off_diag_offset == 19 and plan.equal_windows
    plan = measurement_plan(100, [8, 16], 1, 1, return_raw=True)
    assert measured_steps(100, plan).all() and not plan.gate_off_diag


def test_measurement_plan_rejects_what_the_reference_rejects():
    with pytest.raises(ValueError, match="1.5"):
        measurement_plan(24, [16], 1, 1, return_raw=True)
    with pytest.raises(ValueError, match="multiple of off_diag_stride"):
        measurement_plan(26, [16], 1, 50, return_raw=False)
    with pytest.raises(ValueError, match="divisible by 2"):
        check_even_p({"sx": SpinHalf.create(chain(4).edges, h_x=-1.0, sign_check=False)},
                     [3, 4])
    check_even_p({"one": Identity(4)}, [3, 4])


# -- the oracle's operator application ----------------------------------------------

def _ops():
    e = chain(8).edges
    return {"spsm": SpinHalf.create(e, J_x=2.0, J_y=2.0, J_z=0.0, sign_check=False,
                                    dtype=np.float64),
            "sx": SpinHalf.create(e, h_x=-1.0, sign_check=False, dtype=np.float64),
            "mixed": SpinHalf.create(e, J_x=1.5, J_y=0.5, J_z=0.8, h_x=-0.6, sign_check=False,
                                     dtype=np.float64)}


@pytest.mark.parametrize("name", ["spsm", "sx", "mixed"])
@pytest.mark.parametrize("guided", [False, True])
def test_operator_value_is_the_guide_transformed_column_sum(orc, name, guided):
    """v(x) = sum_x' O_x'x psi(x') / psi(x) over the Kronecker matrix of O,
    every configuration of 8 sites."""
    _, G = golden_system()
    G = G if guided else None
    op = _ops()[name]
    P = orc.Problem(lower(op, G))
    cfg = basis(8)
    _, vals = P.apply_operator(orc.pack(cfg), 5, 3, 16)
    M = spinhalf_matrix(op).tocsc()
    idx = index_of(cfg)
    pos = np.empty_like(idx)
    pos[idx] = np.arange(len(idx))
    lp = jastrow_log_psi(G, cfg) if guided else np.zeros(len(cfg))
    for c in range(0, len(cfg), 3):
        col = M[:, idx[c]]
        want = float(np.sum(col.data * np.exp(lp[pos[col.indices]] - lp[c])))
        assert abs(vals[c] - want) <= 1e-12 * (1 + abs(want)), (c, vals[c], want)


def test_operator_successors_follow_the_operator_entries(orc):
    """One configuration replicated 40 000 times (distinct Philox streams):
    the successor frequencies match O_x'x psi(x') / (psi(x) v) (chi-square)."""
    _, G = golden_system()
    op = _ops()["mixed"]
    P = orc.Problem(lower(op, G))
    x = np.array([[1, -1, -1, 1, 1, 1, -1, 1]], dtype=np.float64)
    n = 40000
    out, vals = P.apply_operator(np.repeat(orc.pack(x), n, axis=0), 9, 1, 17)
    assert np.all(vals == vals[0])
    got = index_of(orc.unpack(out, 8))
    cfg = basis(8)
    idx = index_of(cfg)
    pos = np.empty_like(idx)
    pos[idx] = np.arange(len(idx))
    col = spinhalf_matrix(op).tocsc()[:, index_of(x)[0]]
    lp = jastrow_log_psi(G, cfg)
    c0 = pos[index_of(x)[0]]
    p = col.data * np.exp(lp[pos[col.indices]] - lp[c0]) / vals[0]
    np.testing.assert_allclose(p.sum(), 1.0, rtol=1e-12)
    counts = np.array([np.sum(got == r) for r in col.indices])
    assert counts.sum() == n
    chi2 = np.sum((counts - n * p) ** 2 / (n * p))
    assert chi2 < len(p) + 6 * np.sqrt(2 * len(p)), chi2


def test_diagonal_only_operator_keeps_the_configuration(orc):
    """No off-diagonal slot: the configuration stays and the value is the
    signed diagonal element."""
    e = chain(8).edges
    zz = SpinHalf.create(e, J_z=-1.0, sign_check=False, dtype=np.float64)
    P = orc.Problem(lower(zz, None))
    spins = orc.init_random(8, 50, 3)
    out, vals = P.apply_operator(spins, 1, 1, 16)
    np.testing.assert_array_equal(out, spins)
    cfg = orc.unpack(spins, 8)
    e = np.asarray(e)
    want = -0.25 * np.sum(cfg[:, e[:, 0]] * cfg[:, e[:, 1]], axis=1)
    np.testing.assert_allclose(vals, want, rtol=0, atol=1e-15)
    assert np.any(vals < 0)


def test_zero_weight_configuration_stays_with_value_zero(orc):
    """spsm on a fully polarised configuration: every rate is 0, so the
    walker stays with value 0 (the zero-weight guard)."""
    P = orc.Problem(lower(_ops()["spsm"], None))
    spins = orc.pack(np.ones((4, 8)))
    out, vals = P.apply_operator(spins, 1, 1, 16)
    np.testing.assert_array_equal(out, spins)
    assert np.all(vals == 0.0)


# -- the oracle's shadow chain ---------------------------------------------------------

@pytest.mark.parametrize("method", [0, 1, 2])
def test_identity_shadow_replays_the_main_chain(orc, method):
    """From the main chain's own ensemble with the main chain's future
    draws, the shadow chain's log G are the main chain's, step for step, so
    the identity's estimate is exactly 1 (test_identity_forward_walk.py)."""
    H, G = golden_system()
    h = orc.Problem(lower(H, G))
    beta, lam = (0.2, 0.0) if method == 0 else (4.0, 5.0)
    spins = orc.init_random(8, 48, 2)
    spins, _, _ = orc.run_chain(h, spins, 5, beta, 0.0, 77, 0, method=method, lam=lam)
    _, _, series = orc.run_chain(h, spins.copy(), 6, beta, 0.0, 77, 5, method=method, lam=lam)
    vals = np.ones(48)
    fw, norm, logG, means = orc.forward_walk_operator(h, None, spins, [0, 1, 3, 6], beta, 0.0,
                                                      77, 4, 16, values=vals, method=method,
                                                      lam=lam, history=True)
    np.testing.assert_array_equal(logG[1:], series[:, 0])
    np.testing.assert_array_equal(fw, 1.0)
    np.testing.assert_array_equal(norm, [0.0, series[0, 0], np.sum(series[:3, 0]),
                                         np.cumsum(series[:, 0])[-1]])
    one = Identity(8)
    L = lower_diag_operators({"one": one}, 8)
    assert L.n_cols == 1 and L.c[0] == 1.0


def test_shadow_values_follow_their_walkers(orc):
    """Distinct starting values: after the shadow steps each mean equals the
    mean of the start values indexed through the chain of parents, which
    the oracle's history reproduces."""
    H, G = golden_system()
    h = orc.Problem(lower(H, G))
    spins = orc.init_random(8, 32, 4)
    vals = np.arange(32, dtype=np.float64)
    fw, norm = orc.forward_walk_operator(h, None, spins, [0, 2], 0.2, 0.0, 5, 10, 16,
                                         values=vals)
    assert fw[0] == np.mean(vals)
    s = spins.copy()
    v = vals.copy()
    e = np.zeros(32)
    for k in (11, 12):
        s, lw, e, _ = h.propagate(s, np.zeros(32), e, 0.2, 0.0, 5, k, 0)
        par, _, _, _ = orc.reconfigure(lw, e, orc.reconfig_psi(5, k, 1))
        s, e = orc.gather(8, s, e, par)
        v = v[par]
    assert fw[1] == np.mean(v)
