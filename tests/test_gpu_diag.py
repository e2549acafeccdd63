"""Diagonal operators measured by forward walking on the GPU: the column
kernel, the step loop with the ancestry ring against the oracle chain, HIP
graph replays, run()'s estimates, the lead-in of a continuation, and the
pure ground-state expectation values of exact diagonalisation."""
import sys

import numpy as np
import pytest
import torch

from helpers import basis, golden_system, index_of, spinhalf_matrix
from pygfmc_amd import (JastrowSpinHalf, SpinHalf, SzSzStructureFactor, chain,
                        evaluate_observables, prepare_walkers_random, prepare_walkers_sz, run,
                        run_from_checkpoint, square, write_checkpoint)
from pygfmc_amd import _lib
from pygfmc_amd.diagfw import DiagRecorder
from pygfmc_amd.engine import Chain, StepGraph, set_method
from pygfmc_amd.estimators import estimate_diagonal
from pygfmc_amd.observables import lower_diag_operators
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.walkers import Walker

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def chain_sf(N, n_k):
    return SzSzStructureFactor(np.arange(float(N))[:, None],
                               (2 * np.pi * np.arange(n_k) / N)[:, None], dtype=np.float64)


def op_sets():
    sq = square(4, 4)
    pos = np.array([(x, y) for x in range(4) for y in range(4)], dtype=float)
    ks = 2 * np.pi * np.array([(0, 0), (1, 0), (2, 2), (1, 3)], dtype=float) / 4
    rng = np.random.default_rng(4)
    return [
        ("chain8_zz_sf", 8, {"zz": SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64),
                             "sf": chain_sf(8, 3)}),
        ("chain40_fields", 40, {"d": SpinHalf.create(chain(40).edges,
                                                     J_z=rng.uniform(-1, 1, 40),
                                                     h_z=rng.uniform(-0.5, 0.5, 40),
                                                     dtype=np.float64)}),
        ("square4_sf", 16, {"sf": SzSzStructureFactor(pos, ks),
                            "zz": SpinHalf.create(sq.edges, J_z=-0.75, dtype=np.float64)}),
        ("chain100_sf", 100, {"sf": chain_sf(100, 7)}),
    ]


OPS = op_sets()


@pytest.mark.parametrize("name,N,ops", OPS, ids=[o[0] for o in OPS])
def test_column_values_match_oracle(orc, name, N, ops):
    say("diag values", name)
    spins = orc.init_random(N, 777, 3)
    low = lower_diag_operators(ops, N)
    want = orc.Diag(low).values(spins)
    rec = DiagRecorder(ops, N, [0], 777, "cuda")
    got = rec.values(to_dev(spins)).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


STEP_CASES = [("cont_time", 0.2, None, [0, 1, 1, 3, 6]), ("single_step", 3, 4.0, [1, 2, 5]),
              ("multi_step", 4, 4.0, [2, 2])]


@pytest.mark.parametrize("method,step,lam,m_list", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_step_loop_rows_match_oracle_chain(orc, method, step, lam, m_list):
    """Two consecutive step-loop calls (the ring continues across them)
    against the oracle's chain with the ring of driver.py:240-288."""
    say("diag step loop", method, m_list)
    H, G = golden_system()
    nw, n1, n2 = 203, 5, 9
    prob = Problem(H, G, _lib.GFMC_FP64)
    set_method(prob, method, lam)
    ops = OPS[0][2]
    d = orc.Diag(lower_diag_operators(ops, 8))
    spins = orc.init_random(8, nw, 6)
    beta, alw, seed, step0 = float(step), 0.3, 99, 40
    meth = {"cont_time": 0, "single_step": 1, "multi_step": 2}[method]
    s_ref, e_ref, ser_ref, rows_ref = orc.diag_chain(orc.Problem(lower(H, G)), d, m_list, spins,
                                                     n1 + n2, beta, alw, seed, step0,
                                                     method=meth, lam=lam or 0.0)
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ch = Chain(prob, Walker.from_device(to_dev(spins), z, z.clone(), 8))
    rec = DiagRecorder(ops, 8, m_list, nw, "cuda")
    r1 = rec.series(n1, "cuda")
    r2 = rec.series(n2, "cuda")
    a = ch.steps(n1, beta, alw, seed, step0, diag=rec, dseries=r1)
    b = ch.steps(n2, beta, alw, seed, step0 + n1, diag=rec, dseries=r2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ch.spins.cpu().numpy().view(np.uint32), s_ref)
    ser = np.concatenate([a.cpu().numpy(), b.cpu().numpy()])
    np.testing.assert_allclose(ser, ser_ref, rtol=1e-12, atol=1e-12)
    rows = np.concatenate([r1.cpu().numpy(), r2.cpu().numpy()])
    np.testing.assert_allclose(rows, rows_ref, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(rows) > rows.size // 2


def test_record_entry_point_matches_step_loop():
    """gfmc_diag_record after hand-issued steps equals the fused step loop."""
    say("diag record")
    H, G = golden_system()
    nw = 300
    prob = Problem(H, G, _lib.GFMC_FP64)
    ops = OPS[0][2]
    w = prepare_walkers_random(nw, 8, 2)
    ch1, ch2 = Chain(prob, w), Chain(prob, w)
    m_list = [1, 4]
    rec1 = DiagRecorder(ops, 8, m_list, nw, "cuda")
    rec2 = DiagRecorder(ops, 8, m_list, nw, "cuda")
    rows1 = rec1.series(6, "cuda")
    ch1.steps(6, 0.2, 0.3, 5, 0, diag=rec1, dseries=rows1)
    rows2 = rec2.series(6, "cuda")
    out2 = torch.empty(2, dtype=torch.float64, device="cuda")
    for t in range(6):
        ch2.step(0.2, 0.3, 5, t, out2, evals=False)
        rec2.record(ch2.spins, ch2.parents, rows2[t])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(rows1.cpu().numpy(), rows2.cpu().numpy())


def test_graph_replays_equal_eager_diag_steps():
    say("diag graph")
    H, G = golden_system()
    nw, n = 512, 7
    ops = OPS[0][2]
    m_list = [0, 2, 9]
    w = prepare_walkers_random(nw, 8, 4)
    pe, pg = Problem(H, G, _lib.GFMC_FP64), Problem(H, G, _lib.GFMC_FP64)
    eager, graphed = Chain(pe, w), Chain(pg, w)
    re_, rg = DiagRecorder(ops, 8, m_list, nw, "cuda"), DiagRecorder(ops, 8, m_list, nw, "cuda")
    g = StepGraph(graphed, n, 0.2, 0.3, 17, diag=rg)
    for rep in range(3):
        de = re_.series(n, "cuda")
        eager.steps(n, 0.2, 0.3, 17, 1000 + rep * n, diag=re_, dseries=de)
        dg = torch.empty_like(de)
        g.launch(1000 + rep * n, dout=dg)
        x = torch.randn(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(x)
        np.testing.assert_array_equal(dg.cpu().numpy(), de.cpu().numpy())
        np.testing.assert_array_equal(graphed.spins.cpu().numpy(), eager.spins.cpu().numpy())


def test_run_reports_estimates_of_its_series():
    say("diag run")
    H, G = golden_system()
    ops = {"zz": SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64),
           "sf": chain_sf(8, 3)}
    w = prepare_walkers_random(256, 8, 1)
    _, E, obs, st, raw = run(w, 400, H, 0.2, G, [2, 4, 8], rng_key=3, warmup_steps=40,
                             verbosity=0, diag_operators=ops, lanes=64)
    assert raw["diag"]["zz"].shape == (400, 3) and raw["diag"]["sf"].shape == (400, 3, 3)
    assert raw["additional_diag"]["zz"].shape == (4, 3)
    assert raw["additional_diag"]["sf"].shape == (4, 3, 3)
    want = estimate_diagonal([2, 4, 8], raw["log_weights"], raw["diag"])
    for name in ops:
        for p in (2, 4, 8):
            np.testing.assert_array_equal(obs[name][p], want[name][p])
    e2, o2, ee, oe = evaluate_observables(raw, n_blocks=20)
    assert oe["sf"][4].shape == (3,) and np.all(oe["sf"][4] > 0) and oe["zz"][2] > 0
    assert list(st["parameters/diag_operators"]) == [b"zz", b"sf"]
    # the graphed production equals eager launches
    _, _, obs_e, _, raw_e = run(w, 400, H, 0.2, G, [2, 4, 8], rng_key=3, warmup_steps=40,
                                verbosity=0, diag_operators=ops, lanes=64, use_graph=False)
    np.testing.assert_array_equal(raw_e["diag"]["sf"], raw["diag"]["sf"])
    # the Hamiltonian itself is no measurement operator: its rates are <= 0
    with pytest.raises(ValueError, match="negative rate"):
        run(w, 50, H, 0.2, G, [2], rng_key=3, warmup_steps=8, verbosity=0,
            off_diag_operators={"x": H})


def test_continuation_with_lead_in_equals_uninterrupted_run(tmp_path):
    say("diag resume")
    L = chain(24)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(24, L.edges, beta=0.3, dtype=np.float64)
    ops = {"zz": SpinHalf.create(L.edges, J_z=1.0, dtype=np.float64), "sf": chain_sf(24, 4)}
    kw = dict(rng_key=21, verbosity=0, lanes=64, number_of_outputs=6, diag_operators=ops)
    w = prepare_walkers_sz(384, 24, 8)
    _, _, _, st_all, raw_all = run(w, 240, H, 0.2, G, [4, 10], warmup_steps=40, **kw)
    w1, _, _, st1, raw1 = run(w, 150, H, 0.2, G, [4, 10], warmup_steps=40, **kw)
    f = str(tmp_path / "d.npz")
    write_checkpoint(f, st1, w1, H, G, raw1)
    _, _, _, _, raw2 = run_from_checkpoint(f, H, G, n_steps=90, verbosity=0,
                                           number_of_outputs=5, diag_operators=ops)
    for name in ops:
        np.testing.assert_array_equal(np.concatenate([raw1["diag"][name], raw2["diag"][name]]),
                                      raw_all["diag"][name])
    with pytest.raises(ValueError):
        run_from_checkpoint(f, H, G, n_steps=10, verbosity=0)


def test_forward_walking_reaches_the_pure_ground_state_values():
    """No guide: the walkers sample psi_0, the m = 1 rows are close to the
    mixed estimate sum psi_0 O / sum psi_0, and forward walking over m = 20
    and 30 steps (imaginary time 4 and 6, the gap is 1.24) reaches the pure
    <psi_0|O|psi_0> of exact diagonalisation."""
    say("diag physics")
    H, _ = golden_system(guide=False)
    zz = SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64)
    sf = chain_sf(8, 5)
    M = spinhalf_matrix(H).toarray()
    vals, vecs = np.linalg.eigh(M)
    cfg = basis(8)
    psi = vecs[index_of(cfg), 0]
    psi = psi * np.sign(psi.sum())
    Z = spinhalf_matrix(zz).diagonal()[index_of(cfg)]
    S = sf.diagonal_weight(cfg)
    pure_zz, mixed_zz = float(psi ** 2 @ Z), float(psi @ Z / psi.sum())
    pure_sf = psi ** 2 @ S
    w = prepare_walkers_random(4096, 8, 7)
    _, E, obs, _, raw = run(w, 3000, H, 0.2, None, [2, 40, 60], rng_key=8, warmup_steps=200,
                            verbosity=0, diag_operators={"zz": zz, "sf": sf}, lanes=64)
    _, o, _, oe = evaluate_observables(raw, n_blocks=20)
    say("  zz", {p: (o["zz"][p], oe["zz"][p]) for p in (2, 40, 60)}, "pure", pure_zz,
        "mixed", mixed_zz)
    say("  sf", o["sf"][60], oe["sf"][60], "pure", pure_sf)
    assert abs(o["zz"][2] - mixed_zz) < abs(o["zz"][2] - pure_zz)
    for p in (40, 60):
        assert abs(o["zz"][p] - pure_zz) < 5 * oe["zz"][p] + 0.01
        assert np.all(np.abs(o["sf"][p] - pure_sf) < 5 * oe["sf"][p] + 0.02)
    assert abs(E[60] - vals[0]) < 0.01
