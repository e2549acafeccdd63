"""Off-diagonal operators measured by forward walking on the GPU: the
operator application and the shadow chain against the oracle, the identity
check (the shadow replays the main chain, so the estimate is exactly 1), the
run's estimates against their re-estimation, and the pure ground-state
expectation values of exact diagonalisation."""
import sys

import numpy as np
import pytest
import torch

from helpers import basis, golden_system, index_of, spinhalf_matrix
from pygfmc_amd import (MLPSpinHalf, RBMSpinHalf, SpinHalf, SzSzStructureFactor, chain,
                        evaluate_observables, prepare_walkers_from_configs,
                        prepare_walkers_random, run)
from pygfmc_amd import _lib
from pygfmc_amd.engine import set_method
from pygfmc_amd.observables import Identity
from pygfmc_amd.offdiagfw import OffDiagMeasurer
from pygfmc_amd.problem import Problem, lower

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(spins):
    return torch.as_tensor(np.ascontiguousarray(spins).view(np.int32), device="cuda")


def ops8(dtype=np.float64):
    e = chain(8).edges
    return {"spsm": SpinHalf.create(e, J_x=2.0, J_y=2.0, J_z=0.0, sign_check=False, dtype=dtype),
            "sx": SpinHalf.create(e, h_x=-1.0, sign_check=False, dtype=dtype),
            "sfx": SpinHalf.create(e, J_x=4.0, J_y=0.0, J_z=0.0, sign_check=False,
                                   dtype=dtype),
            "mixed": SpinHalf.create(e, J_x=1.5, J_y=0.5, J_z=0.8, h_x=-0.6, sign_check=False,
                                     dtype=dtype)}


def apply_dev(prob, spins, seed, step, tag):
    nw = spins.shape[0]
    prob.reserve(nw)      # the MLP guide's caller-owned workspace (no-op otherwise)
    s = to_dev(spins)
    out = torch.empty_like(s)
    vals = torch.empty(nw, dtype=torch.float64, device="cuda")
    work = torch.empty(int(_lib.load().gfmc_apply_operator_work_doubles(prob.handle, nw)),
                       dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().gfmc_apply_operator(prob.handle, nw, _lib.ptr(s), seed, step, tag,
                                               _lib.ptr(out), _lib.ptr(vals), _lib.ptr(work),
                                               _lib.stream()), "gfmc_apply_operator")
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32), vals.cpu().numpy()


@pytest.mark.parametrize("name", ["spsm", "sx", "sfx", "mixed"])
@pytest.mark.parametrize("guided", [False, True])
def test_apply_operator_matches_oracle_fp64(orc, name, guided):
    say("apply operator", name, guided)
    _, G = golden_system()
    G = G if guided else None
    op = ops8()[name]
    spins = orc.init_random(8, 1000, 12)
    s_ref, v_ref = orc.Problem(lower(op, G)).apply_operator(spins, 31, 7, 17)
    s, v = apply_dev(Problem(op, G, _lib.GFMC_FP64), spins, 31, 7, 17)
    np.testing.assert_array_equal(s, s_ref)
    np.testing.assert_allclose(v, v_ref, rtol=1e-12, atol=1e-14)
    assert np.mean(np.any(s != spins, axis=1)) > 0.5 or name == "mixed"
    if name == "sfx" and not guided:
        assert np.all(v == 8.0)


@pytest.mark.parametrize("guide", ["jastrow", "rbm", "mlp"])
def test_apply_operator_matches_oracle_fp32(orc, guide):
    """FP32 rates (and the network guides' BF16 weights): values within
    2e-5, the successors equal where the draws do not sit on a rate
    boundary."""
    say("apply operator fp32", guide)
    e = chain(8).edges
    if guide == "jastrow":
        G = golden_system(dtype=np.float32)[1]
    elif guide == "rbm":
        G = RBMSpinHalf(8, alpha=4, seed=3, scale=0.5)
    else:
        G = MLPSpinHalf(8, (64, 32), seed=3, scale=1.5, out_scale=1.0)
    op = SpinHalf.create(e, J_x=1.5, J_y=0.5, J_z=0.8, h_x=-0.6, sign_check=False,
                         dtype=np.float32)
    spins = orc.init_random(8, 800, 13)
    s_ref, v_ref = orc.Problem(lower(op, G)).apply_operator(spins, 3, 9, 18)
    s, v = apply_dev(Problem(op, G, _lib.GFMC_FP32), spins, 3, 9, 18)
    np.testing.assert_allclose(v, v_ref, rtol=2e-5, atol=1e-6)
    assert np.mean(np.all(s == s_ref, axis=1)) > 0.99


CHAIN_CASES = [("cont_time", 0.2, None, False), ("cont_time", 0.2, None, True),
               ("single_step", 3, 5.0, True), ("multi_step", 4, 5.0, True)]


@pytest.mark.parametrize("method,step,lam,guided", CHAIN_CASES,
                         ids=[f"{c[0]}-{'jastrow' if c[3] else 'none'}" for c in CHAIN_CASES])
def test_forward_walk_matches_oracle(orc, method, step, lam, guided):
    say("forward walk", method, guided)
    H, G = golden_system()
    G = G if guided else None
    nw = 300
    spins = orc.init_random(8, nw, 21)
    meth = {"cont_time": 0, "single_step": 1, "multi_step": 2}[method]
    h_ref = orc.Problem(lower(H, G))
    m_list = [0, 1, 3, 3, 7]
    prob = Problem(H, G, _lib.GFMC_FP64)
    set_method(prob, method, lam)
    prob.set_lanes(64)
    od = OffDiagMeasurer({"spsm": ops8()["spsm"], "sx": ops8()["sx"]}, prob, G, m_list, nw,
                         "cuda")
    fw, nrm = od.rows(1, "cuda")
    od.measure(to_dev(spins), float(step), 0.4, 55, 100, fw[0], nrm[0])
    torch.cuda.synchronize()
    fw, nrm = fw.cpu().numpy()[0], nrm.cpu().numpy()[0]
    for i, name in enumerate(("spsm", "sx")):
        op_ref = orc.Problem(lower(ops8()[name], G))
        f_ref, n_ref = orc.forward_walk_operator(h_ref, op_ref, spins, m_list, float(step), 0.4,
                                                 55, 100, _lib.TAG_OPERATOR + i, method=meth,
                                                 lam=lam or 0.0)
        np.testing.assert_allclose(fw[i], f_ref, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(nrm[i], n_ref, rtol=1e-12, atol=1e-12)
    assert np.all(nrm[:, 2:] != 0.0)


def test_diagonal_start_matches_oracle(orc):
    """A diagonal operator starts the shadow from its values (configurations
    unchanged)."""
    say("forward walk diagonal")
    H, G = golden_system()
    nw = 257
    spins = orc.init_random(8, nw, 22)
    zz = SpinHalf.create(chain(8).edges, J_z=1.0, dtype=np.float64)
    prob = Problem(H, G, _lib.GFMC_FP64)
    prob.set_lanes(64)
    od = OffDiagMeasurer({"zz": zz}, prob, G, [0, 2, 5], nw, "cuda")
    fw, nrm = od.rows(1, "cuda")
    od.measure(to_dev(spins), 0.2, 0.1, 8, 3, fw[0], nrm[0])
    torch.cuda.synchronize()
    vals = spinhalf_matrix(zz).diagonal()[index_of(orc.unpack(spins, 8))]
    f_ref, n_ref = orc.forward_walk_operator(orc.Problem(lower(H, G)), None, spins, [0, 2, 5],
                                             0.2, 0.1, 8, 3, 0, values=vals)
    np.testing.assert_allclose(fw.cpu().numpy()[0, 0], f_ref, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(nrm.cpu().numpy()[0, 0], n_ref, rtol=1e-12, atol=1e-12)


def _tfim(L):
    edges = np.asarray([(i, (i + 1) % L) for i in range(L)], dtype=np.int32)
    return SpinHalf.create(edges, J_x=0.0, J_y=0.0, J_z=-4.0, h_x=1.5, dtype=np.float64)


@pytest.mark.parametrize("method,step,lam,stride", [("single_step", 1, 10.0, 1),
                                                    ("multi_step", 8, 10.0, 4),
                                                    ("cont_time", 0.2, None, 3)])
def test_identity_estimate_is_exactly_one(method, step, lam, stride):
    """The shadow chain consumes the main chain's own future draws, so from
    the main chain's ensemble it replays the main chain: every numerator
    term equals its denominator term and <identity>_p = 1 for every p and
    stride (test_identity_forward_walk.py:97-129)."""
    say("identity", method, stride)
    L = 6
    w = prepare_walkers_from_configs(np.ones((64, L)))
    _, _, obs, st, raw = run(w, 600, _tfim(L), step, None, [16, 32, 64, 128], method=method,
                             lam=lam, rng_key=3, off_diag_operators={"one": Identity(L)},
                             off_diag_stride=stride, warmup_steps=100, verbosity=0, lanes=64)
    for p, v in obs["one"].items():
        assert abs(v - 1.0) < 1e-9, (p, v)
    np.testing.assert_array_equal(raw["off_diag"]["one"], 1.0)
    steps = raw["off_diag_steps"]
    np.testing.assert_array_equal(steps, np.arange(stride - 1, 600, stride))
    # the shadow's log weight product is the main chain's, m steps ahead
    C = np.concatenate([[0.0], np.cumsum(raw["log_weights"])])
    ok = steps + 64 < 600
    want = C[steps[ok] + 1 + 64] - C[steps[ok] + 1]
    np.testing.assert_allclose(raw["off_diag_norm"]["one"][ok, 3], want, rtol=0, atol=1e-9)
    assert st["parameters/off_diag_stride"] == stride


def test_run_estimates_equal_their_re_estimation_and_equal_windows():
    """obs of the run equal evaluate_observables(raw) (diagonal and
    off-diagonal together, common trim); return_raw=False measures only the
    band every window can use."""
    say("offdiag run")
    H, G = golden_system()
    e = chain(8).edges
    ops = {"spsm": ops8()["spsm"], "sx": ops8()["sx"]}
    dops = {"zz": SpinHalf.create(e, J_z=1.0, dtype=np.float64)}
    w = prepare_walkers_random(256, 8, 5)
    kw = dict(rng_key=9, warmup_steps=50, verbosity=0, lanes=64, diag_operators=dops,
              off_diag_operators=ops)
    _, E, obs, st, raw = run(w, 120, H, 0.2, G, [2, 8], **kw)
    E2, obs2 = evaluate_observables(raw)
    for name in ("zz", "spsm", "sx"):
        for p in (2, 8):
            assert obs[name][p] == pytest.approx(obs2[name][p], rel=1e-12)
    assert raw["off_diag"]["spsm"].shape == (120, 2)
    _, _, obs3, st3 = run(w, 120, H, 0.2, G, [2, 8], return_raw=False, off_diag_stride=5,
                          **kw)
    assert st3["parameters/off_diag_equal_windows"] == 1
    assert st3["parameters/off_diag_offset"] == 9
    assert set(obs3) == {"zz", "spsm", "sx"}
    with pytest.raises(ValueError, match="divisible by 2"):
        run(w, 120, H, 0.2, G, [3, 8], **kw)
    with pytest.raises(ValueError, match="share name"):
        run(w, 120, H, 0.2, G, [2, 8], rng_key=9, verbosity=0, diag_operators={"sx": dops["zz"]},
            off_diag_operators=ops)
    bad = {"neg": SpinHalf.create(e, J_x=-1.0, J_y=-1.0, sign_check=False, dtype=np.float64)}
    with pytest.raises(ValueError, match="negative rate"):
        run(w, 120, H, 0.2, G, [2, 8], rng_key=9, verbosity=0, off_diag_operators=bad)
    with pytest.raises(TypeError):
        run(w, 120, H, 0.2, G, [2, 8], rng_key=9, verbosity=0,
            off_diag_operators={"sf": SzSzStructureFactor(np.arange(8.0)[:, None],
                                                          np.zeros((1, 1)))})


def test_forward_walking_reaches_the_pure_off_diagonal_values():
    """Chain of 8, J_x = J_y = -1, J_z = 1, h_x = 0.5, Jastrow guide: the
    forward-walking estimates of spsm (sum S+S- + h.c. on the bonds) and sx
    (sum S^x) at m = 12 shadow steps (imaginary time 2.4 per side; the gap
    is 1.24) reach the pure <psi_0|O|psi_0> of exact diagonalisation."""
    say("offdiag physics")
    H, G = golden_system()
    M = spinhalf_matrix(H).toarray()
    vals, vecs = np.linalg.eigh(M)
    psi = vecs[:, 0]
    ops = {"spsm": ops8()["spsm"], "sx": ops8()["sx"]}
    pure = {k: float(psi @ (spinhalf_matrix(op).toarray() @ psi)) for k, op in ops.items()}
    w = prepare_walkers_random(2048, 8, 7)
    _, E, obs, _, raw = run(w, 2400, H, 0.2, G, [4, 24], rng_key=8, warmup_steps=200,
                            verbosity=0, off_diag_operators=ops, off_diag_stride=2, lanes=64)
    _, o, _, oe = evaluate_observables(raw, n_blocks=20)
    for k in ops:
        say(f"  {k}: p=4 {o[k][4]:.5f} +- {oe[k][4]:.5f}, p=24 {o[k][24]:.5f} +- "
            f"{oe[k][24]:.5f}; pure {pure[k]:.5f}")
        assert abs(o[k][24] - pure[k]) < 5 * oe[k][24] + 0.01
    assert abs(E[24] - vals[0]) < 0.01


def test_continuation_equals_uninterrupted_run(tmp_path):
    """A run split at a multiple of the stride and continued from its
    checkpoint measures the same off-diagonal rows as the uninterrupted run
    (the shadow chains use absolute Philox positions); load_raw offsets the
    continuation's steps."""
    say("offdiag resume")
    from pygfmc_amd import load_raw, run_from_checkpoint, write_checkpoint
    H, G = golden_system()
    ops = {"spsm": ops8()["spsm"], "one": Identity(8)}
    kw = dict(rng_key=4, verbosity=0, lanes=64, number_of_outputs=4, off_diag_operators=ops,
              off_diag_stride=3)
    w = prepare_walkers_random(192, 8, 6)
    _, _, _, _, raw_all = run(w, 96, H, 0.2, G, [2, 6], warmup_steps=30, **kw)
    w1, _, _, st1, raw1 = run(w, 60, H, 0.2, G, [2, 6], warmup_steps=30, **kw)
    f = str(tmp_path / "o.npz")
    write_checkpoint(f, st1, w1, H, G, raw1)
    w2, _, _, st2, raw2 = run_from_checkpoint(f, H, G, n_steps=36, verbosity=0,
                                              number_of_outputs=4, off_diag_operators=ops)
    for name in ops:
        for grp in ("off_diag", "off_diag_norm"):
            np.testing.assert_array_equal(np.concatenate([raw1[grp][name], raw2[grp][name]]),
                                          raw_all[grp][name])
    np.testing.assert_array_equal(np.concatenate([raw1["off_diag_steps"],
                                                  raw2["off_diag_steps"] + 60]),
                                  raw_all["off_diag_steps"])
    write_checkpoint(f, st2, w2, H, G, raw2)
    both = load_raw(f)
    np.testing.assert_array_equal(both["off_diag_steps"], raw_all["off_diag_steps"])
    E_a, o_a = evaluate_observables(both)
    E_b, o_b = evaluate_observables(raw_all)
    for name in ops:
        for p in (2, 6):
            assert o_a[name][p] == o_b[name][p]
    with pytest.raises(ValueError, match="off-diagonal operators"):
        run_from_checkpoint(f, H, G, n_steps=12, verbosity=0)
