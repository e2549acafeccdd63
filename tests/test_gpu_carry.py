"""Carried slot weights on the GPU: a step loop that starts each walker from
its parent's stored weights gives bit-identical results to one that
re-evaluates every slot (all propagators, lane counts changed between
calls, graph replays, a diagonal record in the same loop, single
three-launch steps in between)."""
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import JastrowSpinHalf, SpinHalf, SzSzStructureFactor, chain, square
from pygfmc_amd import _lib
from pygfmc_amd.diagfw import DiagRecorder
from pygfmc_amd.engine import Chain, StepGraph, set_method
from pygfmc_amd.problem import Problem
from pygfmc_amd.walkers import prepare_walkers_random, prepare_walkers_sz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(H, G, prec, w):
    """Two chains on one problem: one carrying, one re-evaluating."""
    p = Problem(H, G, prec)
    a, b = Chain(p, w), Chain(p, w)
    b.carry = None
    return p, a, b


def same(a, b):
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.spins.cpu().numpy(), b.spins.cpu().numpy())
    np.testing.assert_array_equal(a.log_w.cpu().numpy(), b.log_w.cpu().numpy())
    np.testing.assert_array_equal(a.e_loc.cpu().numpy(), b.e_loc.cpu().numpy())


CASES = [("cont_time", 0.2, None), ("single_step", 3, 5.0), ("multi_step", 4, 5.0)]


@pytest.mark.parametrize("lattice", ["square8", "chain40"])
@pytest.mark.parametrize("prec", [_lib.GFMC_FP64, _lib.GFMC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("method,step,lam", CASES, ids=[c[0] for c in CASES])
def test_carried_weights_are_bit_identical(lattice, prec, method, step, lam):
    say("carry", lattice, method, prec)
    L = square(8, 8) if lattice == "square8" else chain(40)
    N = 64 if lattice == "square8" else 40
    dt = np.float64 if prec == _lib.GFMC_FP64 else np.float32
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=dt)
    G = JastrowSpinHalf(N, L.edges, beta=0.3, dtype=dt)
    w = prepare_walkers_sz(512, N, 3)
    p, a, b = pair(H, G, prec, w)
    assert a.carry is not None
    set_method(p, method, lam)
    if lattice == "chain40":
        assert set(p.lanes_ok) == {64, 32, 16}
    for lanes, n, s0 in ((64, 7, 0), (64, 5, 7), (32, 6, 12), (16, 4, 18), (64, 3, 22)):
        if lanes not in p.lanes_ok:
            continue
        p.set_lanes(lanes)
        sa = a.steps(n, step, 0.1, 9, s0, evals=True)
        sb = b.steps(n, step, 0.1, 9, s0, evals=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sa.cpu().numpy(), sb.cpu().numpy())
        same(a, b)
    np.testing.assert_array_equal(a.evals.cpu().numpy(), b.evals.cpu().numpy())


def test_graph_replays_and_single_steps_with_carry():
    """Graph replays continue the carry, and so do single three-launch steps
    (gfmc_propagate_carry) between step loops."""
    say("carry graph")
    L = square(6, 6)
    H = SpinHalf.create(L.edges, -1.0, -0.7, 1.0, h_x=0.3, dtype=np.float64)
    G = JastrowSpinHalf(36, L.edges, beta=0.25, h=0.05, dtype=np.float64)
    w = prepare_walkers_random(300, 36, 5)
    p, a, b = pair(H, G, _lib.GFMC_FP64, w)
    p.set_lanes(64)
    ga, gb = StepGraph(a, 5, 0.2, 0.2, 13), StepGraph(b, 5, 0.2, 0.2, 13)
    for rep in range(3):
        ra = ga.launch(100 + 5 * rep).clone()
        rb = gb.launch(100 + 5 * rep).clone()
        x = torch.randn(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(x)
        np.testing.assert_array_equal(ra.cpu().numpy(), rb.cpu().numpy())
        same(a, b)
    # eager loops around single steps (no graph: a single step swaps the
    # chain's buffers)
    p2, c, d = pair(H, G, _lib.GFMC_FP64, w)
    out2c = torch.empty(2, dtype=torch.float64, device="cuda")
    out2d = torch.empty(2, dtype=torch.float64, device="cuda")
    for rep in range(3):
        c.steps(4, 0.2, 0.2, 13, 10 * rep)
        d.steps(4, 0.2, 0.2, 13, 10 * rep)
        c.step(0.2, 0.2, 13, 10 * rep + 4, out2c, evals=False)
        d.step(0.2, 0.2, 13, 10 * rep + 4, out2d, evals=False)
        same(c, d)
    c.steps(4, 0.2, 0.2, 13, 40)
    d.steps(4, 0.2, 0.2, 13, 40)
    same(c, d)


def test_carry_with_diagonal_record():
    say("carry diag")
    L = chain(40)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(40, L.edges, beta=0.3, dtype=np.float64)
    ops = {"zz": SpinHalf.create(L.edges, J_z=1.0, dtype=np.float64),
           "sf": SzSzStructureFactor(np.arange(40.0)[:, None],
                                     (2 * np.pi * np.arange(3) / 40)[:, None])}
    w = prepare_walkers_sz(256, 40, 2)
    p, a, b = pair(H, G, _lib.GFMC_FP64, w)
    ra, rb = DiagRecorder(ops, 40, [0, 2, 5], 256, "cuda"), DiagRecorder(ops, 40, [0, 2, 5], 256,
                                                                        "cuda")
    da, db = ra.series(12, "cuda"), rb.series(12, "cuda")
    a.steps(12, 0.2, 0.0, 3, 50, diag=ra, dseries=da)
    b.steps(12, 0.2, 0.0, 3, 50, diag=rb, dseries=db)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(da.cpu().numpy(), db.cpu().numpy())
    same(a, b)


def test_paths_without_records_carry_nothing():
    """The network guides and non-uniform couplings take the generic paths:
    no carry buffer."""
    from pygfmc_amd import RBMSpinHalf
    L = chain(12)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    p = Problem(H, RBMSpinHalf(12, alpha=2, seed=1), _lib.GFMC_FP32)
    assert int(_lib.load().gfmc_carry_bytes(p.handle, 100)) == 0
    rng = np.random.default_rng(0)
    jxy = -rng.uniform(0.5, 1.5, 12)
    H2 = SpinHalf.create(L.edges, jxy, jxy, rng.uniform(0.5, 1, 12), dtype=np.float64)
    p2 = Problem(H2, None, _lib.GFMC_FP64)
    assert int(_lib.load().gfmc_carry_bytes(p2.handle, 100)) == 0
    w = prepare_walkers_random(64, 12, 1)
    c = Chain(p2, w)
    c.steps(3, 0.2, 0.0, 1, 0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", [_lib.GFMC_FP64, _lib.GFMC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("lanes", [64, 32])
def test_walkers_changed_outside_the_loop_are_re_evaluated(prec, lanes):
    """Spins written into a carrying chain between step loops (here: half of
    the walkers replaced by other configurations) do not reuse their
    parents' stored weights: the chain then steps exactly like a fresh
    chain started from the edited walkers."""
    say("carry after an edit", prec, lanes)
    L = square(8, 8)
    dt = np.float64 if prec == _lib.GFMC_FP64 else np.float32
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=dt)
    G = JastrowSpinHalf(64, L.edges, beta=0.3, dtype=dt)
    w = prepare_walkers_sz(512, 64, 3)
    p = Problem(H, G, prec)
    p.set_lanes(lanes)
    a = Chain(p, w)
    a.steps(4, 0.2, 0.0, 71, 0)
    other = prepare_walkers_sz(512, 64, 99)
    a.spins[:256].copy_(other.spins[:256])
    from pygfmc_amd.walkers import Walker
    b = Chain(p, Walker.from_device(a.spins.clone(), a.log_w.clone(), a.e_loc.clone(), 64))
    sa = a.steps(3, 0.2, 0.0, 71, 4)
    sb = b.steps(3, 0.2, 0.0, 71, 4)
    same(a, b)
    np.testing.assert_array_equal(sa.cpu().numpy(), sb.cpu().numpy())
    # and the untouched walkers did use the carry: the edited chain without
    # the check would differ (the stored rows of the replaced walkers'
    # parents are stale) -- the carry buffer was live during these steps
    assert a.carry is not None and int(a.carry[:4].view(torch.int32)[0].item()) == 1


def test_tied_rbm_carry_is_bit_identical():
    """The tied-RBM kernel carries slot weights and theta (exact in FP64) from
    step to step: a carrying chain equals a re-evaluating one bitwise, also
    across a parameter change (rows of the old parameters lapse)."""
    from pygfmc_amd import RBMSpinHalf
    from pygfmc_amd.training import flat_params
    say("carry tied rbm")
    L = square(6, 6)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    T = L.extras["translations"]
    G0 = RBMSpinHalf.symmetric(T, 4, b=[0.1, -0.2, 0.05, 0.3], seed=3, scale=0.2)
    w = prepare_walkers_sz(700, 36, 2)
    p, a, b = pair(H, G0, _lib.GFMC_FP32, w)
    assert p.tied is not None and a.carry is not None
    G1 = RBMSpinHalf.symmetric(T, 4, b=[-0.1, 0.2, 0.0, 0.1], seed=4, scale=0.25)
    for n, s0, G in ((6, 0, None), (5, 6, None), (4, 11, G1), (3, 15, None)):
        if G is not None:
            flat = torch.as_tensor(flat_params(G), device="cuda")
            _lib.check(_lib.load().gfmc_rbm_set_params(p.handle, _lib.ptr(flat), _lib.stream()),
                       "gfmc_rbm_set_params")
        sa = a.steps(n, 0.2, 0.1, 9, s0, evals=True)
        sb = b.steps(n, 0.2, 0.1, 9, s0, evals=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sa.cpu().numpy(), sb.cpu().numpy())
        same(a, b)
    np.testing.assert_array_equal(a.evals.cpu().numpy(), b.evals.cpu().numpy())
