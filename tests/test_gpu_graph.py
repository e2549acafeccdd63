"""Step loops captured into HIP graphs: every replay, with eager work and
synchronisations between replays, equals the eager launches bit for bit."""
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import (JastrowSpinHalf, SpinHalf, chain, prepare_walkers_random,
                        prepare_walkers_sz, run, run_batch, square, transverse_field_ising)
from pygfmc_amd.batch import Batch, BatchChain
from pygfmc_amd.engine import Chain, StepGraph
from pygfmc_amd.problem import Problem, precision_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def busy_eager_work():
    a = torch.randn(4096, device="cuda")
    b = (a * 2).sum()
    torch.cuda.synchronize()
    return float(b.item())


@pytest.mark.parametrize("dtype,n", [(np.float64, 4), (np.float32, 3)])
def test_graph_replays_equal_eager_steps(dtype, n):
    say("graph vs eager", dtype.__name__, n)
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 0.9, dtype=dtype)
    G = JastrowSpinHalf(16, L.edges, beta=0.2, dtype=dtype)
    p = Problem(H, G, precision_of(H))
    w = prepare_walkers_sz(700, 16, 3)
    a, b = Chain(p, w), Chain(p, w)
    g = StepGraph(b, n, 0.2, 1.3, 11)
    for r in range(4):
        sa = a.steps(n, 0.2, 1.3, 11, 100 + r * n).cpu().numpy()
        busy_eager_work()
        sb = g.launch(100 + r * n).cpu().numpy()
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a.spins.cpu().numpy(), b.spins.cpu().numpy())
        np.testing.assert_array_equal(a.e_loc.cpu().numpy(), b.e_loc.cpu().numpy())
        np.testing.assert_array_equal(a.log_w.cpu().numpy(), b.log_w.cpu().numpy())


def test_batch_graph_replays_equal_eager_steps():
    say("batch graph vs eager")
    Ls = [16, 30, 50]
    Hs = [transverse_field_ising(chain(L).edges, 1.0, 0.8 + 0.1 * i, dtype=np.float32)
          for i, L in enumerate(Ls)]
    Gs = [JastrowSpinHalf(L, chain(L).edges, beta=-0.2, dtype=np.float32) for L in Ls]
    nws = [300, 1100, 64]
    ws = [prepare_walkers_random(nw, L, i) for i, (L, nw) in enumerate(zip(Ls, nws))]
    bt = Batch(Hs, Gs, nws)
    a, b = BatchChain(bt, ws), BatchChain(bt, ws)
    alw = torch.tensor([0.1, 0.2, 0.3], dtype=torch.float64, device="cuda")
    g = StepGraph(b, 5, 0.2, alw, 21)
    for r in range(3):
        sa = a.steps(5, 0.2, alw, 21, r * 5).cpu().numpy()
        busy_eager_work()
        sb = g.launch(r * 5).cpu().numpy()
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a.spins.cpu().numpy(), b.spins.cpu().numpy())
        np.testing.assert_array_equal(a.e_loc.cpu().numpy(), b.e_loc.cpu().numpy())


def test_run_with_and_without_graph_are_identical():
    say("run graph on/off")
    L = chain(40)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(40, L.edges, beta=0.3, dtype=np.float64)
    out = []
    for use_graph in (False, True):
        w = prepare_walkers_sz(512, 40, 5)
        _, E, _, st, raw = run(w, 300, H, 0.2, G, [4, 16], rng_key=8, warmup_steps=40,
                               verbosity=0, number_of_outputs=10, use_graph=use_graph)
        assert st["hip_graph"] == use_graph
        out.append(raw)
    np.testing.assert_array_equal(out[0]["log_weights"], out[1]["log_weights"])
    np.testing.assert_array_equal(out[0]["energy"], out[1]["energy"])
