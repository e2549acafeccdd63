"""The warm-up ramp and its shift updates against oracle chains (FP64).

Each stage of engine._warmup runs `count` branching steps of its step size
with the shift of the stage before, then measures the new shift
alw + (LSE(log G of the stage's steps) - log count) / step
(driver.py:733-736, estimators.py:722-751 for the starting value).  The
oracle replays the same stages from the same walkers with the same Philox
positions (warm-up tags) by its own chain; the stage sizes and counts are
the reference rule's, recomputed here.
"""
import numpy as np
import pytest
import torch

from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain, square, _lib
from pygfmc_amd.engine import Chain, _warmup, set_method
from pygfmc_amd.problem import Problem, lower
from pygfmc_amd.walkers import PhiloxKey, Walker

pytestmark = pytest.mark.gpu


def stages_of(step_f, n, method):
    base, extra = divmod(n, 4)
    counts = [base + (i < extra) for i in range(4)]
    if method == "cont_time":
        sizes = [step_f / 8, step_f / 4, step_f / 2, step_f]
    else:
        k = int(step_f)
        sizes = [max(1, k // 8), max(1, k // 4), max(1, k // 2), k]
    return [(float(s), c) for s, c in zip(sizes, counts) if c > 0]


CASES = [
    ("cont_time", 0.2, None, 14),
    ("cont_time", 0.3, None, 7),
    ("multi_step", 8, 6.0, 13),
    ("single_step", 4, 6.0, 9),
]


@pytest.mark.parametrize("method,step_f,lam,n_warm", CASES,
                         ids=[f"{c[0]}-{c[3]}" for c in CASES])
def test_warmup_stage_shifts_match_oracle_chains(orc, method, step_f, lam, n_warm):
    L = chain(12)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(12, L.edges, beta=0.3, dtype=np.float64)
    nw, seed, pos0 = 128, 2024, 500
    spins = orc.init_sz(12, nw, 6, 31)
    p = Problem(H, G, _lib.GFMC_FP64)
    m = {"cont_time": 0, "single_step": 1, "multi_step": 2}[method]
    set_method(p, method, lam)
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ch = Chain(p, Walker.from_device(torch.as_tensor(spins.view(np.int32), device="cuda"),
                                     z, z.clone(), 12))
    trace = []
    alw_gpu = _warmup(ch, PhiloxKey(seed, pos0), step_f, None, n_warm, 0, method, lam,
                      trace=trace)
    torch.cuda.synchronize()
    # the oracle: starting shift from its rates, then the stages
    op = orc.Problem(lower(H, G))
    _, e0, _ = op.rates(spins)
    alw = float(np.mean(-e0)) if lam is None else float(np.mean(np.log(lam - e0)))
    s = spins
    counter = pos0
    want = stages_of(step_f, n_warm, method)
    assert [(a, b) for a, b, _ in trace] == want
    for (step_, n_w), (_, _, got) in zip(want, trace):
        s, _, ser = orc.run_chain(op, s, n_w, step_, alw, seed, counter, warmup=True,
                                  method=m, lam=0.0 if lam is None else lam)
        counter += n_w
        lg = ser[:, 0]
        mx = lg.max()
        alw = alw + (mx + np.log(np.exp(lg - mx).sum()) - np.log(n_w)) / step_
        assert got == pytest.approx(alw, rel=1e-12, abs=1e-12)
    assert alw_gpu == pytest.approx(alw, rel=1e-12, abs=1e-12)
    np.testing.assert_array_equal(ch.spins.cpu().numpy().view(np.uint32), s)


def test_pinned_shift_is_kept_while_the_ramp_runs(orc):
    """A pinned avg_log_weight is used by every stage and returned unchanged
    (driver.py:765-777); the walkers still equilibrate."""
    L = square(4, 4)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(16, L.edges, beta=0.2, dtype=np.float64)
    nw = 64
    spins = orc.init_sz(16, nw, 8, 5)
    p = Problem(H, G, _lib.GFMC_FP64)
    z = torch.zeros(nw, dtype=torch.float64, device="cuda")
    ch = Chain(p, Walker.from_device(torch.as_tensor(spins.view(np.int32), device="cuda"),
                                     z, z.clone(), 16))
    trace = []
    out = _warmup(ch, PhiloxKey(9, 0), 0.2, 11.5, 8, 0, trace=trace)
    assert out == 11.5
    op = orc.Problem(lower(H, G))
    s = spins
    counter = 0
    for step_, n_w, got in trace:
        s, _, ser = orc.run_chain(op, s, n_w, step_, 11.5, 9, counter, warmup=True)
        counter += n_w
        lg = ser[:, 0]
        mx = lg.max()
        assert got == pytest.approx(11.5 + (mx + np.log(np.exp(lg - mx).sum()) - np.log(n_w))
                                    / step_, rel=1e-12, abs=1e-12)
    np.testing.assert_array_equal(ch.spins.cpu().numpy().view(np.uint32), s)
