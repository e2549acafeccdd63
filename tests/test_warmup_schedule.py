"""The warm-up ramp's stage sizes and counts (no GPU).

The rule of the reference (codes/pygfmc/engine/driver.py:567-584): the
warm-up steps split over four stages by divmod(warmup_steps, 4), the first
`extra` stages one step longer; step sizes beta/8, beta/4, beta/2, beta in
continuous time, max(1, n_bra // 8), max(1, n_bra // 4), max(1, n_bra // 2),
n_bra sub-steps for the discrete methods; empty stages dropped.  The
expected tables below are written out by hand from that rule.
"""
import pytest

from pygfmc_amd.engine import warmup_schedule


@pytest.mark.parametrize("n,counts", [
    (1000, [250, 250, 250, 250]),
    (1003, [251, 251, 251, 250]),
    (1001, [251, 250, 250, 250]),
    (14, [4, 4, 3, 3]),
    (5, [2, 1, 1, 1]),
])
def test_cont_time_stages(n, counts):
    got = warmup_schedule(0.2, n)
    assert [c for _, c in got] == counts
    assert [s for s, _ in got] == [0.025, 0.05, 0.1, 0.2]


def test_short_ramps_drop_empty_stages():
    assert warmup_schedule(0.2, 3) == [(0.025, 1), (0.05, 1), (0.1, 1)]
    assert warmup_schedule(0.4, 1) == [(0.05, 1)]
    assert warmup_schedule(0.2, 0) == []
    assert warmup_schedule(0.2, -5) == []


@pytest.mark.parametrize("n_bra,sizes", [
    (6, [1.0, 1.0, 3.0, 6.0]),
    (20, [2.0, 5.0, 10.0, 20.0]),
    (1, [1.0, 1.0, 1.0, 1.0]),
    (9, [1.0, 2.0, 4.0, 9.0]),
])
@pytest.mark.parametrize("method", ["single_step", "multi_step"])
def test_discrete_stage_sizes(n_bra, sizes, method):
    got = warmup_schedule(n_bra, 10, method)
    assert [s for s, _ in got] == sizes
    assert [c for _, c in got] == [3, 3, 2, 2]
