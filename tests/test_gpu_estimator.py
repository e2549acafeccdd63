"""The p_array energy estimator streamed on the GPU (ring buffer of
cumulative log normalisation factors) against the reference's golden
estimates and the host formula."""
import os
import sys

import numpy as np
import pytest
import torch

from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain, prepare_walkers_sz, run
from pygfmc_amd.estimators import DeviceEnergyEstimator, estimate_energy

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden_spinhalf.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def stream(p_array, lw, E, chunks):
    est = DeviceEnergyEstimator(p_array, torch.device("cuda"))
    rows = torch.as_tensor(np.stack([lw, E], axis=1), device="cuda")
    edges = np.linspace(0, len(lw), chunks + 1).astype(int)
    for a, b in zip(edges[:-1], edges[1:]):
        est.update(rows[a:b])
    return est.result()


def test_streamed_estimator_reproduces_the_golden_energies():
    """The reference's own energy/{2,4} of its golden runs, from their raw
    series (the host formula matches these to the last bit)."""
    g = np.load(GOLDEN, allow_pickle=False)
    cases = sorted({k.split("/raw/")[0] for k in g.files if k.endswith("/raw/log_weights")})
    assert len(cases) >= 6
    for c in cases:
        out = stream([2, 4], g[f"{c}/raw/log_weights"], g[f"{c}/raw/energy"], chunks=7)
        for p in (2, 4):
            ref = float(g[f"{c}/energy/{p}"])
            assert out[p] == pytest.approx(ref, rel=1e-12, abs=1e-12), (c, p)


@pytest.mark.parametrize("chunks", [1, 13, 997])
def test_streamed_estimator_matches_host_formula(chunks):
    rng = np.random.default_rng(chunks)
    n = 5000
    lw = np.cumsum(rng.normal(scale=0.3, size=n)) * 0.01 + rng.normal(scale=2.0, size=n)
    E = rng.normal(loc=-40.0, scale=3.0, size=n)
    P = [1, 4, 16, 64, 128, 1000]
    out = stream(P, lw, E, chunks)
    ref = estimate_energy(P, lw, E)
    for p in P:
        assert out[p] == pytest.approx(ref[p], rel=1e-12), p


def test_streamed_estimator_flags_nonfinite_values():
    lw = np.zeros(50)
    E = np.ones(50)
    E[17] = np.nan
    est = DeviceEnergyEstimator([4], torch.device("cuda"))
    est.update(torch.as_tensor(np.stack([lw, E], axis=1), device="cuda"))
    with pytest.raises(ValueError):
        est.result()


def test_run_reports_the_streamed_estimate():
    say("run energy_device")
    L = chain(40)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(40, L.edges, beta=0.3, dtype=np.float64)
    w = prepare_walkers_sz(512, 40, 5)
    _, E, _, st, _ = run(w, 400, H, 0.2, G, [4, 16, 64], rng_key=8, warmup_steps=40,
                         verbosity=0, number_of_outputs=10)
    for p in (4, 16, 64):
        assert st["energy_device"][p] == pytest.approx(E[p], rel=1e-12)
