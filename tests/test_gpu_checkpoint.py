"""A run split into a run plus a continuation from the checkpoint is
bit-identical to the uninterrupted run."""
import numpy as np
import pytest

from pygfmc_amd import (JastrowSpinHalf, SpinHalf, chain, prepare_walkers_sz, run,
                        run_from_checkpoint, write_checkpoint)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,step,lam", [("cont_time", 0.2, None), ("multi_step", 4, 25.0)])
def test_resume_equals_uninterrupted_run(tmp_path, method, step, lam):
    L = chain(50)
    H = SpinHalf.create(L.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(50, L.edges, beta=0.3, dtype=np.float64)
    kw = dict(method=method, lam=lam, rng_key=11, verbosity=0, lanes=64, number_of_outputs=7)
    w = prepare_walkers_sz(512, 50, 2)
    _, _, _, st_all, raw_all = run(w, 300, H, step, G, [4, 16], warmup_steps=60, **kw)
    w1, _, _, st1, raw1 = run(w, 180, H, step, G, [4, 16], warmup_steps=60, **kw)
    f = str(tmp_path / "c.npz")
    write_checkpoint(f, st1, w1, H, G, raw1)
    _, _, _, st2, raw2 = run_from_checkpoint(f, H, G, n_steps=120, verbosity=0,
                                             number_of_outputs=5)
    np.testing.assert_array_equal(np.concatenate([raw1["log_weights"], raw2["log_weights"]]),
                                  raw_all["log_weights"])
    np.testing.assert_array_equal(np.concatenate([raw1["energy"], raw2["energy"]]),
                                  raw_all["energy"])
    np.testing.assert_array_equal(st2["parameters/rng_key_next"],
                                  st_all["parameters/rng_key_next"])
    with pytest.raises(ValueError):
        run_from_checkpoint(f, H, JastrowSpinHalf(50, L.edges, beta=0.2), n_steps=10)
