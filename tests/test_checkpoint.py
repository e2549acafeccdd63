"""Checkpoint file format (CPU): round trip, run appending, and the
parameters a continuation must share with the stored run."""
import numpy as np
import pytest
import torch

from pygfmc_amd import SpinHalf, chain, load_raw, read_checkpoint, write_checkpoint


class _Walkers:
    """The attributes write_checkpoint reads from a Walker."""

    def __init__(self, nw, N, seed):
        rng = np.random.default_rng(seed)
        self.n_sites, self.n_walkers = N, nw
        self.spins = torch.as_tensor(rng.integers(0, 2 ** 31, (nw, (N + 31) // 32),
                                                  dtype=np.int64).astype(np.int32))
        self.log_weight = np.zeros(nw)
        self.e_loc = rng.normal(size=nw)


def _stats(n_steps, pos, step=0.2):
    return {"parameters/step_size": step, "parameters/p_array": np.array([4, 16]),
            "parameters/method": "cont_time", "parameters/lam": float("nan"),
            "parameters/precision": "fp64", "parameters/lanes_per_walker": 64,
            "parameters/rng_key": np.array([7, pos], dtype=np.uint64),
            "parameters/rng_key_next": np.array([7, pos + n_steps], dtype=np.uint64),
            "avg_log_weight": 1.25, "parameters/n_steps": n_steps}


def test_checkpoint_round_trip_and_append(tmp_path):
    f = str(tmp_path / "chain.npz")
    H = SpinHalf.create(chain(40).edges, -1.0, -1.0, 1.0)
    w = _Walkers(16, 40, 0)
    rng = np.random.default_rng(1)
    raw1 = dict(log_weights=rng.normal(size=30), energy=rng.normal(size=30))
    raw2 = dict(log_weights=rng.normal(size=20), energy=rng.normal(size=20))
    assert write_checkpoint(f, _stats(30, 0), w, H, None, raw1) == 0
    w2 = _Walkers(16, 40, 2)
    assert write_checkpoint(f, _stats(20, 30), w2, H, None, raw2) == 1
    d = read_checkpoint(f)
    np.testing.assert_array_equal(d["current_walkers/spins"], w2.spins.numpy().view(np.uint32))
    np.testing.assert_array_equal(d["current_walkers/e_loc"], w2.e_loc)
    np.testing.assert_array_equal(d["run_1/rng_key_next"], [7, 50])
    raw = load_raw(f)
    np.testing.assert_array_equal(raw["log_weights"], np.concatenate([raw1["log_weights"],
                                                                      raw2["log_weights"]]))
    assert raw["p_array"] == [4, 16]
    with np.load(f, allow_pickle=False) as z:       # no pickled objects inside
        assert len(z.files) > 0


def test_checkpoint_rejects_a_different_calculation(tmp_path):
    f = str(tmp_path / "chain.npz")
    H = SpinHalf.create(chain(40).edges, -1.0, -1.0, 1.0)
    w = _Walkers(16, 40, 0)
    write_checkpoint(f, _stats(10, 0), w, H)
    with pytest.raises(ValueError, match="step_size"):
        write_checkpoint(f, _stats(10, 10, step=0.1), w, H)
    H2 = SpinHalf.create(chain(40).edges, -1.0, -1.0, 0.5)
    with pytest.raises(ValueError, match="hamiltonian"):
        write_checkpoint(f, _stats(10, 10), w, H2)
