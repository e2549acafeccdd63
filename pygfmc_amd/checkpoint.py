"""Checkpoint / resume of a GFMC chain and its raw series.

The reference stores runs in HDF5 (write_to_HDF5 / run_from_HDF5 /
load_raw_from_HDF5, engine/io.py:370-1473).  h5py is not part of this
engine's environment, so the same information goes to one NumPy ``.npz``
archive (no pickled objects; ``numpy.load(allow_pickle=False)`` reads it):

* ``parameters/*``: what must match for a continuation to be the same
  calculation (step size, p_array, method, lam, precision, lanes per
  walker, sites, walkers, and SHA-1 fingerprints of the Hamiltonian and the
  guide);
* ``current_walkers/{spins,log_weight,e_loc}``: the ensemble after the
  newest run (bit-packed spins);
* ``run_<n>/{log_weights,energy,rng_key,rng_key_next,avg_log_weight}`` per
  appended run, and with diagonal operators ``run_<n>/diag/<name>`` and the
  lead-in ``run_<n>/additional_diag/<name>`` (the diagonal rows of the
  max(m) steps after the run, traced with its ancestry); with off-diagonal
  operators ``run_<n>/off_diag/<name>``, ``run_<n>/off_diag_norm/<name>``
  and ``run_<n>/off_diag_steps`` (the run's own 0-based steps; ``load_raw``
  offsets them by the steps of the runs before, io.py:505-515).

``run_from_checkpoint`` continues the stored chain without a warm-up, with
the stored shift, walkers and Philox position: the series of a run split
into a run plus a continuation is bit-identical to the uninterrupted run.
The continuation's ancestry ring starts empty, so its first max(m)
diagonal rows in ``raw['diag']`` are replaced by the stored lead-in
(io.py:1428-1470), which makes those series identical to the uninterrupted
run's as well (the returned ``obs`` are the call's own estimates, as in the
reference; re-estimate the spliced series with ``evaluate_observables``).
"""
from __future__ import annotations

import hashlib
import os

import numpy as np
import torch

from .walkers import Walker

__all__ = ["write_checkpoint", "read_checkpoint", "load_raw", "run_from_checkpoint"]

MUST_MATCH = ("step_size", "p_array", "method", "lam", "precision", "n_sites", "n_walkers",
              "hamiltonian_sha1", "guide_sha1", "diag_operators", "diag_operator_hashes",
              "off_diag_operators", "off_diag_operator_hashes", "off_diag_stride",
              "off_diag_equal_windows")


def _joined(stats, key):
    return np.asarray(",".join(x.decode() if isinstance(x, bytes) else str(x)
                               for x in stats.get(key, [])))


def _sha1(obj):
    if obj is None:
        return "none"
    return hashlib.sha1(obj.fingerprint()).hexdigest()


def _read(filename):
    if not os.path.exists(filename):
        return {}
    with np.load(filename, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _runs(data):
    return sorted({int(k.split("/")[0][4:]) for k in data if k.startswith("run_")})


def write_checkpoint(filename, stats, walkers, hamiltonian, guide=None, raw=None):
    """Append the run described by ``stats`` / ``raw`` (as returned by
    ``run``) and replace the stored ensemble by ``walkers``.  The first write
    records the parameters; later writes must match them."""
    params = {
        "step_size": np.float64(stats["parameters/step_size"]),
        "p_array": np.asarray(stats["parameters/p_array"], dtype=np.int64),
        "method": np.asarray(stats["parameters/method"]),
        "lam": np.float64(stats["parameters/lam"]),
        "precision": np.asarray(stats["parameters/precision"]),
        "lanes_per_walker": np.int64(stats["parameters/lanes_per_walker"]),
        "n_sites": np.int64(walkers.n_sites),
        "n_walkers": np.int64(walkers.n_walkers),
        "hamiltonian_sha1": np.asarray(_sha1(hamiltonian)),
        "guide_sha1": np.asarray(_sha1(guide)),
        "diag_operators": _joined(stats, "parameters/diag_operators"),
        "diag_operator_hashes": _joined(stats, "parameters/diag_operator_hashes"),
        "off_diag_operators": _joined(stats, "parameters/off_diag_operators"),
        "off_diag_operator_hashes": _joined(stats, "parameters/off_diag_operator_hashes"),
        "off_diag_stride": np.int64(stats.get("parameters/off_diag_stride", 1)),
        "off_diag_equal_windows": np.int64(stats.get("parameters/off_diag_equal_windows", 0)),
    }
    data = _read(filename)
    if data:
        for k in MUST_MATCH:
            old, new = data.get(f"parameters/{k}", params[k]), params[k]
            same = (np.array_equal(old, new) or
                    (np.asarray(old).dtype.kind == "f" and np.isnan(old) and np.isnan(new)))
            if not same:
                raise ValueError(f"{filename}: parameter {k!r} differs from the stored run "
                                 f"({old!r} vs {new!r}); a continuation must be the same "
                                 f"calculation")
    for k, v in params.items():
        data[f"parameters/{k}"] = v
    n = (max(_runs(data)) + 1) if _runs(data) else 0
    if raw is not None:
        data[f"run_{n}/log_weights"] = np.asarray(raw["log_weights"], dtype=np.float64)
        data[f"run_{n}/energy"] = np.asarray(raw["energy"], dtype=np.float64)
        for group in ("diag", "additional_diag", "off_diag", "off_diag_norm"):
            for name, arr in (raw.get(group) or {}).items():
                if "/" in name or "," in name:
                    raise ValueError(f"operator name {name!r} contains '/' or ','")
                data[f"run_{n}/{group}/{name}"] = np.asarray(arr, dtype=np.float64)
        if raw.get("off_diag_steps") is not None:
            data[f"run_{n}/off_diag_steps"] = np.asarray(raw["off_diag_steps"], dtype=np.int64)
    data[f"run_{n}/rng_key"] = np.asarray(stats["parameters/rng_key"], dtype=np.uint64)
    data[f"run_{n}/rng_key_next"] = np.asarray(stats["parameters/rng_key_next"],
                                               dtype=np.uint64)
    data[f"run_{n}/avg_log_weight"] = np.float64(stats["avg_log_weight"])
    data[f"run_{n}/n_steps"] = np.int64(stats["parameters/n_steps"])
    data["current_walkers/spins"] = walkers.spins.cpu().numpy().view(np.uint32)
    data["current_walkers/log_weight"] = walkers.log_weight
    data["current_walkers/e_loc"] = walkers.e_loc
    tmp = filename + ".tmp.npz"
    np.savez(tmp, **data)
    os.replace(tmp, filename)
    return n


def read_checkpoint(filename):
    """The stored arrays as a flat dict (``parameters/*``,
    ``current_walkers/*``, ``run_<n>/*``)."""
    data = _read(filename)
    if not data:
        raise FileNotFoundError(filename)
    return data


def load_raw(filename):
    """The raw series of every stored run, concatenated in run order
    (load_raw_from_HDF5, engine/io.py:742)."""
    data = read_checkpoint(filename)
    runs = [n for n in _runs(data) if f"run_{n}/log_weights" in data]
    if not runs:
        raise ValueError(f"{filename} holds no raw series")
    lw = np.concatenate([data[f"run_{n}/log_weights"] for n in runs])
    E = np.concatenate([data[f"run_{n}/energy"] for n in runs])
    p = [int(v) for v in data["parameters/p_array"]]
    names = _diag_names(data)
    diag = ({name: np.concatenate([data[f"run_{n}/diag/{name}"] for n in runs])
             for name in names} if names else None)
    onames = _diag_names(data, "off_diag_operators")
    off = off_norm = off_steps = None
    if onames:
        off = {name: np.concatenate([data[f"run_{n}/off_diag/{name}"] for n in runs])
               for name in onames}
        off_norm = {name: np.concatenate([data[f"run_{n}/off_diag_norm/{name}"] for n in runs])
                    for name in onames}
        starts = np.cumsum([0] + [len(data[f"run_{n}/energy"]) for n in runs])
        off_steps = np.concatenate([data[f"run_{n}/off_diag_steps"] + starts[i]
                                    for i, n in enumerate(runs)]).astype(np.int64)
    return dict(log_weights=lw, energy=E, p_array=p, m_array=[q // 2 for q in p], diag=diag,
                diag_operators=names or None, off_diag=off, off_diag_norm=off_norm,
                off_diag_steps=off_steps, off_diag_operators=onames or None)


def _diag_names(data, key="diag_operators"):
    s = str(data.get(f"parameters/{key}", ""))
    return [x for x in s.split(",") if x]


def run_from_checkpoint(filename, hamiltonian, guide=None, n_steps=None, **run_kwargs):
    """Continue the stored chain (run_from_HDF5, engine/io.py:1303-1473):
    stored walkers, shift and Philox position, no warm-up; step size,
    p_array, method, lam and lanes per walker come from the file.  The
    Hamiltonian and guide are checked against the stored fingerprints."""
    from .engine import run_gfmc
    data = read_checkpoint(filename)
    if str(data["parameters/hamiltonian_sha1"]) != _sha1(hamiltonian):
        raise ValueError(f"{filename}: the Hamiltonian differs from the stored one")
    if str(data["parameters/guide_sha1"]) != _sha1(guide):
        raise ValueError(f"{filename}: the guide differs from the stored one")
    for k in ("step_size", "p_array", "method", "lam", "lanes", "precision", "rng_key",
              "avg_log_weight", "warmup_steps"):
        if k in run_kwargs:
            raise ValueError(f"{k!r} comes from the checkpoint and cannot be overridden")
    last = max(_runs(data))
    if n_steps is None:
        n_steps = int(data[f"run_{last}/n_steps"])
    N = int(data["parameters/n_sites"])
    spins = torch.as_tensor(data["current_walkers/spins"].view(np.int32), device="cuda")
    w = Walker.from_device(spins,
                           torch.as_tensor(data["current_walkers/log_weight"], device="cuda"),
                           torch.as_tensor(data["current_walkers/e_loc"], device="cuda"), N)
    method = str(data["parameters/method"])
    lam = float(data["parameters/lam"])
    key = tuple(int(v) for v in data[f"run_{last}/rng_key_next"])
    names = _diag_names(data)
    given = list(run_kwargs.get("diag_operators") or {})
    if given != names:
        raise ValueError(f"{filename}: the stored run measured diagonal operators {names}, "
                         f"this call {given}; a continuation must measure the same ones")
    onames = _diag_names(data, "off_diag_operators")
    ogiven = list(run_kwargs.get("off_diag_operators") or {})
    if ogiven != onames:
        raise ValueError(f"{filename}: the stored run measured off-diagonal operators "
                         f"{onames}, this call {ogiven}; a continuation must measure the same "
                         f"ones")
    for k in ("off_diag_stride", "return_raw"):
        if k in run_kwargs:
            raise ValueError(f"{k!r} comes from the checkpoint and cannot be overridden")
    if "parameters/off_diag_stride" in data:
        run_kwargs["off_diag_stride"] = int(data["parameters/off_diag_stride"])
    if int(data.get("parameters/off_diag_equal_windows", 0)):
        run_kwargs["return_raw"] = False
    out = run_gfmc(w, n_steps, hamiltonian, float(data["parameters/step_size"]), guide,
                   [int(v) for v in data["parameters/p_array"]], method=method,
                   lam=None if np.isnan(lam) else lam, rng_key=key,
                   avg_log_weight=float(data[f"run_{last}/avg_log_weight"]), warmup_steps=0,
                   lanes=int(data["parameters/lanes_per_walker"]),
                   precision=str(data["parameters/precision"]), **run_kwargs)
    if names and len(out) > 4 and out[4] is not None and out[4].get("diag") is not None:
        raw = out[4]
        spliced = {}
        for name in names:
            arr = np.array(raw["diag"][name], copy=True)
            lead = data.get(f"run_{last}/additional_diag/{name}")
            if lead is not None:
                if lead.shape[1:] != arr.shape[1:]:
                    raise ValueError(f"{filename}: the stored lead-in of {name!r} has rows "
                                     f"{lead.shape[1:]}, this run {arr.shape[1:]}")
                k = min(lead.shape[0], arr.shape[0])
                arr[:k] = lead[:k]
            spliced[name] = arr
        raw["diag"] = spliced
    return out
