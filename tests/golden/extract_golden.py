"""Write tests/golden/golden_spinhalf.npz from the reference's committed
fixture tests/core/reference_data/golden_cpu.npz (PyGFMC, Apache-2.0,
Copyright 2026 Robin Schaefer; jax 0.9.2, cpu, float64).

Kept (data only):
* ``<case>/config`` (64, 8) and ``<case>/e_loc`` (64,) of the SpinHalf runs
  without a guide and with the Jastrow guide, all three methods: the
  final walkers of a 300-step run on chain(8), J_x = J_y = -1, J_z = 1,
  h_x = 0.5, Jastrow beta = 0.3, h = 0.05 (test_golden_outputs.py:191-215);
  e_loc is the local energy of config after the last reconfiguration.
* ``<case>/raw/log_weights``, ``<case>/raw/energy`` (300,) and
  ``<case>/energy/<p>`` of every case that has them: the windowed energy
  estimator's inputs and outputs.
* ``spinhalf_*/raw/diag/{zz,sf}`` (300, [3,] 2), ``spinhalf_*/obs/{zz,sf}/<p>``
  and ``spinhalf_*/raw/additional_diag/{zz,sf}``: the forward-walking series
  of the diagonal operators zz = SpinHalf.create(edges, J_z=1) and sf =
  SzSzStructureFactor(positions 0..7, k = 2 pi m / 8, m = 0, 1, 2) and
  their estimates at p = 2, 4 (test_golden_outputs.py:191-211; the runs also
  measured off-diagonal operators, so the estimates use the common window
  trim i >= max(m) = 2).
* ``spinhalf_*/raw/off_diag/{spsm,sx,sfx}``, ``.../raw/off_diag_norm/...``
  (300, 2), ``.../raw/off_diag_steps`` (300,) and ``spinhalf_*/obs/{spsm,
  sx,sfx}/<p>``: the off-diagonal forward-walking rows (walker mean of the
  operator values after m shadow steps, the shadow's log weight product) of
  spsm = SpinHalf.create(edges, J_x=2, J_y=2), sx = SpinHalf.create(edges,
  h_x=-1) and sfx = 4 S^x_i S^x_j (1 + cos(pi/2)) on the bonds
  (test_golden_outputs.py:181-205), and their pure estimates at p = 2, 4.

Run in the build container only (the reference is not on the GPU box):
    python tests/golden/extract_golden.py
"""
import os

import numpy as np

SRC = "/root/reference/tests/core/reference_data/golden_cpu.npz"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_spinhalf.npz")


def main():
    g = np.load(SRC, allow_pickle=False)
    out = {}
    for guide in ("none", "jastrow"):
        for method in ("cont_time", "multi_step", "single_step"):
            case = f"spinhalf_{guide}_{method}"
            out[f"{case}/config"] = g[f"{case}/config"]
            out[f"{case}/e_loc"] = g[f"{case}/e_loc"]
    cases = sorted({k.split("/")[0] for k in g.files if "/raw/log_weights" in k})
    for case in cases:
        out[f"{case}/raw/log_weights"] = g[f"{case}/raw/log_weights"]
        out[f"{case}/raw/energy"] = g[f"{case}/raw/energy"]
        for k in g.files:
            if k.startswith(f"{case}/energy/"):
                out[k] = g[k]
        if case.startswith("spinhalf_"):
            for op in ("zz", "sf"):
                for k in g.files:
                    if (k in (f"{case}/raw/diag/{op}", f"{case}/raw/additional_diag/{op}")
                            or k.startswith(f"{case}/obs/{op}/")):
                        out[k] = g[k]
            for op in ("spsm", "sx", "sfx"):
                for k in g.files:
                    if (k in (f"{case}/raw/off_diag/{op}", f"{case}/raw/off_diag_norm/{op}")
                            or k.startswith(f"{case}/obs/{op}/")):
                        out[k] = g[k]
            if f"{case}/raw/off_diag_steps" in g.files:
                out[f"{case}/raw/off_diag_steps"] = g[f"{case}/raw/off_diag_steps"]
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays from {len(cases)} cases")


if __name__ == "__main__":
    main()
