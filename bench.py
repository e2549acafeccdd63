"""Throughput of the continuous-time GFMC branching step on one MI355X.

Workload (default): XXZ model on the 32 x 32 periodic square lattice
(N = 1024 sites, 2048 nearest-neighbour edges), J_x = J_y = -1 (sign-free
frame), J_z = Delta = 1.0, Jastrow guide beta = 0.3 on the lattice edges,
65 536 walkers started with Sz = 0, imaginary-time step 0.2, FP32 rates and
ratios with FP64 weights, prefix sums and accumulators.  A step = one
propagation of every walker + systematic reconfiguration + gather.
Synthetic input: random Sz = 0 walkers from a fixed seed, equilibrated by
the untimed warm-up steps.

Prints ONE JSON line (see the keys below).  `roofline` is measured live for
the propagation kernel with HIP events on the launch stream; `cpu_baseline`
times the FP64 C oracle (the reference algorithm, full re-enumeration of the
slots at every hop) on the host cores on a bounded sample of the same
walkers.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps after the workload's own GFMC warm-up")
    ap.add_argument("--settle", type=int, default=500,
                    help="untimed production steps after the GFMC warm-up (the population "
                         "keeps relaxing for a few hundred steps after the ramp)")
    ap.add_argument("--gfmc-warmup", type=int, default=1000,
                    help="the workload's GFMC warm-up ramp (driver.py DEFAULT_WARMUP_STEPS), "
                         "always run before any timing")
    ap.add_argument("--delta", type=float, default=1.0)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", default="auto", help="lanes per walker: auto, 64, 32, 16")
    ap.add_argument("--kernel-steps", type=int, default=10)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the chain (L = 100) and batched Ising-chain sections")
    ap.add_argument("--no-race", action="store_true",
                    help="skip the second metric (GPU-seconds to a fixed energy error bar, "
                         "learned RBM guide vs Jastrow, 10x10 Heisenberg)")
    ap.add_argument("--no-mlp", action="store_true",
                    help="skip the large learned-guide section (kagome 12x12, two-layer MLP, "
                         "262 144 walkers)")
    ap.add_argument("--only-mlp", action="store_true",
                    help="run the kagome MLP section alone (for profiling its kernels)")
    ap.add_argument("--only-race", action="store_true",
                    help="run the learned-guide section alone (10x10 Heisenberg, RBM vs "
                         "Jastrow; for profiling its kernels)")
    ap.add_argument("--mlp-walkers", type=int, default=262144)
    ap.add_argument("--mlp-steps", type=int, default=3)
    ap.add_argument("--race-train-iters", type=int, default=1500)
    ap.add_argument("--race-steps", type=int, default=20000,
                    help="production steps of each run of the learned-guide section; the "
                         "jackknife takes blocks of 200 steps (longer than the p = 128 "
                         "estimator's window): 100 blocks at 20000 steps, a ~14 %% relative "
                         "error of each run's sigma^2")
    ap.add_argument("--summarize-profile", default=None,
                    help="print a text summary of a rocprofv3 database and exit")
    ap.add_argument("--summarize-kernel", action="append", default=None,
                    help="PREFIX@N[,SKIP]: summarise the N dispatches (0: all) before the "
                         "last SKIP of the kernels whose name starts with PREFIX (repeatable)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before the timed steps, copy a tensor of this many MiB once "
                         "(a plain copy kernel of known size, to scale HBM counters)")
    return ap.parse_args()


def summarize_profile(path, kernels):
    """Text summary of a rocprofv3 CSV output directory (--output-format
    csv) of this script: the kernel statistics table, then per kernel of
    ``kernels`` [(name prefix, n_last, skip)] the average duration (kernel
    trace) or counter value (--pmc) of the n_last dispatches before its last
    ``skip`` ones (every dispatch when n_last = 0), and their sum."""
    import csv
    import glob

    def rows(suffix):
        out = []
        for f in sorted(glob.glob(os.path.join(path, "**", "*" + suffix), recursive=True)):
            with open(f, newline="") as fh:
                out += list(csv.DictReader(fh))
        return out

    def short(name):
        name = name.split("(")[0]
        return name[5:] if name.startswith("void ") else name

    lines = [f"profile: {os.path.basename(os.path.normpath(path))}"]
    stats = rows("kernel_stats.csv")
    if stats:
        stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
        lines.append("| kernel | calls | avg us | % |")
        lines.append("|---|---|---|---|")
        for r in stats[:16]:
            lines.append(f"| {short(r['Name'])[:90]} | {r['Calls']} | "
                         f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    trace = rows("kernel_trace.csv")
    pmc = rows("counter_collection.csv")
    for prefix, n_last, skip in kernels:
        tr = sorted((r for r in trace if short(r["Kernel_Name"]).startswith(prefix)),
                    key=lambda r: int(r["Start_Timestamp"]))
        if tr:
            d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in tr]
            d = d[:len(d) - skip]
            last = d[-n_last:] if n_last else d
            lines.append("")
            lines.append(f"{prefix}...: {len(d) + skip} launches; {len(last)} before the last "
                         f"{skip}: avg "
                         f"{np.mean(last):.4f} ms, total {np.sum(last):.3f} ms")
        by = {}
        for r in pmc:
            if short(r["Kernel_Name"]).startswith(prefix):
                by.setdefault(r["Counter_Name"], []).append(
                    (int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        for cn, v in sorted(by.items()):
            v.sort()
            v = v[:len(v) - skip]
            last = [x[1] for x in (v[-n_last:] if n_last else v)]
            lines.append(f"{prefix}... {cn}: {len(v) + skip} dispatches; {len(last)} before the "
                         f"last {skip}: avg "
                         f"{np.mean(last):.1f}, sum {np.sum(last):.1f}")
    # counters of the copy kernel (calibration of the HBM byte counters)
    cp = [r for r in pmc if "copyBuffer" in r["Kernel_Name"]]
    if cp:
        top = max(cp, key=lambda r: float(r["Counter_Value"]))
        lines.append(f"largest device copy (grid {top['Grid_Size']}): {top['Counter_Name']} = "
                     f"{float(top['Counter_Value']):.1f}")
    return "\n".join(lines)


def calib_copy(mib):
    """One device copy of ``mib`` MiB (a plain copy kernel of known size, to
    scale the HBM byte counters of a --pmc run)."""
    import torch
    src = torch.ones(mib * (1 << 18), dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()


def profiled_traffic(stem, key="traffic_bytes_per_launch"):
    """The HBM bytes recorded in profiles/<stem>.md (a line ``<key>: <int>``),
    or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", stem + ".md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return {"traffic_bytes_per_launch": float(line.split(":")[1]),
                        "source": "profiles/" + stem + ".md"}
    return None


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.summarize_profile:
        ks = []
        for k in args.summarize_kernel or ["k_propagate<float, 2, 64@60"]:
            pre, _, n = k.rpartition("@") if "@" in k else (k, "", "0")
            n_last, _, skip = n.partition(",")
            ks.append((pre, int(n_last), int(skip or 0)))
        print(summarize_profile(args.summarize_profile, ks))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; runs on more than one GPU "
                                   "are not implemented (walkers would shard by contiguous "
                                   "ranges with a scalar scan of the weight sums)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if args.only_mlp:
        if args.calib_copy_mib > 0:
            calib_copy(args.calib_copy_mib)
        print(json.dumps({"kagome12_mlp": kagome_mlp_section(args)}), flush=True)
        return
    if args.only_race:
        if args.calib_copy_mib > 0:
            calib_copy(args.calib_copy_mib)
        print(json.dumps({"sigma_race": sigma_race(args)}), flush=True)
        return
    from pygfmc_amd import JastrowSpinHalf, SpinHalf, prepare_walkers_sz, square
    from pygfmc_amd import _lib
    from pygfmc_amd.engine import Chain
    from pygfmc_amd.problem import get_problem

    L, nw, beta = args.L, args.walkers, 0.2
    lat = square(L, L)
    N, E = lat.N, len(lat.edges)
    H = SpinHalf.create(lat.edges, J_x=-1.0, J_y=-1.0, J_z=args.delta,
                        dtype=np.float32 if args.precision == "fp32" else np.float64)
    G = JastrowSpinHalf(N, lat.edges, beta=0.3)
    p = get_problem(H, G)
    w = prepare_walkers_sz(nw, N, 12345)
    chain = Chain(p, w)
    seed = 777
    series = torch.empty((max(args.steps, args.warmup, 1), 2), dtype=torch.float64,
                         device=chain.spins.device)

    def one_step(t, row, ev_pair=None):
        chain.step(beta, alw, seed, t, series[row], events=ev_pair)

    # the workload's own GFMC warm-up (untimed, whatever --warmup says): the
    # four-stage step ramp of run_gfmc (250 steps each at step 0.025, 0.05,
    # 0.1, 0.2), which also measures the shift.
    from pygfmc_amd.engine import _warmup
    from pygfmc_amd.walkers import PhiloxKey
    p.set_lanes(64)
    alw = _warmup(chain, PhiloxKey(seed, 0), beta, None, args.gfmc_warmup, 0)
    pos0 = 1 << 20             # production positions, past the warm-up's
    # untimed production steps (shift held): after the ramp the population
    # keeps relaxing (the evaluations per walker-step can sit near 28 for a
    # hundred steps or more before settling near 17), then the harness's
    # untimed steps
    for t in range(args.settle + args.warmup):
        one_step(pos0 + t, t % series.shape[0])
    pos0 += args.settle + args.warmup
    torch.cuda.synchronize()
    # lanes per walker of the propagation kernel: the autotuner, or pinned
    if args.lanes == "auto":
        from pygfmc_amd.autotune import tune_problem
        shape, tune_times = tune_problem(p, chain.spins, beta, alw)
    else:
        p.set_lanes(int(args.lanes))
        shape, tune_times = (int(args.lanes), 1, 0), None
    lanes = shape[0]

    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib)

    # kernel pass: per-step launches, HIP events around the propagation
    # kernel on its launch stream (the roofline's kernel time and work)
    n_k = max(1, args.kernel_steps)
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(n_k)]
    chain.evals.zero_()
    for k in range(n_k):
        one_step(pos0 + k, k, events[k])
    torch.cuda.synchronize()
    prop_ms = [a.elapsed_time(b) for a, b in events]
    prop_avg_s = float(np.mean(prop_ms)) / 1e3
    evals_kernel = float(chain.evals.to(torch.float64).sum().item())

    # timed region: the production loop as replays of one captured HIP
    # graph of `chunk` steps, inputs resident in HBM
    from pygfmc_amd.engine import StepGraph
    chunk = max(1, min(args.graph_steps, args.steps))
    n_rep = max(1, args.steps // chunk)
    steps_timed = n_rep * chunk
    graph = StepGraph(chain, chunk, beta, alw, seed, evals=True)
    pos = pos0 + n_k
    graph.launch(pos)          # untimed: first replay
    pos += chunk
    chain.evals.zero_()
    out_series = torch.empty((steps_timed, 2), dtype=torch.float64, device=chain.spins.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(n_rep):
        graph.launch(pos + r * chunk, out=out_series[r * chunk:(r + 1) * chunk])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lg = out_series[:, 0].cpu().numpy()
    if not np.all(np.isfinite(lg)):
        raise SystemExit("population extinct during the bench")
    evals_total = float(chain.evals.to(torch.float64).sum().item())
    evals_per_ws = evals_total / (nw * steps_timed)
    ms_per_step = elapsed / steps_timed * 1e3
    value = nw * steps_timed / elapsed

    # algorithmic FP32 work of one launch (reference algorithm: every slot of
    # every rate evaluation): n_edges * (4 (d - 1) + 6) per evaluation
    d = G.max_deg
    flop_per_eval = E * (4 * (d - 1) + 6)
    flops_launch = evals_kernel / n_k * flop_per_eval
    achieved_tf = flops_launch / prop_avg_s / 1e12
    peak_tf = 157.3

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(H, G, chain, N, beta, alw, seed, args.cpu_seconds)

    # HBM bytes per launch: two rocprofv3 --pmc passes (FETCH_SIZE,
    # WRITE_SIZE, each scaled by what it reads for a device copy of known
    # size) of this same command, kept in the kernel's profile summary; not
    # measurable inside this run, so null for another workload
    workload = f"xxz_square{L}x{L}_delta{args.delta:g}_jastrow0.3_w{nw}_step0.2_{args.precision}"
    prof = profiled_traffic(f"k_propagate_{workload}_lanes{lanes}")
    traffic = prof["traffic_bytes_per_launch"] if prof else None

    out = {
        "metric": "walker-steps/s; GPU-sec to fixed energy σ, learned BF16 guide vs plain",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": steps_timed,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else "fp64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "lattice": f"square {L}x{L} periodic", "n_sites": N, "n_edges": E,
                   "n_walkers": nw, "delta": args.delta, "step_size": beta,
                   "guide": "Jastrow beta=0.3", "global_batch": nw, "seq_len": N,
                   "parallelism": "none (1 GPU)", "hip_graph_steps": chunk,
                   "lanes_per_walker": lanes, "waves_per_workgroup": shape[1],
                   "xcd_contiguous_walkers": bool(shape[2]), "autotune_seconds": tune_times,
                   "evals_per_walker_step": evals_per_ws},
        "roofline": {"bound": "valu",
                     "compute_unit": "FP32 vector ALU (no GEMM on this path; MI355X's FP32 "
                                     "vector and matrix peaks are both 157.3 TFLOP/s)",
                     "achieved": achieved_tf, "peak": peak_tf, "unit": "TFLOP/s",
                     "frac": achieved_tf / peak_tf,
                     "work": "reference-equivalent FLOPs: every slot of every rate "
                             "evaluation, n_edges (4 (d - 1) + 6) per evaluation; the kernel "
                             "itself re-evaluates only the slots a move changes",
                     "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE + "
                                     "WRITE_SIZE, copy-calibrated)",
                     "traffic_source": prof["source"] if prof else None,
                     "hbm_GBps": traffic / prop_avg_s / 1e9 if traffic else None,
                     "hbm_frac": traffic / prop_avg_s / 8e12 if traffic else None,
                     "kernel": f"k_propagate<float, local-field records, {lanes} lanes>",
                     "kernel_avg_ms": prop_avg_s * 1e3, "kernel_steps": n_k,
                     "evals_per_walker_step": evals_kernel / (nw * n_k),
                     "flop_per_eval": flop_per_eval},
        "cpu_baseline": cpu,
    }
    progress(f"headline done: {value:.4g} walker-steps/s")
    if not args.no_configs:
        out["xxz_square32_deltas"] = xxz_deltas_section(args)
        progress("xxz deltas done")
        out["heisenberg_chain100"] = chain100_section()
        progress("chain100 done")
        out["batched_tfim_chains"] = tfim_batch_section()
        progress("tfim batch done")
        out["published_throughput_config"] = fig3_section()
        progress("published configuration done")
    if not args.no_mlp:
        out["kagome12_mlp"] = kagome_mlp_section(args)
        progress("kagome mlp done")
    if not args.no_race:
        out["sigma_race"] = sigma_race(args)
        progress("sigma race done")
    print(json.dumps(out), flush=True)


def xxz_deltas_section(args):
    """The 32 x 32 XXZ configuration at its three anisotropies (Delta = 0.5,
    1, 2): the same protocol as the headline (the 1000-step warm-up ramp,
    500 untimed steps, then 20 steps as replays of one captured 10-step
    graph) with the autotuned lanes per walker; walker-steps/s and the rate
    evaluations per walker-step of the timed steps."""
    import torch

    from pygfmc_amd import JastrowSpinHalf, SpinHalf, prepare_walkers_sz, square
    from pygfmc_amd.autotune import tune_problem
    from pygfmc_amd.engine import Chain, StepGraph, _warmup
    from pygfmc_amd.problem import get_problem
    from pygfmc_amd.walkers import PhiloxKey
    lat = square(32, 32)
    res = {}
    for delta in (0.5, 1.0, 2.0):
        H = SpinHalf.create(lat.edges, J_x=-1.0, J_y=-1.0, J_z=delta, dtype=np.float32)
        G = JastrowSpinHalf(1024, lat.edges, beta=0.3)
        p = get_problem(H, G)
        p.set_lanes(64)
        chain = Chain(p, prepare_walkers_sz(args.walkers, 1024, 12345))
        alw = _warmup(chain, PhiloxKey(777, 0), 0.2, None, args.gfmc_warmup, 0)
        chain.steps(500, 0.2, alw, 777, 1 << 20)
        (lanes, wpb, xcd), _ = tune_problem(p, chain.spins, 0.2, alw)
        graph = StepGraph(chain, 10, 0.2, alw, 777, evals=True)
        pos = (1 << 20) + 500
        graph.launch(pos)
        chain.evals.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(2):
            graph.launch(pos + 10 * (r + 1))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[f"delta{delta:g}"] = {
            "walker_steps_per_s": args.walkers * 20 / dt, "ms_per_step": dt / 20 * 1e3,
            "evals_per_walker_step": float(chain.evals.to(torch.float64).sum().item())
            / (args.walkers * 20), "lanes_per_walker": lanes, "waves_per_workgroup": wpb,
            "xcd_contiguous_walkers": bool(xcd)}
    return {"workload": f"xxz_square32x32_jastrow0.3_w{args.walkers}_step0.2_fp32", **res}


def chain100_section():
    """The reference's plain method: Heisenberg ring L = 100 (sign-free frame),
    Jastrow beta 0.3, 1024 walkers Sz = 0, step 0.2, p = [4, 16, 64, 128],
    2000 warm-up + 1000 production steps, FP64; E/L against the Bethe-ansatz
    value -0.4432296."""
    from pygfmc_amd import JastrowSpinHalf, SpinHalf, chain, prepare_walkers_sz
    lat = chain(100)
    H = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    G = JastrowSpinHalf(100, lat.edges, beta=0.3, dtype=np.float64)
    w = prepare_walkers_sz(1024, 100, 0)
    secs, E, err, st = timed_run(w, 1000, H, G, [4, 16, 64, 128], 0, 2000)
    return {"workload": "heisenberg_chain100_jastrow0.3_w1024_step0.2_fp64",
            "walker_steps_per_s": 1024 * 1000 / st["runtime"], "it_per_sec": st["it_per_sec"],
            "gpu_s_total": secs, "E_per_site": E / 100, "sigma_per_site": err / 100,
            "exact_per_site": -0.4432296,
            "z": (E / 100 + 0.4432296) / (err / 100) if err > 0 else None,
            "lanes_per_walker": st["parameters/lanes_per_walker"],
            "evals_per_step": st["evals_per_step_mean"]}


# The reference's published throughput (paper Fig. 3 data, it/s x walkers):
# L = 100 Heisenberg ring, no guide, Neel start, single_step, n_bra = 6,
# lam = 25; walker-steps/s on one GPU.
FIG3_PUBLISHED = {65536: {"A100": 356.04 * 65536, "H200": 554.18 * 65536},
                  2097152: {"A100": 14.234 * 2097152, "H200": 27.647 * 2097152}}


def fig3_section():
    """The reference's own throughput benchmark configuration (Heisenberg
    ring L = 100, J_xy = -1, J_z = 1, no guide, Neel initial state,
    method single_step, n_bra = 6, lam = 25, FP32 rates) at 65 536 and
    2 097 152 walkers, against its published one-GPU A100 / H200 numbers.
    Shorter than the published protocol (10 000 + 10 000 steps): 2000 warm-up
    + 3000 timed steps at 65 536 walkers, 200 + 300 at 2 097 152."""
    from pygfmc_amd import SpinHalf, Walker, chain, prepare_walkers_from_configs, run
    lat = chain(100)
    H = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    neel = np.tile([1.0, -1.0], 50)
    out = {"workload": "heisenberg_chain100_noguide_neel_single_step_nbra6_lam25_fp32"}
    for nw, warm, steps in ((65536, 2000, 3000), (2097152, 200, 300)):
        one = prepare_walkers_from_configs(neel[None])
        w = Walker.from_device(one.spins.repeat(nw, 1).contiguous(),
                               one.log_weight_dev.repeat(nw), one.e_loc_dev.repeat(nw), 100)
        _, E, _, st, _ = run(w, steps, H, 6, None, [4], method="single_step", lam=25.0,
                             rng_key=1, warmup_steps=warm, verbosity=0, number_of_outputs=10)
        v = nw * st["it_per_sec"]
        pub = FIG3_PUBLISHED[nw]
        out[f"w{nw}"] = {"walker_steps_per_s": v, "it_per_sec": st["it_per_sec"],
                         "steps": steps, "warmup": warm, "E_per_site_p4": E[4] / 100,
                         "lanes_per_walker": st["parameters/lanes_per_walker"],
                         "published_A100": pub["A100"], "published_H200": pub["H200"],
                         "vs_A100": v / pub["A100"], "vs_H200": v / pub["H200"]}
        del w
    return out


def tfim_batch_section():
    """64 transverse-field Ising chains in one launch (J = 1, h ~ U[0.5, 1.5]
    and L ~ U{16..256} from seed 0), Jastrow guides, 2048 walkers each, step
    0.2, p = [4, 16, 64], 400 warm-up + 1500 steps, FP32 ratios; z-scores of
    E(p = 64) against the free-fermion ground energies.

    Each chain's Jastrow beta is chosen by a batched scan (the four candidates
    of every chain in one launch of 256 lattices x 256 walkers, 300 + 300
    steps): the one whose step energies vary least.  Chains in the ordered
    phase (h < 1) start fully polarised: from random configurations their
    populations keep a pair of domain walls (energy 4 (J - h)) for far longer
    than the run.  Reported next to the fixed beta = -0.25 guide from random
    configurations."""
    import warnings

    from pygfmc_amd import (JastrowSpinHalf, chain, evaluate_observables,
                            prepare_walkers_random, run_batch, tfim_chain_ground_energy,
                            transverse_field_ising)
    from pygfmc_amd.walkers import Walker
    rng = np.random.default_rng(0)
    Ls = rng.integers(16, 257, 64)
    hs = rng.uniform(0.5, 1.5, 64)
    Hs = [transverse_field_ising(chain(int(L)).edges, 1.0, h, dtype=np.float32)
          for L, h in zip(Ls, hs)]

    def start(nw, l, L, h, ordered):
        if ordered and h < 1.0:
            return Walker(np.ones((nw, int(L))))
        return prepare_walkers_random(nw, int(L), l)

    def zscores(raws, pick=64):
        zs = []
        for l in range(64):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                E, _, err, _ = evaluate_observables(raws[l], n_blocks=20)
            zs.append((E[pick] - tfim_chain_ground_energy(int(Ls[l]), 1.0, hs[l])) / err[pick])
        return np.asarray(zs)

    # guide scan
    cands = [-0.2, -0.3, -0.45, -0.6]
    t0 = time.perf_counter()
    Hc = [H for H in Hs for _ in cands]
    Gc = [JastrowSpinHalf(int(L), chain(int(L)).edges, beta=b, dtype=np.float32)
          for L in Ls for b in cands]
    wc = [start(256, l, L, h, True) for l, (L, h) in enumerate(zip(Ls, hs)) for _ in cands]
    _, raws_c, _ = run_batch(wc, 300, Hc, 0.2, Gc, [4], rng_key=5, warmup_steps=300)
    var = np.array([np.var(np.asarray(r["energy"])) for r in raws_c]).reshape(64, len(cands))
    betas = np.array(cands)[np.argmin(var, axis=1)]
    scan_s = time.perf_counter() - t0
    out = {"workload": "tfim_64chains_L16-256_w2048_step0.2_fp32"}
    for tag, bet, ordered in (("selected", betas, True), ("beta-0.25_random", None, False)):
        Gs = [JastrowSpinHalf(int(L), chain(int(L)).edges,
                              beta=-0.25 if bet is None else float(bet[l]), dtype=np.float32)
              for l, L in enumerate(Ls)]
        ws = [start(2048, l, L, h, ordered) for l, (L, h) in enumerate(zip(Ls, hs))]
        _, raws, st = run_batch(ws, 1500, Hs, 0.2, Gs, [4, 16, 64], rng_key=7,
                                warmup_steps=400)
        zs = zscores(raws)
        near = (hs > 0.8) & (hs < 1.0)
        res = {"z_rms_all": float(np.sqrt(np.mean(zs ** 2))),
               "frac_abs_z_lt_3": float(np.mean(np.abs(zs) < 3)),
               "z_rms_L_le_64": float(np.sqrt(np.mean(zs[Ls <= 64] ** 2))),
               "z_rms_h_outside_0.8_1.0": float(np.sqrt(np.mean(zs[~near] ** 2))),
               "z_rms_h_in_0.8_1.0": float(np.sqrt(np.mean(zs[near] ** 2))),
               "z_rms_L_gt_128_h_outside_0.8_1.0": float(np.sqrt(np.mean(
                   zs[(Ls > 128) & ~near] ** 2)))}
        if tag == "selected":
            out.update({"walker_steps_per_s": st["walker_steps_per_sec"],
                        "it_per_sec": st["it_per_sec"], "n_walkers_total": st["n_walkers_total"],
                        "lanes_per_walker": st["lanes_per_walker"],
                        "autotune_seconds": st["autotune_seconds"], "hip_graph": st["hip_graph"],
                        "evals_per_step": st["evals_per_step_mean"],
                        "guide_scan_seconds": scan_s,
                        "jastrow_beta_counts": {f"{b:g}": int(np.sum(betas == b)) for b in cands}})
            out.update(res)
        else:
            out[tag] = res
    return out


def kagome_mlp_section(args):
    """Large learned guide: kagome XXZ 12 x 12 unit cells (432 sites, 864
    edges, J_x = J_y = -1, J_z = 1, Sz = 0), two-layer MLP guide 432 -> 512 ->
    256 -> 1 (BF16 MFMA, random initialisation: trained-guide quality does
    not change the work per row), step 0.2, FP32 rates, FP64 weights.  One
    untimed step, then ``--mlp-steps`` timed branching steps; the layer-2
    MFMA work is counted from the row blocks the propagation processed.
    Then one training gradient (variance minimisation) over 4096 of the
    walkers."""
    import torch

    from pygfmc_amd import MLPSpinHalf, SpinHalf, kagome, prepare_walkers_sz
    from pygfmc_amd import _lib
    from pygfmc_amd.engine import Chain
    from pygfmc_amd.problem import get_problem
    from pygfmc_amd.training import vmin_gradient
    import ctypes
    lat = kagome(12, 12)
    H = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    G = MLPSpinHalf(lat.N, (512, 256), seed=0, scale=1.0, out_scale=0.1)
    p = get_problem(H, G)
    nw, beta = args.mlp_walkers, 0.2
    w = prepare_walkers_sz(nw, lat.N, 3)
    chain = Chain(p, w)
    _, e, _ = p.rates(chain.spins, want_rates=False)
    alw = float(torch.mean(-e).item())
    chain.steps(1, beta, alw, 11, 0)
    cnt0 = (ctypes.c_int64 * 2)()
    _lib.check(_lib.load().gfmc_mlp_counters(p.handle, cnt0), "gfmc_mlp_counters")
    chain.evals.zero_()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.mlp_steps)]
    ser = torch.empty((args.mlp_steps, 2), dtype=torch.float64, device=chain.spins.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.mlp_steps):
        chain.step(beta, alw, 11, 1 + k, ser[k], events=ev[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prop_s = float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3
    cnt1 = (ctypes.c_int64 * 2)()
    _lib.check(_lib.load().gfmc_mlp_counters(p.handle, cnt1), "gfmc_mlp_counters")
    blocks = cnt1[0] - cnt0[0]
    evals = float(chain.evals.to(torch.float64).sum().item())
    flop = blocks * 16 * 2.0 * 512 * 256
    if not np.all(np.isfinite(ser[:, 0].cpu().numpy())):
        raise SystemExit("population extinct in the MLP section")
    evals_ws = evals / (nw * args.mlp_steps)
    prof = profiled_traffic(f"k_mlp_gemm32_kagome12x12_mlp512x256_w{nw}", "traffic_bytes_per_step")
    roof = {"bound": "mfma", "compute_unit": "BF16 MFMA 32x32x16, FP32 accumulation (layer 2)",
            "achieved": flop / args.mlp_steps / prop_s / 1e12, "peak": 2500.0,
            "unit": "TFLOP/s", "frac": flop / args.mlp_steps / prop_s / 2.5e15,
            "work": "layer-2 FLOPs 2 * 512 * 256 per network row over the 16-row blocks the "
                    "propagation processed (padding rows included)",
            "kernel": "k_mlp_gemm32<8, true> (the propagation of one branching step, timed "
                      "with HIP events around it, includes its row selection and scans)",
            "kernel_avg_ms": prop_s * 1e3,
            "traffic": prof["traffic_bytes_per_launch"] if prof else None,
            "traffic_unit": "HBM bytes per branching step (rocprofv3 --pmc FETCH_SIZE + "
                            "WRITE_SIZE over the step's k_mlp_gemm32 launches, copy-calibrated)",
            "traffic_source": prof["source"] if prof else None}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(H, G, chain, lat.N, beta, alw, 11, args.cpu_seconds,
                           beta_sample=0.002, evals_per_ws=evals_ws, max_walkers=256)
    n_s = 4096
    sub = chain.spins[:n_s].contiguous()
    vmin_gradient(H, G, sub)   # untimed (workspace)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    vmin_gradient(H, G, sub)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t1
    return {"workload": f"kagome12x12_xxz_mlp512x256_bf16_w{nw}_step0.2",
            "walker_steps_per_s": nw * args.mlp_steps / dt,
            "s_per_step": dt / args.mlp_steps, "steps": args.mlp_steps,
            "evals_per_walker_step": evals_ws,
            "rows_per_eval": blocks * 16 / evals,
            "hop_iterations_per_step": (cnt1[1] - cnt0[1]) / args.mlp_steps,
            "mfma_layer2_tflops_per_s": flop / dt / 1e12,
            "mfma_layer2_frac_of_bf16_peak": flop / dt / 2.5e15,
            "train_gradient_seconds_4096_samples": t_grad,
            "note": "layer-2 FLOP counted over 16-row blocks (padding included); the "
                    "per-row layer-1 activations (512 tanh + BF16) run on VALU",
            "roofline": roof, "cpu_baseline": cpu}


def timed_run(walkers, n_steps, H, G, p_array, seed, warmup, n_blocks=20):
    """run() bracketed by device synchronisation: (GPU seconds, E, sigma at
    max p, stats); the jackknife over ``n_blocks`` blocks."""
    import warnings

    import torch

    from pygfmc_amd import evaluate_observables, run
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, _, st, raw = run(walkers, n_steps, H, 0.2, G, p_array, rng_key=seed,
                           warmup_steps=warmup, verbosity=0)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        E, _, err, _ = evaluate_observables(raw, n_blocks=n_blocks)
    pm = max(p_array)
    return secs, float(E[pm]), float(err[pm]), st


def sigma_race(args):
    """Second metric on the 10 x 10 periodic Heisenberg lattice (J_x = J_y =
    -1, J_z = 1, Sz = 0), p = [4, 16, 64, 128], 16 384 walkers, step 0.2:
    GPU-seconds until the jackknife error bar (blocks of 200 steps) of E at
    p = 128 reaches that of the Jastrow-FP64 run (20 000 production steps
    after 400 warm-up steps), for the learned BF16 RBM guide (alpha = 4: four
    translation-invariant filters, 400 hidden units; started as the
    nearest-neighbour Jastrow guide and trained on the GPU by local-energy
    variance minimisation with natural-gradient steps on the walkers its own
    GFMC generates) with and without its training time.  The RBM's time to
    the target error bar is extrapolated from its own run by sigma^2 ~ 1 /
    production steps.  Two Jastrow baselines: beta = 0.15 (the best of a
    scan) and beta = 0.3 (the reference configurations' guide, which is
    biased on this lattice at this walker count: population-control bias
    from its large local-energy variance)."""
    import torch

    from pygfmc_amd import (JastrowSpinHalf, RBMSpinHalf, SpinHalf, prepare_walkers_sz, square,
                            train_variance)
    from pygfmc_amd.problem import get_problem
    lat = square(10, 10)
    P = [4, 16, 64, 128]
    nw = 16384
    H64 = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    H32 = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    out = {"workload": "heisenberg_square10x10_w16384_step0.2_p4-128",
           "exact_E": -67.1549, "exact_source": "Sandvik, PRB 56, 11678 (1997): E/N = -0.671549 "
                                                 "at L = 10"}
    jas = {}
    for beta in (0.15, 0.3):
        G = JastrowSpinHalf(100, lat.edges, beta=beta, dtype=np.float64)
        w = prepare_walkers_sz(nw, 100, 0)
        secs, E, err, st = timed_run(w, args.race_steps, H64, G, P, 1, 400,
                                     n_blocks=args.race_steps // 200)
        jas[beta] = dict(beta=beta, gpu_s=secs, E=E, sigma=err, steps=args.race_steps,
                         warmup=400, walkers=nw, evals_per_step=st["evals_per_step_mean"],
                         ms_per_step=st["runtime"] / args.race_steps * 1e3)
        progress(f"jastrow beta={beta}: E {E:.4f} +- {err:.4f}")
    out["jastrow_fp64_beta0.15"] = jas[0.15]
    out["jastrow_fp64_beta0.3"] = jas[0.3]
    # training: the RBM starts as the beta = 0.15 Jastrow guide
    rbm0 = RBMSpinHalf.from_jastrow(lat.edges, lat.extras["translations"], 0.15, 4, seed=1,
                                    scale=0.01)
    w = prepare_walkers_sz(4096, 100, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G, hist = train_variance(H32, rbm0, w, n_iter=args.race_train_iters, steps_per_iter=2,
                             lr=1.0, clip=3.0, rng_key=3, sampler="self", warmup_steps=100,
                             optimizer="sr", sr_shift=1e-2, sr_abs=1e-4, sr_max_norm=0.05)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    v = hist["variance"]
    progress(f"rbm trained: variance {v[:5].mean():.3f} -> {v[-20:].mean():.3f} in {train_s:.1f}s")
    # production with the trained guide
    w = prepare_walkers_sz(nw, 100, 0)
    warm = 100
    secs, E, err, st = timed_run(w, args.race_steps, H32, G, P, 1, warm,
                                 n_blocks=args.race_steps // 200)
    frac_prod = args.race_steps / (args.race_steps + warm)
    rbm = dict(alpha=4, hidden=400, filters=4, train_gpu_s=train_s,
               train_iters=args.race_train_iters, train_walkers=4096, optimizer="natural gradient",
               variance_start=float(v[:5].mean()), variance_end=float(v[-20:].mean()),
               gpu_s=secs, E=E, sigma=err, steps=args.race_steps, warmup=warm, walkers=nw,
               evals_per_step=st["evals_per_step_mean"],
               ms_per_step=st["runtime"] / args.race_steps * 1e3)
    # the propagation kernel's roofline on this workload and the CPU oracle
    rbm["roofline"], rbm["cpu_baseline"] = rbm_kernel_measure(H32, G, args.cpu_seconds)
    out["rbm_bf16"] = rbm
    for beta, J in jas.items():
        t_warm = secs * (1 - frac_prod)
        t_need = t_warm + secs * frac_prod * (err / J["sigma"]) ** 2
        z = (E - J["E"]) / np.hypot(err, J["sigma"])
        out[f"vs_jastrow_beta{beta:g}"] = dict(
            target_sigma=J["sigma"], jastrow_gpu_s=J["gpu_s"], rbm_gpu_s=t_need,
            rbm_gpu_s_with_training=t_need + train_s,
            speedup=J["gpu_s"] / t_need, speedup_with_training=J["gpu_s"] / (t_need + train_s),
            energy_z=float(z),
            per_step_variance_ratio=(J["sigma"] ** 2 * J["steps"]) / (err ** 2 * rbm["steps"]),
            per_step_cost_ratio=rbm["ms_per_step"] / J["ms_per_step"],
            # each sigma^2 from n jackknife blocks carries a relative error of
            # about sqrt(2 / n); the speed-up is their ratio
            speedup_rel_err=float(np.sqrt(2.0 / (J["steps"] // 200) + 2.0 / (rbm["steps"] // 200))))
    return out


def rbm_kernel_measure(H, G, cpu_seconds):
    """The RBM propagation kernel (k_propagate_rbt, table factors, for the
    translation-tied guide; k_propagate_rbh otherwise) on 16 384 walkers of
    the trained guide, equilibrated:
    average launch time by HIP events on the launch stream against the
    algorithmic FP32 work of the reference algorithm (every slot's log-ratio
    over all hidden units: n_edges H 8 flop per rate evaluation), and the FP64
    oracle propagating a bounded sample of the same walkers on the host."""
    import torch

    from pygfmc_amd import prepare_walkers_sz
    from pygfmc_amd.engine import Chain
    from pygfmc_amd.problem import get_problem
    p = get_problem(H, G)
    nw = 16384
    ch = Chain(p, prepare_walkers_sz(nw, H.N, 5))
    _, e, _ = p.rates(ch.spins, want_rates=False)
    alw = float(torch.mean(-e).item())
    ch.steps(200, 0.2, alw, 9, 0)
    n_k = 10
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(n_k)]
    out2 = torch.empty(2, dtype=torch.float64, device=ch.spins.device)
    ch.evals.zero_()
    for k in range(n_k):
        ch.step(0.2, alw, 9, 1000 + k, out2, events=ev[k])
    torch.cuda.synchronize()
    t = float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3
    evals = float(ch.evals.to(torch.float64).sum().item()) / n_k
    flop = evals * len(H.edge_i) * G.hidden * 8.0
    kname = "k_propagate_rbt" if p.tied is not None else "k_propagate_rbh"
    prof = profiled_traffic(f"{kname}_heisenberg_square10x10_rbm400_w16384")
    if p.tied is not None:
        td = p.tied
        kfull = f"{kname}<{td['n_pc']}, {td['n_sc']}, {'true' if td['packed'] else 'false'}>"
    else:
        kfull = f"{kname}<{(G.hidden + 63) // 64}>"
    roof = {"bound": "valu", "compute_unit": "FP32 vector ALU (no GEMM in the hop; the MI355X "
                                              "FP32 vector and matrix peaks are both 157.3 "
                                              "TFLOP/s)",
            "achieved": flop / t / 1e12, "peak": 157.3, "unit": "TFLOP/s",
            "frac": flop / t / 1e12 / 157.3,
            "work": "reference-equivalent FLOPs: every slot's log-ratio over all hidden units, "
                    "n_edges H 8 per rate evaluation",
            "traffic": prof["traffic_bytes_per_launch"] if prof else None,
            "traffic_unit": "HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, "
                            "copy-calibrated)",
            "traffic_source": prof["source"] if prof else None,
            "kernel": kfull, "kernel_avg_ms": t * 1e3,
            "evals_per_walker_step": evals / nw, "flop_per_eval": len(H.edge_i) * G.hidden * 8.0,
            "transcendentals_per_eval": len(H.edge_i) * G.hidden * 2.0}
    cpu = cpu_baseline(H, G, ch, H.N, 0.2, alw, 9, cpu_seconds)
    return roof, cpu


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(H, G, chain, N, beta, alw, seed, seconds, beta_sample=None, evals_per_ws=None,
                 max_walkers=None):
    """The FP64 oracle on the host cores: propagation of a bounded sample of
    the equilibrated walkers (same Hamiltonian, guide and step).  With
    ``beta_sample`` < ``beta`` (a guide too slow on the host for whole
    steps) the sample propagates for beta_sample only and walker-steps/s is
    the oracle's rate evaluations per second divided by the GPU run's
    evaluations per walker-step ``evals_per_ws``."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    from pygfmc_amd.problem import lower
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    prob = oracle.Problem(lower(H, G))
    spins = chain.spins.cpu().numpy().view(np.uint32)
    if max_walkers:
        spins = spins[:max_walkers]
    bs = beta if beta_sample is None else beta_sample
    n = min(16, len(spins))
    done = 0
    evals = 0
    t0 = time.perf_counter()
    step = 0
    while True:
        take = spins[:n]
        _, _, _, ev = prob.propagate(take, np.zeros(n), np.zeros(n), bs, alw, seed,
                                     10_000_000 + step, 0, n_threads=threads)
        done += n
        evals += int(ev.sum())
        step += 1
        el = time.perf_counter() - t0
        if el > seconds:
            break
        if el < seconds / 8 and n < len(spins):
            n = min(len(spins), n * 2)
    el = time.perf_counter() - t0
    if beta_sample is None:
        value = done / el
        sample = (f"{done} walker-propagations of the equilibrated bench walkers ({el:.1f} s, "
                  f"FP64 C oracle, OpenMP over walkers)")
    else:
        value = evals / el / evals_per_ws
        sample = (f"{done} walker-propagations over imaginary time {bs:g} of the equilibrated "
                  f"bench walkers ({evals} rate evaluations in {el:.1f} s, FP64 C oracle, OpenMP "
                  f"over walkers), scaled by the GPU run's {evals_per_ws:.2f} evaluations per "
                  f"walker-step")
    return {"value": value, "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "sample": sample}


if __name__ == "__main__":
    main()
