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
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--delta", type=float, default=1.0)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before the timed steps, copy a tensor of this many MiB once "
                         "(a plain copy kernel of known size, to scale HBM counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; runs on more than one GPU "
                                   "are not implemented (walkers would shard by contiguous "
                                   "ranges with a scalar scan of the weight sums)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

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
    _, e, _ = p.rates(chain.spins, want_rates=False)
    alw = float(torch.mean(-e).item())
    seed = 777
    series = torch.empty((max(args.steps, args.warmup, 1), 2), dtype=torch.float64,
                         device=chain.spins.device)

    def one_step(t, row, ev_pair=None):
        chain.step(beta, alw, seed, t, series[row], events=ev_pair)

    # warm-up: equilibrate and settle the shift (untimed)
    for t in range(args.warmup):
        one_step(t, t)
        if t % 5 == 4:
            lg = series[t - 4: t + 1, 0].cpu().numpy()
            alw += float(np.mean(lg)) / beta
    torch.cuda.synchronize()

    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib * (1 << 18), dtype=torch.float32,
                         device=chain.spins.device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    chain.evals.zero_()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(args.warmup + k, k, events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lg = series[: args.steps, 0].cpu().numpy()
    if not np.all(np.isfinite(lg)):
        raise SystemExit("population extinct during the bench")
    prop_ms = [a.elapsed_time(b) for a, b in events]
    prop_avg_s = float(np.mean(prop_ms)) / 1e3
    evals_total = float(chain.evals.to(torch.float64).sum().item())
    evals_per_ws = evals_total / (nw * args.steps)
    ms_per_step = elapsed / args.steps * 1e3
    value = nw * args.steps / elapsed

    # algorithmic FP32 work of one launch (reference algorithm: every slot of
    # every rate evaluation): n_edges * (4 (d - 1) + 6) per evaluation
    d = G.max_deg
    flop_per_eval = E * (4 * (d - 1) + 6)
    flops_launch = evals_total / args.steps * flop_per_eval
    achieved_tf = flops_launch / prop_avg_s / 1e12
    peak_tf = 157.3

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(H, G, chain, N, beta, alw, seed, args.cpu_seconds)

    out = {
        "metric": "walker-steps/s; GPU-sec to fixed energy σ, learned BF16 guide vs plain",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else "fp64",
        "data": "synthetic",
        "config": {"workload": f"xxz_square{L}x{L}_delta{args.delta:g}_jastrow0.3_"
                               f"w{nw}_step0.2_{args.precision}",
                   "lattice": f"square {L}x{L} periodic", "n_sites": N, "n_edges": E,
                   "n_walkers": nw, "delta": args.delta, "step_size": beta,
                   "guide": "Jastrow beta=0.3", "global_batch": nw, "seq_len": N,
                   "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "mfma", "compute_unit": "VALU FP32 (no GEMM on this path)",
                     "achieved": achieved_tf, "peak": peak_tf, "unit": "TFLOP/s",
                     "frac": achieved_tf / peak_tf, "traffic": None,
                     "kernel": "k_propagate<float>", "kernel_avg_ms": prop_avg_s * 1e3,
                     "evals_per_walker_step": evals_per_ws,
                     "flop_per_eval": flop_per_eval},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(H, G, chain, N, beta, alw, seed, seconds):
    """The FP64 oracle on the host cores: propagation of a bounded sample of
    the equilibrated walkers (same Hamiltonian, guide and step)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    from pygfmc_amd.problem import lower
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    prob = oracle.Problem(lower(H, G))
    spins = chain.spins.cpu().numpy().view(np.uint32)
    n = 16
    done = 0
    t0 = time.perf_counter()
    step = 0
    while True:
        take = spins[:n]
        prob.propagate(take, np.zeros(n), np.zeros(n), beta, alw, seed, 10_000_000 + step, 0,
                       n_threads=threads)
        done += n
        step += 1
        el = time.perf_counter() - t0
        if el > seconds:
            break
        if el < seconds / 8 and n < len(spins):
            n = min(len(spins), n * 2)
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "sample": f"{done} walker-propagations of the equilibrated bench walkers "
                      f"({el:.1f} s, FP64 C oracle, OpenMP over walkers)"}


if __name__ == "__main__":
    main()
