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
    ap.add_argument("--no-race", action="store_true",
                    help="skip the second metric (GPU-seconds to a fixed energy error bar, "
                         "learned RBM guide vs Jastrow, 10x10 Heisenberg)")
    ap.add_argument("--race-train-iters", type=int, default=600)
    ap.add_argument("--race-rbm-steps", type=int, default=300)
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
    if not args.no_race:
        out["sigma_race"] = sigma_race(args)
    print(json.dumps(out), flush=True)


def timed_run(walkers, n_steps, H, G, p_array, seed, warmup):
    """run() bracketed by device synchronisation: (GPU seconds, E, sigma at
    max p, stats)."""
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
        E, _, err, _ = evaluate_observables(raw, n_blocks=20)
    pm = max(p_array)
    return secs, float(E[pm]), float(err[pm]), st


def sigma_race(args):
    """Second metric on the 10 x 10 periodic Heisenberg lattice (J_x = J_y =
    -1, J_z = 1, Sz = 0), p = [4, 16, 64, 128], 16 384 walkers, step 0.2:
    GPU-seconds until the jackknife error bar (20 blocks) of E at p = 128
    reaches that of the Jastrow-FP64 run (1000 production steps after 400
    warm-up steps), for the learned BF16 RBM guide (alpha = 4, trained on the
    GPU by local-energy variance minimisation on GFMC samples of the beta =
    0.15 Jastrow guide) with and without its training time.  The RBM's time
    to the target error bar is extrapolated from its own run by sigma^2 ~
    1 / production steps.  Two Jastrow baselines: beta = 0.3 (the reference
    configurations' guide) and beta = 0.15 (the best of a scan)."""
    import torch

    from pygfmc_amd import (JastrowSpinHalf, RBMSpinHalf, SpinHalf, prepare_walkers_sz, square,
                            train_variance)
    lat = square(10, 10)
    P = [4, 16, 64, 128]
    nw = 16384
    H64 = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float64)
    H32 = SpinHalf.create(lat.edges, -1.0, -1.0, 1.0, dtype=np.float32)
    out = {"workload": "heisenberg_square10x10_w16384_step0.2_p4-128",
           "exact_E": -67.1549, "exact_source": "Sandvik, PRB 56, 11678 (1997): E/N = -0.671549 "
                                                 "at L = 10"}
    jas = {}
    for beta in (0.3, 0.15):
        G = JastrowSpinHalf(100, lat.edges, beta=beta, dtype=np.float64)
        w = prepare_walkers_sz(nw, 100, 0)
        secs, E, err, st = timed_run(w, 1000, H64, G, P, 1, 400)
        jas[beta] = dict(beta=beta, gpu_s=secs, E=E, sigma=err, steps=1000, warmup=400,
                         walkers=nw, evals_per_step=st["evals_per_step_mean"])
    out["jastrow_fp64_beta0.3"] = jas[0.3]
    out["jastrow_fp64_beta0.15"] = jas[0.15]
    # training
    sampler = JastrowSpinHalf(100, lat.edges, beta=0.15, dtype=np.float32)
    rbm0 = RBMSpinHalf(100, alpha=4, seed=0, scale=0.01)
    w = prepare_walkers_sz(4096, 100, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G, hist = train_variance(H32, rbm0, w, n_iter=args.race_train_iters, steps_per_iter=2,
                             lr=1e-3, clip=3.0, rng_key=3, sampler=sampler)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    v = hist["variance"]
    # production with the trained guide
    w = prepare_walkers_sz(nw, 100, 0)
    warm = 100
    secs, E, err, st = timed_run(w, args.race_rbm_steps, H32, G, P, 1, warm)
    frac_prod = args.race_rbm_steps / (args.race_rbm_steps + warm)
    out["rbm_bf16"] = dict(alpha=4, hidden=400, train_gpu_s=train_s,
                           train_iters=args.race_train_iters, train_walkers=4096,
                           variance_start=float(v[:5].mean()), variance_end=float(v[-20:].mean()),
                           gpu_s=secs, E=E, sigma=err, steps=args.race_rbm_steps, warmup=warm,
                           walkers=nw, evals_per_step=st["evals_per_step_mean"])
    for beta, J in jas.items():
        t_warm = secs * (1 - frac_prod)
        t_need = t_warm + secs * frac_prod * (err / J["sigma"]) ** 2
        z = (E - J["E"]) / np.hypot(err, J["sigma"])
        out[f"vs_jastrow_beta{beta:g}"] = dict(
            target_sigma=J["sigma"], jastrow_gpu_s=J["gpu_s"], rbm_gpu_s=t_need,
            rbm_gpu_s_with_training=t_need + train_s,
            speedup=J["gpu_s"] / t_need, speedup_with_training=J["gpu_s"] / (t_need + train_s),
            energy_z=float(z))
    return out


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
