"""Benchmark: particles unbound per second on one MI355X.

Workload (default): a power-law subhalo population drawn until it holds
5x10^7 particles (dN/dN_p ~ N_p^-1.9 on [20, 10^6], log-normal concentrations,
10% interlopers), theta=0.7, eps=1 kpc/h, BoundMassPrecision 0.995, no
subsampling.  One step = one complete unbinding of the whole population
(every subhalo iterated to convergence, final ordering and summary outputs),
with inputs resident in HBM.

Reported beside the headline value:
  * the plain GPU loop (one subhalo at a time, host-driven iterations: the
    drop-in Subhalo_t::Unbind called per subhalo) timed on a stratified sample
    (every subhalo above 10^4 particles plus a random share of the rest) and
    extrapolated by sampling weight;
  * the CPU oracle (the reference algorithm restated in C++, OpenMP over
    subhaloes like RefineParticles) on a bounded random sample;
  * the roofline of the dominant kernel (tree walk / direct sum).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hbt-herons_amd"))

METRIC = "particles unbound/s: batched+learned vs plain GPU at ≥99.9% bound-mass match"
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def walk_flops(st):
    # 10 FLOP per opening test (3 sub, 5 for r^2, 2 for the test), +6 per accepted
    # interaction (rsqrt, mul, add, 3 compensation), 14 per direct pair.
    return 10.0 * st["tree_lane_visits"] + 6.0 * st["tree_lane_accepts"] + 14.0 * st["direct_lane_pairs"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=float, default=5e7)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--direct-cutoff", type=int, default=-1,
                    help="sources <= cutoff use the exact sum; -1 = tuned value, 0 = tree everywhere")
    ap.add_argument("--plain-sample", type=float, default=0.01,
                    help="fraction of subhaloes below 1e4 particles in the plain-loop sample")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; multi-GPU sharding is described "
                                   "in DESIGN.md but not built"}))
        sys.exit(2)

    import hbtx
    import synth

    t0 = time.time()
    pop = synth.population(seed=args.seed, target_particles=int(args.particles))
    log(f"generated {pop.n} particles in {pop.nsub} subhaloes in {time.time() - t0:.1f}s")
    prm = hbtx.workload_params(direct_cutoff=args.direct_cutoff)

    ctx = hbtx.Context(0)
    ctx.load(pop)
    for _ in range(args.warmup):
        ctx.run(prm)
    ctx.sync()
    times = []
    for k in range(args.steps):
        ctx.sync()
        w0 = time.perf_counter()
        ms = ctx.run(prm)
        ctx.sync()
        w1 = time.perf_counter()
        times.append((ms, (w1 - w0) * 1e3))
        log(f"step {k}: {ms:.2f} ms (events) {times[-1][1]:.2f} ms (wall)")
    walk_ms = ctx.stats()["walk_ms"]
    out = ctx.fetch()
    # work counters: one extra, untimed run of the counting instantiation (same results, deterministic work)
    ctx.run(prm, hbtx.COUNT)
    st = ctx.stats()
    st["walk_ms"] = walk_ms
    ms_step = float(np.mean([t[1] for t in times]))
    value = pop.n / (ms_step / 1e3)
    res = out["result"]
    flops = walk_flops(st)
    achieved = flops / (walk_ms / 1e3) / 1e12 if walk_ms > 0 else 0.0
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "power-law NFW subhalo population, batched unbinding",
                   "n_particles": int(pop.n), "n_subhalos": int(pop.nsub), "theta": 0.7,
                   "softening": 0.001, "bound_mass_precision": 0.995, "max_sample_size": 0,
                   "direct_cutoff": int(args.direct_cutoff), "seed": args.seed,
                   "rounds": int(st["rounds"]), "bound_particles": int(res["nbound"].sum()),
                   "disrupted": int(res["disrupted"].sum())},
        "roofline": {"bound": "valu", "kernel": "k_walk_energy", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": None, "kernel_ms_per_step": walk_ms,
                     "flop_per_step": flops},
        "walk_stats": {k: v for k, v in st.items() if not k.startswith("unused")},
    }

    if not args.no_plain:
        counts = pop.counts()
        rng = np.random.default_rng(123)
        big = np.where(counts > 10000)[0]
        small = np.where(counts <= 10000)[0]
        pick = rng.choice(small, max(1, int(args.plain_sample * len(small))), replace=False) if len(small) else []
        sample = np.concatenate([big, np.sort(pick)]).astype(np.int64)
        weight_small = len(small) / max(1, len(pick))
        sub = pop.subset(sample)
        c2 = hbtx.Context(0)
        c2.load(sub)
        c2.run(prm, hbtx.PLAIN)  # warm-up
        w0 = time.perf_counter()
        c2.run(prm, hbtx.PLAIN)
        c2.sync()
        # per-subhalo timing is not exposed; split the sample run by re-running the two strata
        t_all = time.perf_counter() - w0
        subb = pop.subset(big)
        c2.load(subb)
        w0 = time.perf_counter()
        c2.run(prm, hbtx.PLAIN)
        c2.sync()
        t_big = time.perf_counter() - w0
        t_small = max(0.0, t_all - t_big)
        plain_est = t_big + t_small * weight_small
        c2.close()
        plain_rate = pop.n / plain_est
        line["plain_gpu"] = {"value": plain_rate, "unit": "particles/s", "seconds_estimated": plain_est,
                             "sample": f"all {len(big)} subhaloes with >1e4 particles plus {len(pick)} of "
                                       f"{len(small)} smaller ones (weight {weight_small:.1f})"}
        line["speedup_vs_plain"] = value / plain_rate
        log(f"plain GPU estimate {plain_est:.2f}s -> {plain_rate:.3e} particles/s")

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        cores = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(321)
        order = rng.permutation(pop.nsub)
        # grow the random sample until its CPU time reaches the target
        take = 64
        secs = 0.0
        n_done = 0
        while True:
            ids = np.sort(order[:take])
            sub = pop.subset(ids)
            r = pyoracle.unbind_batch(sub, pyoracle.default_params(bound_mass_precision=0.995),
                                      nthreads=cores)
            secs = r["seconds"]
            n_done = sub.n
            if secs >= args.cpu_seconds * 0.5 or take >= pop.nsub:
                break
            take = min(pop.nsub, int(take * max(2.0, min(8.0, args.cpu_seconds / max(secs, 1e-3)))))
        line["cpu_baseline"] = {"value": n_done / secs, "unit": "particles/s", "cores": cores, "kind": "port",
                                "sample": f"{take} random subhaloes ({n_done} particles) of the same "
                                          f"population, {secs:.1f}s of unbinding"}
        log(f"cpu baseline {n_done / secs:.3e} particles/s on {cores} cores")
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
