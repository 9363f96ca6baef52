"""Benchmark: particles unbound per second on one MI355X, batched + learned.

Workload (default): a power-law subhalo population drawn until it holds
5x10^7 particles (dN/dN_p ~ N_p^-1.9 on [20, 10^6], log-normal concentrations,
10% interlopers), evolved through 8 ID-preserving snapshots (tidal stripping by
1.5 v_esc kicks outside r_200 * 0.9^k, 2% jitter).  The 8-64-64-1 classifier is
trained on the GPU on the plain unbinding's bound flags of snapshots 0-5; the
timed step is the unbinding of snapshot 6 with the learned warm start:
per-particle features (from snapshot 5's outputs) + inference + the verified,
warm-started batched unbinding (every subhalo to convergence, final ordering
and summary outputs), inputs resident in HBM.  theta=0.7, eps=1 kpc/h,
BoundMassPrecision 0.995, no subsampling.

Reported beside the headline value:
  * the same batched unbinding without the warm start, whose result (identical
    to the plain iteration's) judges the learned one: bound-mass mismatch;
  * the plain GPU loop (one subhalo at a time, host-driven iterations: the
    drop-in Subhalo_t::Unbind called per subhalo) timed on a stratified sample
    (every subhalo above 10^4 particles plus a random share of the rest) and
    extrapolated by sampling weight;
  * the CPU oracle (the reference algorithm restated in C++, OpenMP over
    subhaloes like RefineParticles) on a bounded random sample of snapshot 6;
  * the roofline of the dominant kernel (the tree walk / direct sum).
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


def walk_traffic():
    """HBM bytes per walk launch on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (each scaled by a copy kernel's reading)."""
    path = os.path.join(ROOT, "profiles", "pmc_k_walk_energy_evolved_5e7.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def walk_flops(st):
    # 10 FLOP per opening test (3 sub, 5 for r^2, 2 for the test), +6 per accepted
    # interaction (rsqrt, mul, add, 3 compensation), 14 per direct pair.
    return 10.0 * st["tree_lane_visits"] + 6.0 * st["tree_lane_accepts"] + 14.0 * st["direct_lane_pairs"]


def timed_steps(ctx, steps, warmup, step):
    for _ in range(warmup):
        step()
    ctx.sync()
    times = []
    for k in range(steps):
        ctx.sync()
        w0 = time.perf_counter()
        step()
        ctx.sync()
        times.append((time.perf_counter() - w0) * 1e3)
    return times


def plain_gpu_full(hbtx, pop, prm):
    """The plain GPU loop on the whole evaluated snapshot: Subhalo_t::Unbind one
    subhalo at a time (HBTX_PLAIN, host-driven iterations, the same kernels and
    results as the batched path), every subhalo timed, nothing sampled.  Inputs
    are resident before the clock starts."""
    c2 = hbtx.Context(0)
    c2.load(pop)
    c2.sync()
    w0 = time.perf_counter()
    c2.run(prm, hbtx.PLAIN)
    c2.sync()
    secs = time.perf_counter() - w0
    out = c2.fetch()
    c2.close()
    return secs, out


def other_configs(hbtx, synth, ctx, log):
    """The other workloads, timed as batched unbinding (plain method, inputs
    resident): one NFW halo of 10,000 particles + 2,000 interlopers, 20,000
    power-law subhaloes, and 5,000 hydro-mixture subhaloes with thermal energy;
    and particle reassignment over 200 host haloes of about 3.5e7 particles
    (their parity with the oracle is tested in tests/)."""
    out = {}
    cases = [
        ("single_nfw_12k", synth.single_nfw(seed=0), {}),
        ("population_20000_subhalos", synth.population(seed=2, nsub=20000), {}),
        ("hydro_mixture_5000_subhalos", synth.population(seed=3, nsub=5000, hydro=True), dict(dm_only=0, thermal=1)),
    ]
    for name, pop, kw in cases:
        prm = hbtx.workload_params(direct_cutoff=-1, **kw)
        ctx.load(pop)
        ms = timed_steps(ctx, 3, 1, lambda: ctx.run(prm))
        res = ctx.fetch()
        rec = {"particles": int(pop.n), "subhalos": int(pop.nsub), "ms_per_step": float(np.mean(ms)),
               "particles_per_s": pop.n / (float(np.mean(ms)) / 1e3),
               "bound_particles": int(res["result"]["nbound"].sum())}
        out[name] = rec
        log(f"{name}: {rec['ms_per_step']:.2f} ms, {rec['particles_per_s']:.3e} particles/s")
    out["reassign_host_halos"] = reassign_line(synth, ctx, log)
    return out


def reassign_line(synth, ctx, log, reps=3):
    """Particle reassignment (SubhaloSnapshot_t::ReassignParticles) over 200 host
    haloes with the hydro mixture's types: 79% tracers (dark matter, stars), 21% gas
    and black holes; device time from the uploaded inputs to the targets."""
    h = synth.host_haloes(seed=21, ngroups=200, subs=(1, 12), n_sub=(50, 50000), tracer_frac=0.79)
    ms = []
    for rep in range(reps + 1):
        tgt, m = ctx.reassign_targets(h.pos, h.is_tracer, h.sub, h.group_offsets, k=10)
        if rep:
            ms.append(m)
    rec = {"particles": int(h.n), "host_halos": int(len(h.group_offsets) - 1),
           "subhalos": int(len(h.sub_offsets) - 1), "non_tracers": int((h.is_tracer == 0).sum()),
           "k": 10, "ms_per_step": float(np.mean(ms)), "particles_per_s": h.n / (float(np.mean(ms)) / 1e3),
           "moved": int((tgt >= 0).sum()), "timing": "device, inputs uploaded"}
    log(f"reassign_host_halos: {rec['ms_per_step']:.2f} ms, {rec['particles_per_s']:.3e} particles/s, "
        f"{rec['moved']} moved")
    return rec


CPU_TOP_BIN = 18  # subhaloes of 2^19 particles and more are left out of the CPU sample


def cpu_baseline(pop, args, log):
    """The oracle (the reference's plain method restated in C++) on the host cores,
    on a sample stratified by subhalo size: per log2-size bin b up to CPU_TOP_BIN,
    min(64, 2^(17-b)) random subhaloes (at least one), each bin run as one batch on
    all cores.  The process CPU time per subhalo of a bin, times the bin's count,
    estimates the CPU time of the whole population of those bins; the rate is their
    particles over that time spread over the cores (perfect packing)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = min(16, os.cpu_count() or 1)
    sizes = np.diff(pop.offsets)
    bins = np.floor(np.log2(np.maximum(sizes, 1))).astype(int)
    rng = np.random.default_rng(321)
    cpu_total = n_total = 0.0
    sampled = sampled_n = 0
    wall = 0.0
    for b in range(int(bins.min()), CPU_TOP_BIN + 1):
        members = np.nonzero(bins == b)[0]
        if len(members) == 0:
            continue
        k = int(min(len(members), max(1, min(64, 2 ** (17 - b)))))
        ids = np.sort(rng.choice(members, k, replace=False))
        sub = pop.subset(ids)
        c0, w0 = time.process_time(), time.perf_counter()
        pyoracle.unbind_batch(sub, pyoracle.default_params(bound_mass_precision=0.995), nthreads=cores)
        cpu = time.process_time() - c0
        wall += time.perf_counter() - w0
        cpu_total += cpu / k * len(members)
        n_total += float(sizes[members].sum())
        sampled += k
        sampled_n += sub.n
    rate = n_total / (cpu_total / cores)
    left = int((bins > CPU_TOP_BIN).sum())
    left_n = int(sizes[bins > CPU_TOP_BIN].sum())
    log(f"cpu baseline {rate:.3e} particles/s on {cores} cores ({sampled} subhaloes sampled, {wall:.1f}s)")
    return {"value": rate, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"{sampled} subhaloes ({sampled_n} particles) of snapshot {args.snapshot}, stratified by "
                      f"log2 size up to 2^{CPU_TOP_BIN + 1} particles, plain method, {wall:.1f}s; the rate is the "
                      f"sampled bins' population ({int(n_total)} particles) over its estimated CPU time on "
                      f"{cores} cores; the {left} subhaloes above 2^{CPU_TOP_BIN + 1} particles ({left_n} "
                      f"particles) are neither timed nor extrapolated: the rate covers the other bins only"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=float, default=5e7)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--snapshot", type=int, default=6, help="evaluated snapshot (trained on 0..5)")
    ap.add_argument("--threshold", type=float, default=0.1, help="proposal: P(bound) > threshold")
    ap.add_argument("--warm-min", type=int, default=30000,
                    help="subhaloes of fewer particles run the plain iteration (no proposal).  The warm "
                         "start and the plain iteration take different paths, so particles within the tree "
                         "approximation of E = 0 can fall either way; below ~1e4 bound particles a single "
                         "one exceeds the 1e-3 per-subhalo tolerance.  30000 is the smallest threshold of "
                         "--warm-min-sweep at which every subhalo of the headline workload is within it at "
                         "BoundMassPrecision 0.995 and 1.0")
    ap.add_argument("--warm-min-sweep", default="",
                    help="comma-separated warm-start thresholds: per threshold, the learned step's time and "
                         "its per-subhalo agreement with the plain iteration at both BoundMassPrecision "
                         "settings (an extra line, 'warm_min_sweep')")
    ap.add_argument("--train-per-snapshot", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--direct-cutoff", type=int, default=-1,
                    help="sources <= cutoff use the exact sum; -1 = tuned value, 0 = tree everywhere")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tree", action="store_true", help="skip the host+subhaloes tree-potential line")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the single-halo, 20,000-subhalo and hydro-mixture lines")
    ap.add_argument("--state", default=None,
                    help="npz cache of the trained weights and the previous snapshot's outputs: written "
                         "after training, read instead of labelling and training when it exists")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every round's kernels directly instead of replaying the captured round graphs")
    ap.add_argument("--learned-only", action="store_true",
                    help="time only the learned step (for kernel profiles); implies --no-plain --no-cpu")
    ap.add_argument("--reassign-only", action="store_true",
                    help="time only particle reassignment (for kernel profiles) and print its record")
    args = ap.parse_args()
    if args.learned_only:
        args.no_plain = args.no_cpu = args.no_tree = args.no_configs = True
    if args.gpus != 1:
        print(json.dumps({"error": "hbtx runs on exactly one GPU: sharding the unbinding over several GPUs "
                                   "is designed but not built"}))
        sys.exit(2)
    if args.reassign_only:
        import hbtx
        import synth
        print(json.dumps({"reassign_host_halos": reassign_line(synth, hbtx.Context(0), log)}))
        return

    import hbtx
    import learned
    import synth

    t0 = time.time()
    pop0 = synth.population(seed=args.seed, target_particles=int(args.particles))
    log(f"generated {pop0.n} particles in {pop0.nsub} subhaloes in {time.time() - t0:.1f}s")
    prm = hbtx.workload_params(direct_cutoff=args.direct_cutoff)
    ctx = hbtx.Context(0)

    # --- labels (plain unbinding) of snapshots 0..5 and the training set --------------
    t_lab = t_train = 0.0
    loss = float("nan")
    if args.state and os.path.exists(args.state):
        pop = None
        for k, snap in enumerate(synth.evolve_iter(pop0, nsnap=args.snapshot + 1, seed=4)):
            pop = snap
        z = np.load(args.state)
        prev = {k: z[k] for k in ("rank", "bound", "energy", "com_pos", "com_vel", "mbound", "rscale")}
        net = hbtx.MLP(ctx, 0)
        net.set_weights(z["weights"])
        log(f"loaded trained weights and snapshot {args.snapshot - 1} outputs from {args.state}")
    else:
        rng = np.random.default_rng(0)
        Xs, ys = [], []
        prev = None
        pop = None
        t_lab = time.time()
        for k, snap in enumerate(synth.evolve_iter(pop0, nsnap=args.snapshot + 1, seed=4)):
            pop = snap
            if k == args.snapshot:
                break
            if prev is not None:
                X = learned.snapshot_features(ctx, pop, prev, prm)
            out = ctx.unbind(pop, prm, hbtx.BATCHED)
            if prev is not None:
                take = rng.choice(pop.n, size=min(args.train_per_snapshot, pop.n), replace=False)
                Xs.append(X[take])
                ys.append(learned.labels(pop, out)[take])
                del X
            prev = learned.previous_state(pop, out)
            log(f"snapshot {k}: plain labels, {int(out['result']['nbound'].sum())} bound")
        t_lab = time.time() - t_lab
        X, y = np.concatenate(Xs), np.concatenate(ys)
        del Xs, ys
        t1 = time.time()
        net, loss = learned.train(ctx, X, y, epochs=args.epochs)
        ctx.sync()
        t_train = time.time() - t1
        log(f"trained on {len(X)} samples ({y.mean():.3f} bound) in {t_train:.1f}s, final loss {loss:.4f}")
        del X, y
        if args.state:
            np.savez(args.state, weights=net.weights()[0], **prev)

    # --- evaluated snapshot: batched without and with the learned warm start -----------
    gflag = hbtx.NO_GRAPH if args.no_graph else 0
    ctx.load(pop)
    ctx.set_previous(prev)
    if args.learned_only:
        ms_plain_b, ref = [float("nan")], None
    else:
        ms_plain_b = timed_steps(ctx, args.steps, args.warmup, lambda: ctx.run(prm, gflag))
        ref = ctx.fetch()
        log(f"batched (no warm start): {np.mean(ms_plain_b):.2f} ms/step")

    def learned_step():
        ctx.warm_proposal(net, prm, args.threshold, args.warm_min, count=False)
        ctx.run(prm, hbtx.WARMSTART | gflag)

    ms_learned = timed_steps(ctx, args.steps, args.warmup, learned_step)
    st_time = ctx.stats()
    walk_ms = st_time["walk_ms"]
    rounds = int(st_time["rounds"])
    warm = ctx.fetch()
    npro, inf_ms = ctx.warm_proposal(net, prm, args.threshold, args.warm_min)
    log(f"batched + learned: {np.mean(ms_learned):.2f} ms/step (walk {walk_ms:.2f} ms, {rounds} rounds, "
        f"{npro} proposed, inference {inf_ms:.2f} ms)")
    if args.learned_only:
        log(json.dumps({"ms_per_step": float(np.mean(ms_learned)), "walk_ms": walk_ms, "rounds": rounds,
                        "walk_avg_launch_ms": walk_ms / max(1, rounds), "tree_entries": st_time["tree_entries"],
                        "tree_entries_first": st_time["tree_entries_first"]}))
        ctx.close()
        return
    mismatch = learned.bound_mass_mismatch(pop, warm, ref)
    mism = learned.mismatch_summary(pop, warm, ref)
    if mism["subhalos_within"] < 1.0:
        log(f"PER-SUBHALO GATE FAILED: {mism['subhalos_within']:.5f} of the subhaloes within "
            f"{mism['tolerance']} of the plain iteration's bound mass")
    # the same comparison at BoundMassPrecision 1.0 (iterate until no particle leaves)
    prm1 = hbtx.workload_params(direct_cutoff=args.direct_cutoff, bound_mass_precision=1.0)
    ms_plain1 = timed_steps(ctx, args.steps, args.warmup, lambda: ctx.run(prm1, gflag))
    ref1 = ctx.fetch()

    def learned_step1():
        ctx.warm_proposal(net, prm1, args.threshold, args.warm_min, count=False)
        ctx.run(prm1, hbtx.WARMSTART | gflag)

    ms_learned1 = timed_steps(ctx, args.steps, args.warmup, learned_step1)
    warm1 = ctx.fetch()
    prec1 = {"batched_ms_per_step": float(np.mean(ms_plain1)), "learned_ms_per_step": float(np.mean(ms_learned1)),
             "learned_gain": float(np.mean(ms_plain1) / np.mean(ms_learned1)),
             "value": pop.n / (float(np.mean(ms_learned1)) / 1e3),
             "mismatch": learned.mismatch_summary(pop, warm1, ref1)}
    log(f"BoundMassPrecision 1.0: batched {prec1['batched_ms_per_step']:.2f} ms, learned "
        f"{prec1['learned_ms_per_step']:.2f} ms, mismatch {prec1['mismatch']['summed']:.2e}")
    gate = {"tolerance": 1e-3, "rule": "bound-mass symmetric difference over the plain iteration's bound mass, "
                                       "per subhalo and summed, at BoundMassPrecision 0.995 and 1.0",
            "passed": bool(mism["subhalos_within"] >= 1.0 and mism["summed"] <= 1e-3
                           and prec1["mismatch"]["subhalos_within"] >= 1.0
                           and prec1["mismatch"]["summed"] <= 1e-3)}
    if not gate["passed"]:
        log("PER-SUBHALO GATE FAILED at BoundMassPrecision 1.0 or summed")
    sweep = None
    if args.warm_min_sweep:
        sweep = {}
        for wm in [int(x) for x in args.warm_min_sweep.split(",")]:
            row = {}
            for tag, p, rf, ms_b in (("p0.995", prm, ref, ms_plain_b), ("p1.0", prm1, ref1, ms_plain1)):
                def step_wm():
                    ctx.warm_proposal(net, p, args.threshold, wm, count=False)
                    ctx.run(p, hbtx.WARMSTART | gflag)
                ms_wm = timed_steps(ctx, args.steps, args.warmup, step_wm)
                mm = learned.mismatch_summary(pop, ctx.fetch(), rf)
                row[tag] = {"learned_ms_per_step": float(np.mean(ms_wm)),
                            "learned_gain": float(np.mean(ms_b) / np.mean(ms_wm)),
                            "subhalos_within": mm["subhalos_within"],
                            "bound_mass_in_subhalos_within": mm["bound_mass_in_subhalos_within"],
                            "summed": mm["summed"]}
            sweep[str(wm)] = row
            log(f"warm_min {wm}: " + ", ".join(f"{t} {r['learned_ms_per_step']:.1f} ms gain "
                                                f"{r['learned_gain']:.2f} within {r['subhalos_within']:.5f}"
                                                for t, r in row.items()))
    # work counters: one extra, untimed run of the counting instantiation (same results, deterministic work),
    # from the headline's proposal
    ctx.warm_proposal(net, prm, args.threshold, args.warm_min, count=False)
    ctx.run(prm, hbtx.WARMSTART | hbtx.COUNT)
    st = ctx.stats()
    ctx.run(prm, hbtx.COUNT)
    st_plain = ctx.stats()
    # post-unbinding properties (CalculateProfileProperties + CalculateShape) of that run
    ctx.profiles(prm)
    ctx.sync()
    t0 = time.perf_counter()
    prof = ctx.profiles(prm)
    prof_ms = (time.perf_counter() - t0) * 1e3
    dev_bytes = hbtx.device_bytes()
    st["walk_ms"] = walk_ms
    ms_step = float(np.mean(ms_learned))
    value = pop.n / (ms_step / 1e3)
    batched_value = pop.n / (float(np.mean(ms_plain_b)) / 1e3)
    flops = walk_flops(st)
    achieved = flops / (walk_ms / 1e3) / 1e12 if walk_ms > 0 else 0.0
    res = warm["result"]
    traffic, traffic_src = walk_traffic()
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
        "config": {"workload": "evolved power-law NFW subhalo population (5e7 particles, 8 ID-preserving "
                               f"snapshots), snapshot {args.snapshot} unbound with the learned warm start",
                   "n_particles": int(pop.n), "n_subhalos": int(pop.nsub), "theta": 0.7,
                   "softening": 0.001, "bound_mass_precision": 0.995, "max_sample_size": 0,
                   "direct_cutoff": int(args.direct_cutoff), "seed": args.seed, "snapshot": args.snapshot,
                   "trained_on_snapshots": f"0-{args.snapshot - 1}", "threshold": args.threshold,
                   "warm_min_particles": args.warm_min,
                   "tuning": ctx.tuning(), "round_graphs": not args.no_graph,
                   "rounds": rounds, "bound_particles": int(res["nbound"].sum()),
                   "disrupted": int(res["disrupted"].sum()), "proposed": int(npro)},
        "roofline": {"bound": "valu", "kernel": "k_walk_energy", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_source": traffic_src, "kernel_ms_per_step": walk_ms,
                     "launches_per_step": rounds,
                     "kernel_avg_launch_ms": walk_ms / max(1, rounds), "flop_per_step": flops},
        "bound_mass_mismatch_vs_plain": mismatch,
        "mismatch_detail": mism,
        "per_subhalo_gate": gate,
        "profiles_and_shape": {"ms": prof_ms, "subhalos_with_r200": int(prof["has_r200"].sum()),
                               "note": "host wall time of hbtx_batch_profiles, results copied to the host"},
        "bound_mass_precision_1": prec1,
        "batched_without_learned": {"value": batched_value, "unit": "particles/s",
                                    "ms_per_step": float(np.mean(ms_plain_b)), "rounds": int(st_plain["rounds"]),
                                    "walk_work": st_plain["tree_lane_visits"] + st_plain["direct_lane_pairs"]},
        "learned_gain": value / batched_value,
        "learned": {"inference_ms": inf_ms, "train_seconds": t_train, "label_seconds": t_lab,
                    "training_samples": 5 * args.train_per_snapshot if not args.state else None,
                    "final_loss": loss, "walk_work": st["tree_lane_visits"] + st["direct_lane_pairs"],
                    "mean_iterations": float(res["iterations"].mean()),
                    "mean_iterations_plain": float(ref["result"]["iterations"].mean())},
        "walk_stats": {k: v for k, v in st.items() if not k.startswith("unused")},
        "device_memory": {"bytes": dev_bytes, "bytes_per_particle": dev_bytes / pop.n,
                          "note": "device memory the library's buffers held after the timed runs "
                                  "(batch arrays, tree entry reservation, sort and reduction scratch, network)"},
    }

    if sweep is not None:
        line["warm_min_sweep"] = sweep

    if not args.no_plain:
        secs, plain_out = plain_gpu_full(hbtx, pop, prm)
        plain_rate = pop.n / secs
        same = bool(np.array_equal(plain_out["order"], ref["order"])
                    and np.array_equal(plain_out["result"]["nbound"], ref["result"]["nbound"]))
        line["plain_gpu"] = {"value": plain_rate, "unit": "particles/s", "seconds": secs,
                             "sample": f"all {pop.nsub} subhaloes of snapshot {args.snapshot}, one at a time",
                             "identical_to_batched": same}
        line["speedup_vs_plain"] = value / plain_rate
        log(f"plain GPU loop {secs:.2f}s -> {plain_rate:.3e} particles/s (identical to batched: {same})")

    if not args.no_tree:
        # one tree over a 2e6-particle NFW host with 50 embedded subhaloes, potential at every particle
        host = synth.host_with_subs(seed=1)
        tms = []
        for _ in range(3):
            w0 = time.perf_counter()
            tree = ctx.tree(host.pos, host.mass, prm)
            w1 = time.perf_counter()
            tree.potential(host.pos, host.mass)
            w2 = time.perf_counter()
            tree.close()
            tms.append(((w1 - w0) * 1e3, (w2 - w1) * 1e3))
        build_ms, walk_ms_c1 = np.median([t[0] for t in tms]), np.median([t[1] for t in tms])
        line["tree_potential_host_with_subs"] = {
            "particles": int(host.n), "build_ms": float(build_ms), "potential_ms": float(walk_ms_c1),
            "particles_per_s": host.n / ((build_ms + walk_ms_c1) / 1e3),
            "note": "GravityTree_t Build + EvaluatePotential at every particle, host<->device transfers included"}
        log(f"host+subs tree: build {build_ms:.1f} ms, potential {walk_ms_c1:.1f} ms for {host.n} particles")

    if not args.no_configs:
        line["other_configs"] = other_configs(hbtx, synth, ctx, log)

    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pop, args, log)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
