"""Autotuner of the batched unbinding on the local GPU.

Picks, by timing complete unbindings:
  * direct_cutoff: the source count at or below which a subhalo's potentials
    come from the exact O(N^2) sum rather than the tree walk.  Populations of
    equal-size subhaloes are unbound with every subhalo on one path, then on the
    other; the crossover is the geometric midpoint between the largest size where
    the direct sum wins and the next size.  The cutoff is then the fastest of
    the crossover and its neighbours at factors of sqrt(2) on a power-law
    population, where both paths share the GPU;
  * walk_chunks: how many contiguous chunks the walk's item list is cut into
    for its XCD-aware work distribution, timed on a power-law population.

Writes hbtx_tune.json beside the library (or --out), which every new context
reads (hbtx_create; HBTX_TUNE_FILE points elsewhere).

    python hbt-herons_amd/autotune.py [--particles 4e6] [--out PATH]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import hbtx  # noqa: E402
import synth  # noqa: E402

SIZES = (64, 128, 256, 512, 1024, 2048, 4096, 8192)
CHUNKS = (1, 4, 8, 16)


def best_ms(ctx, prm, flags=hbtx.BATCHED, reps=3):
    ctx.run(prm, flags)  # warm-up (allocations, caches)
    return min(ctx.run(prm, flags) for _ in range(reps))


def tune_cutoff(ctx, particles, log):
    rows = []
    for n in SIZES:
        nsub = max(16, int(particles // (1.1 * n)))
        pop = synth.population(seed=100 + n, nsub=nsub, nmin=n, nmax=n)
        ctx.load(pop)
        t_direct = best_ms(ctx, hbtx.workload_params(direct_cutoff=1 << 40))
        t_tree = best_ms(ctx, hbtx.workload_params(direct_cutoff=0))
        rows.append(dict(size=n, subhalos=nsub, direct_ms=t_direct, tree_ms=t_tree))
        log(f"  {n:6d} sources x {nsub:6d}: direct {t_direct:8.2f} ms  tree {t_tree:8.2f} ms")
    wins = [r["size"] for r in rows if r["direct_ms"] <= r["tree_ms"]]
    if not wins:
        cutoff = 0
    elif wins[-1] == SIZES[-1]:
        cutoff = SIZES[-1]
    else:
        nxt = SIZES[SIZES.index(wins[-1]) + 1]
        cutoff = int(round(math.sqrt(wins[-1] * nxt)))
    return cutoff, rows


def refine_cutoff(ctx, crossover, particles, log):
    """The crossover and its neighbours timed on a power-law population."""
    pop = synth.population(seed=7, target_particles=int(particles))
    ctx.load(pop)
    cands = sorted({max(0, int(round(crossover * 2 ** (k / 2)))) for k in (-2, -1, 0, 1, 2)})
    rows = []
    for cut in cands:
        ms = best_ms(ctx, hbtx.workload_params(direct_cutoff=cut))
        rows.append(dict(direct_cutoff=cut, ms=ms))
        log(f"  power-law population, cutoff {cut:6d}: {ms:8.2f} ms")
    return min(rows, key=lambda r: r["ms"])["direct_cutoff"], rows


def tune_chunks(ctx, cutoff, particles, log):
    pop = synth.population(seed=7, target_particles=int(particles))
    ctx.load(pop)
    prm = hbtx.workload_params(direct_cutoff=cutoff)
    rows = []
    for ch in CHUNKS:
        ctx.set_tuning(cutoff, ch)
        ms = best_ms(ctx, prm)
        rows.append(dict(walk_chunks=ch, ms=ms))
        log(f"  walk chunks {ch:3d}: {ms:8.2f} ms")
    best = min(rows, key=lambda r: r["ms"])["walk_chunks"]
    return best, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=float, default=4e6, help="particles per timed population")
    ap.add_argument("--out", default=os.path.join(HERE, "lib", "hbtx_tune.json"))
    args = ap.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = hbtx.Context(0)
    t0 = time.time()
    log("direct-sum / tree cutoff:")
    crossover, rows_c = tune_cutoff(ctx, args.particles, log)
    log(f"-> crossover {crossover}")
    cutoff, rows_m = refine_cutoff(ctx, crossover, args.particles, log)
    log(f"-> direct_cutoff {cutoff}")
    log("walk item chunks:")
    chunks, rows_w = tune_chunks(ctx, cutoff, args.particles, log)
    log(f"-> walk_chunks {chunks}")
    out = dict(direct_cutoff=cutoff, walk_chunks=chunks, seconds=round(time.time() - t0, 1),
               crossover=crossover, cutoff_timings=rows_c, mixture_timings=rows_m, chunk_timings=rows_w)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(dict(direct_cutoff=cutoff, walk_chunks=chunks, file=args.out)))
    ctx.close()


if __name__ == "__main__":
    main()
