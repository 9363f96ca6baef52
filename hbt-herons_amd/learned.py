"""Learned warm start of Subhalo_t::Unbind (src/subhalo_unbind.cpp:463-808).

Pipeline over an ID-preserving snapshot sequence (same slot = same particle):
  * labels of snapshot k = bound flags of the plain unbinding of snapshot k;
  * features of snapshot k = current phase space + snapshot k-1's outputs
    (binding-energy rank, bound flag, energy, frame, bound mass, bound rms radius),
    computed on the device (hbtx_batch_features);
  * the 8-64-64-1 classifier is trained on the device (hbtx_mlp_train) and its
    proposal starts the unbinding (HBTX_WARMSTART); the exact potential of the
    proposed set re-verifies every particle, and a rejected proposal restarts the
    plain iteration.
"""
import numpy as np

import hbtx


def previous_state(pop, out, box=None):
    """Per-slot arrays of one snapshot's unbinding outputs, for the next snapshot's features."""
    n, nsub = pop.n, pop.nsub
    off = np.asarray(pop.offsets, dtype=np.int64)
    counts = np.diff(off)
    base = np.repeat(off[:-1], counts)
    local = np.arange(n, dtype=np.int64) - base
    slot = base + out["order"].astype(np.int64)
    res = out["result"]
    nb = res["nbound"].astype(np.int64)
    rank = np.empty(n, np.int32)
    rank[slot] = local
    bound = np.empty(n, np.uint8)
    bound[slot] = (local < np.repeat(nb, counts)).astype(np.uint8)
    energy = np.empty(n, np.float32)
    energy[slot] = out["energy"]
    com_pos = res["com_pos"].astype(np.float32)
    com_vel = res["com_vel"].astype(np.float32)
    # rms radius of the bound set about its centre of mass (all particles when none is bound)
    d = pop.pos.astype(np.float64) - np.repeat(com_pos.astype(np.float64), counts, axis=0)
    if box:
        d = d - box * np.round(d / box)
    r2 = (d * d).sum(1)
    w = bound.astype(np.float64)
    none = np.repeat(nb == 0, counts)
    w[none] = 1.0
    nz = counts > 0
    sw = np.zeros(nsub)
    sr = np.zeros(nsub)
    sw[nz] = np.add.reduceat(w, off[:-1][nz])
    sr[nz] = np.add.reduceat(w * r2, off[:-1][nz])
    rscale = np.sqrt(sr / np.maximum(sw, 1.0)).astype(np.float32)
    mbound = res["mbound"].astype(np.float32).copy()
    mtot = np.zeros(nsub)
    mtot[nz] = np.add.reduceat(pop.mass.astype(np.float64), off[:-1][nz])
    mbound[nb == 0] = mtot[nb == 0]
    return dict(rank=rank, bound=bound, energy=energy, com_pos=com_pos, com_vel=com_vel, mbound=mbound,
                rscale=rscale)


def labels(pop, out):
    """Per-slot bound flags of an unbinding output."""
    return previous_state(pop, out)["bound"]


def bound_mass_mismatch(pop, out, ref):
    """Mass of the symmetric difference of the bound sets over the bound mass of ref."""
    a, b = labels(pop, out).astype(bool), labels(pop, ref).astype(bool)
    m = pop.mass.astype(np.float64)
    return float(m[a ^ b].sum() / max(m[b].sum(), 1e-300))


def mismatch_per_subhalo(pop, out, ref):
    """Per subhalo: mass of the symmetric difference of the bound sets over the
    bound mass of ref (0 when both are empty, inf when only ref is empty)."""
    a, b = labels(pop, out).astype(bool), labels(pop, ref).astype(bool)
    m = pop.mass.astype(np.float64)
    off = np.asarray(pop.offsets, dtype=np.int64)
    sub = np.repeat(np.arange(pop.nsub), np.diff(off))
    diff = np.bincount(sub, weights=m * (a ^ b), minlength=pop.nsub)
    mref = np.bincount(sub, weights=m * b, minlength=pop.nsub)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(mref > 0, diff / np.where(mref > 0, mref, 1.0), np.where(diff > 0, np.inf, 0.0))
    return r


def mismatch_summary(pop, out, ref, tol=1e-3):
    """Summed mismatch and the per-subhalo distribution against tol."""
    per = mismatch_per_subhalo(pop, out, ref)
    b = labels(pop, ref).astype(bool)
    off = np.asarray(pop.offsets, dtype=np.int64)
    sub = np.repeat(np.arange(pop.nsub), np.diff(off))
    mref = np.bincount(sub, weights=pop.mass.astype(np.float64) * b, minlength=pop.nsub)
    ok = per <= tol
    return {"summed": bound_mass_mismatch(pop, out, ref), "tolerance": tol,
            "subhalos_within": float(ok.mean()),
            "bound_mass_in_subhalos_within": float(mref[ok].sum() / max(mref.sum(), 1e-300)),
            "per_subhalo_median": float(np.median(per)),
            "per_subhalo_p99": float(np.quantile(per[np.isfinite(per)], 0.99)) if np.isfinite(per).any() else None,
            "subhalos_disagreeing_on_disruption": int(((per == np.inf) | ((mref > 0) & (per >= 1.0))).sum())}


def snapshot_features(ctx, pop, prev, params):
    ctx.load(pop)
    ctx.set_previous(prev)
    return ctx.features(params)


def collect_training_set(ctx, snaps, params, train_snaps=range(6), per_snapshot=2_000_000, seed=0,
                         flags=hbtx.BATCHED):
    """Plain unbinding of the training snapshots for labels; features of snapshot k
    from snapshot k-1's outputs (snapshot 0 only provides the previous state).
    Returns (X, y, outputs of the last training snapshot)."""
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    prev = None
    last = None
    for k in train_snaps:
        pop = snaps[k]
        if prev is not None:
            X = snapshot_features(ctx, pop, prev, params)
        out = ctx.unbind(pop, params, flags)
        lab = labels(pop, out)
        if prev is not None:
            take = rng.choice(pop.n, size=min(per_snapshot, pop.n), replace=False)
            Xs.append(X[take])
            ys.append(lab[take])
        prev = previous_state(pop, out)
        last = out
    return np.concatenate(Xs), np.concatenate(ys), last


def train(ctx, X, y, seed=0, epochs=4, batch=8192, lr=3e-3):
    net = hbtx.MLP(ctx, seed)
    loss = net.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, loss


def warm_unbind(ctx, pop, prev, net, params, threshold=0.1, flags=hbtx.BATCHED, min_particles=0):
    """Features -> proposal -> verified warm-started unbinding of one snapshot.
    Returns (output, run milliseconds, inference milliseconds, proposed count)."""
    ctx.load(pop)
    ctx.set_previous(prev)
    npro, inf_ms = ctx.warm_proposal(net, params, threshold, min_particles)
    ms = ctx.run(params, flags | hbtx.WARMSTART)
    out = ctx.fetch()
    return out, ms, inf_ms, npro
