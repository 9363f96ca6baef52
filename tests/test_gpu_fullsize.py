"""The unbinding at full size, checked through properties that do not depend on
size: graph replay and direct launches agree bit for bit, every subhalo's output
is a permutation of its particles, the bound part is ascending in energy with
every energy negative, every removed particle was unbound when it was removed,
and the per-subhalo sums match the bound sets.  Workloads: the power-law
population drawn to 5x10^7 particles; the 5,000-subhalo hydro mixture with
thermal energy (per-type sums, tracers); and the learned step on the evolved
5x10^7 population, whose bound sets the plain iteration judges per subhalo."""
import numpy as np
import pytest

import hbtx
import learned
import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    return synth.population(seed=2, target_particles=50_000_000)


def check_invariants(pop, a, thermal=False):
    """Size-independent properties of one unbinding output (see the module docstring)."""
    res = a["result"]
    off = pop.offsets
    counts = np.diff(off)
    sub = np.repeat(np.arange(pop.nsub), counts)
    local = np.arange(pop.n) - off[sub]
    # a permutation of every subhalo's particles
    g = off[sub] + a["order"]
    assert np.all(a["order"] >= 0) and np.all(a["order"] < counts[sub])
    assert np.all(np.bincount(g, minlength=pop.n) == 1)
    alive = res["disrupted"][sub] == 0
    nb = res["nbound"][sub]
    bound = alive & (local < nb)
    e = a["energy"]
    # bound part: negative energies in ascending order
    assert np.all(e[bound] < 0)
    same = bound[1:] & bound[:-1] & (sub[1:] == sub[:-1])
    assert np.all(e[1:][same] >= e[:-1][same])
    # removed particles were unbound (E >= 0) in the round that removed them
    assert np.all(e[alive & ~bound] >= 0)
    # the bound mass is the mass of the bound set
    m = pop.mass.astype(np.float64)[g]
    mb = np.bincount(sub[bound], weights=m[bound], minlength=pop.nsub)
    ok = res["disrupted"] == 0
    np.testing.assert_allclose(res["mbound"][ok], mb[ok], rtol=1e-5)
    return sub, g, bound, ok


def test_fullsize_invariants(ctx, big):
    pop = big
    assert pop.n >= 50_000_000
    prm = hbtx.workload_params(direct_cutoff=-1)
    a = ctx.unbind(pop, prm)
    b = ctx.unbind(pop, prm, hbtx.NO_GRAPH)
    assert np.array_equal(a["order"], b["order"])
    assert np.array_equal(a["energy"], b["energy"])
    assert np.array_equal(a["result"], b["result"])
    _, _, _, ok = check_invariants(pop, a)
    assert a["result"]["nbound"][ok].sum() > 0.3 * pop.n


def test_fullsize_hydro_mixture(ctx):
    # the 5,000-subhalo hydro mixture: 70% DM, 20% gas with thermal energy, 9% stars,
    # 1% black holes at 100x mass
    pop = synth.population(seed=3, nsub=5000, hydro=True)
    prm = hbtx.workload_params(direct_cutoff=-1, dm_only=0, thermal=1)
    a = ctx.unbind(pop, prm, hbtx.SAVE_POTENTIAL)
    b = ctx.unbind(pop, prm, hbtx.NO_GRAPH | hbtx.SAVE_POTENTIAL)
    assert np.array_equal(a["order"], b["order"]) and np.array_equal(a["energy"], b["energy"])
    assert np.array_equal(a["result"], b["result"]) and np.array_equal(a["potential"], b["potential"])
    sub, g, bound, ok = check_invariants(pop, a)
    res = a["result"]
    t = pop.ptype[g]
    m = pop.mass.astype(np.float64)[g]
    for ty in range(6):
        sel = bound & (t == ty)
        nbt = np.bincount(sub[sel], minlength=pop.nsub)
        mbt = np.bincount(sub[sel], weights=m[sel], minlength=pop.nsub)
        assert np.array_equal(res["nbound_type"][ok, ty], nbt[ok]), ty
        np.testing.assert_allclose(res["mbound_type"][ok, ty], mbt[ok], rtol=1e-5, atol=1e-7)
    assert np.array_equal(res["nbound_type"][ok].sum(1), res["nbound"][ok])
    # TracerIndex: the first bound tracer (types 1 and 4) in the final order
    tracer = np.isin(t, (1, 4))
    for s in np.nonzero(ok)[0][:2000]:
        lo = pop.offsets[s]
        first = np.nonzero(tracer[lo:lo + res["nbound"][s]])[0]
        assert res["tracer_index"][s] == (first[0] if len(first) else 0), s
    assert res["nbound_type"][ok, 0].sum() > 0 and res["nbound_type"][ok, 5].sum() > 0
    # the thermal term: a bound particle's binding energy is its potential energy
    # (ParticlePotentialEnergies, src/subhalo_unbind.cpp:114-140) plus its kinetic
    # energy in the final frame plus u (zero for every type but gas)
    fv = res["com_vel"][sub].astype(np.float64)
    kin = 0.5 * ((pop.vel[g].astype(np.float64) - fv) ** 2).sum(1)
    u = pop.u[g].astype(np.float64)
    lhs = a["energy"][bound].astype(np.float64)
    rhs = a["potential"][bound] + kin[bound] + u[bound]
    scale = np.abs(a["potential"][bound]).astype(np.float64)
    assert np.quantile(np.abs(lhs - rhs) / scale, 0.999) < 1e-4
    gas = bound & (t == 0)
    assert np.all(u[gas] > 0) and np.all(u[bound & (t != 0)] == 0)


def test_fullsize_learned_step(ctx):
    """The headline's learned step: the evolved 5e7 population, trained on the plain
    unbinding of snapshots 0-5, snapshot 6 warm-started from 30,000 particles; every
    subhalo within 1e-3 of the plain iteration's bound mass, and the invariants."""
    pop0 = synth.population(seed=2, target_particles=50_000_000)
    prm = hbtx.workload_params(direct_cutoff=-1)
    rng = np.random.default_rng(0)
    Xs, ys, prev, pop = [], [], None, None
    for k, snap in enumerate(synth.evolve_iter(pop0, nsnap=7, seed=4)):
        pop = snap
        if k == 6:
            break
        if prev is not None:
            X = learned.snapshot_features(ctx, pop, prev, prm)
        out = ctx.unbind(pop, prm)
        if prev is not None:
            take = rng.choice(pop.n, size=2_000_000, replace=False)
            Xs.append(X[take])
            ys.append(learned.labels(pop, out)[take])
            del X
        prev = learned.previous_state(pop, out)
    del pop0
    net, loss = learned.train(ctx, np.concatenate(Xs), np.concatenate(ys), epochs=4)
    del Xs, ys
    warm, _, _, npro = learned.warm_unbind(ctx, pop, prev, net, prm, min_particles=30000)
    plain = ctx.unbind(pop, prm)
    per = learned.mismatch_per_subhalo(pop, warm, plain)
    print(f"learned step: {npro} proposed, per-subhalo mismatch max {per.max():.2e}, "
          f"summed {learned.bound_mass_mismatch(pop, warm, plain):.2e}")
    assert npro > 0
    assert np.all(per <= 1e-3), (np.nonzero(per > 1e-3)[0][:10], per.max())
    assert np.array_equal(warm["result"]["disrupted"], plain["result"]["disrupted"])
    check_invariants(pop, warm)
