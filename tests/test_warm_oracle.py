"""Pins the oracle's warm-start restatement (the checker of the GPU warm start)
against a brute-force numpy loop with exact potentials (theta = 0 makes the
oracle's tree a direct sum), and checks that an ineligible proposal leaves the
plain iteration untouched."""
import numpy as np
import pytest

import pyoracle as O
import synth
from test_oracle import EPS, G

MIN_PART = 20


def phi_exact(x, m, targets, sources, selfm):
    h, hinv = 2.8 * EPS, 1.0 / (2.8 * EPS)
    d = x[targets][:, None, :] - x[sources][None, :, :]
    r = np.sqrt((d * d).sum(-1))
    u = r * hinv
    wp = np.where(u < 0.5, -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6)),
                  -3.2 + 0.066666666667 / np.maximum(u, 1e-30)
                  + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u))))
    k = np.where(r >= h, -1.0 / np.maximum(r, 1e-30), hinv * wp)
    return ((k * m[sources].astype(np.float64)[None, :]).sum(1) + selfm / np.float32(EPS)) * G


def frame(x, v, m, idx):
    mb = m[idx].astype(np.float64)
    rp = ((x[idx] * mb[:, None]).sum(0) / mb.sum()).astype(np.float32)
    rv = ((v[idx].astype(np.float64) * mb[:, None]).sum(0) / mb.sum()).astype(np.float32)
    return rp, rv


def brute_force_warm(pop, prec, prop):
    """Warm start rules: the first iteration = proposed sources, every particle a target, in
    the proposal's frame; bound = E < 0, proposed or not; accepted if the bound
    particles from outside the proposal are at most 0.1% of the bound set, at least
    half the proposal is bound and the bound proposed count is at least
    2 MinNumPartOfSub.  Then, when a correction follows and outside particles are
    bound, one more iteration adds their exact potential to every candidate's
    energy (frames unchanged); the correction sources are the proposed, removed
    particles (and any the added potential left unbound), and the convergence test
    runs against every particle.  Rejected or ineligible proposals run the plain loop.
    Returns (bound set, energies, iterations, accepted)."""
    x = pop.pos.astype(np.float64)
    v = pop.vel.astype(np.float32)
    m = pop.mass.astype(np.float32)
    n = len(x)
    E = np.zeros(n, np.float32)
    pidx = np.where(prop)[0]
    iters = 0
    accepted = False
    corr = False
    old = None
    removed = np.array([], int)
    if len(pidx) >= 2 * MIN_PART:
        iters = 1
        rp, rv = frame(x, v, m, pidx)
        allp = np.arange(n)
        dv = v - rv
        k2 = (dv * dv).sum(1).astype(np.float32)
        E[:] = (k2.astype(np.float64) * 0.5 + phi_exact(x, m, allp, pidx, np.where(prop, m, 0))).astype(np.float32)
        bound = allp[E < 0]
        outside = int(((E < 0) & ~prop).sum())
        nin = len(bound) - outside
        if (nin >= 2 * MIN_PART and 2 * nin >= len(pidx) and outside * 1000 <= len(bound)):
            accepted = True
            removed = pidx[E[pidx] >= 0]
            if len(removed) < len(bound):
                corr = True
                old = (rp.copy(), rv.copy())
            rp, rv = frame(x, v, m, bound)
            if corr and outside > 0:
                iters += 1
                ext = bound[~prop[bound]]
                E[bound] = (E[bound].astype(np.float64)
                            + phi_exact(x, m, bound, ext, np.where(prop[bound], 0, m[bound]))).astype(np.float32)
                left = bound[E[bound] >= 0]
                if len(left):
                    removed = np.concatenate([removed, left])
                    bound = bound[E[bound] < 0]
                    rp, rv = frame(x, v, m, bound)
            cand = bound
            if not (corr and outside > 0) and np.float32(len(bound)) >= np.float32(n) * np.float32(prec):
                return set(bound.tolist()), E, iters, accepted
    if not accepted:
        rp, rv = pop.ref_pos[0].copy(), pop.ref_vel[0].copy()
        cand = np.arange(n)
        E[:] = 0
    while True:
        iters += 1
        nlast = len(cand)
        if corr:
            dV = (old[1] - rv).astype(np.float32)
            dK = np.float32(0.5 * np.float32((dV * dV).sum()))
            ov = v[cand] - old[1]
            E[cand] = (E[cand].astype(np.float64) + ((ov * dV).sum(1).astype(np.float32) + dK)
                       - phi_exact(x, m, cand, removed, 0.0)).astype(np.float32)
        else:
            dv = v[cand] - rv
            k2 = (dv * dv).sum(1).astype(np.float32)
            E[cand] = (k2.astype(np.float64) * 0.5 + phi_exact(x, m, cand, cand, m[cand])).astype(np.float32)
        bound = cand[E[cand] < 0]
        removed = cand[E[cand] >= 0]
        if len(bound) < MIN_PART:
            return None
        if len(removed) < len(bound):  # the correction loop, once entered, stays on
            corr = True
            old = (rp.copy(), rv.copy())
        rp, rv = frame(x, v, m, bound)
        cand = bound
        if np.float32(len(bound)) >= np.float32(nlast) * np.float32(prec):
            return set(bound.tolist()), E, iters, accepted


@pytest.fixture(scope="module")
def small_halo():
    pop = synth.single_nfw(6, n_halo=300, n_inter=60)
    plain = O.unbind_batch(pop, O.default_params(theta=0.0), nthreads=1)
    nb = int(plain["result"][0]["nbound"])
    bound = np.zeros(pop.n, bool)
    bound[plain["order"][:nb]] = True
    return pop, bound


def proposals(bound, rng):
    unb = np.where(~bound)[0]
    fp = bound.copy()
    fp[rng.choice(unb, size=min(len(unb), 25), replace=False)] = True
    fn = bound.copy()
    fn[np.where(bound)[0][:3]] = False
    return {"exact": bound, "false_positives": fp, "false_negatives": fn,
            "everything": np.ones_like(bound)}


@pytest.mark.parametrize("prec", [0.995, 1.0])
@pytest.mark.parametrize("kind", ["exact", "false_positives", "false_negatives", "everything"])
def test_warm_oracle_matches_brute_force(small_halo, prec, kind):
    pop, bound = small_halo
    prop = proposals(bound, np.random.default_rng(1))[kind]
    r = O.unbind_batch(pop, O.default_params(theta=0.0, bound_mass_precision=prec), nthreads=1,
                       proposal=prop.astype(np.uint8))
    bf = brute_force_warm(pop, prec, prop)
    assert bf is not None
    ref, E, iters, accepted = bf
    if kind == "false_negatives":
        assert not accepted  # bound particles outside a small proposal: restart
    if kind == "exact":
        assert accepted
    nb = int(r["result"][0]["nbound"])
    assert set(r["order"][:nb].tolist()) == ref
    assert int(r["result"][0]["iterations"]) == iters
    np.testing.assert_allclose(r["energy"][:nb], E[r["order"][:nb]], rtol=2e-5, atol=1e-3)


def test_ineligible_proposal_is_the_plain_iteration():
    pop = synth.population(seed=4, nsub=30, nmax=3000)
    plain = O.unbind_batch(pop, O.default_params(), nthreads=1)
    for prop in (np.zeros(pop.n, np.uint8), (np.arange(pop.n) % 97 == 0).astype(np.uint8)):
        warm = O.unbind_batch(pop, O.default_params(), nthreads=1, proposal=prop)
        for s in range(pop.nsub):
            a, b = pop.offsets[s], pop.offsets[s + 1]
            if prop[a:b].sum() >= 2 * MIN_PART:
                continue
            assert np.array_equal(warm["order"][a:b], plain["order"][a:b])
            assert np.array_equal(warm["energy"][a:b], plain["energy"][a:b])
            assert warm["result"][s] == plain["result"][s]


@pytest.fixture(scope="module")
def larger_halo():
    pop = synth.single_nfw(7, n_halo=1500, n_inter=300)
    plain = O.unbind_batch(pop, O.default_params(theta=0.0), nthreads=1)
    nb = int(plain["result"][0]["nbound"])
    bound = np.zeros(pop.n, bool)
    bound[plain["order"][:nb]] = True
    return pop, bound, plain["order"][:nb]


@pytest.mark.parametrize("prec", [0.995, 1.0])
def test_warm_oracle_readmits_bound_particles_outside_the_proposal(larger_halo, prec):
    # one bound particle left out of the proposal (0.07% of the bound set): the
    # warm round keeps it, its potential is added in one more iteration, and the
    # result matches the brute-force rules (the particle stays bound)
    pop, bound, order = larger_halo
    assert len(order) >= 1000
    prop = bound.copy()
    missing = int(order[len(order) // 3])
    prop[missing] = False
    r = O.unbind_batch(pop, O.default_params(theta=0.0, bound_mass_precision=prec), nthreads=1,
                       proposal=prop.astype(np.uint8))
    ref, E, iters, accepted = brute_force_warm(pop, prec, prop)
    assert accepted
    assert missing in ref
    nb = int(r["result"][0]["nbound"])
    assert set(r["order"][:nb].tolist()) == ref
    assert int(r["result"][0]["iterations"]) == iters
    np.testing.assert_allclose(r["energy"][:nb], E[r["order"][:nb]], rtol=2e-5, atol=1e-3)
    # the same with the exact proposal: one iteration fewer (no outside particle to add)
    r2 = O.unbind_batch(pop, O.default_params(theta=0.0, bound_mass_precision=prec), nthreads=1,
                        proposal=bound.astype(np.uint8))
    ref2, _, iters2, _ = brute_force_warm(pop, prec, bound)
    assert int(r2["result"][0]["iterations"]) == iters2 == iters - 1
    assert set(r2["order"][:int(r2["result"][0]["nbound"])].tolist()) == ref2
