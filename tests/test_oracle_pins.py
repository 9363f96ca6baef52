"""Pins for the oracle functions that the unbinding tests surface only through
their results: the subsampling mass upscale (src/subhalo_unbind.cpp:290-391),
RefineBindingEnergyOrder (:395-425), AverageKinematics (:200-249), the Hubble
term of Snapshot_t::RelativeVelocity (src/snapshot.h:135-148) and
Parameter_t::GetCurrentSoftening (src/config_parser.cpp:641-649).

Each expectation comes from what the mathematics fixes (conservation of mass,
closed-form potentials and kinetic terms of a rigidly rotating ring, a brute-force
exact sort, the potential at zero separation), not from the oracle's own code."""
import numpy as np
import pytest

import pyoracle as O
from test_warm_oracle import phi_exact
import synth

G = 43.0071


# ---------------------------------------------------------------------------
# mass upscale under subsampling
# ---------------------------------------------------------------------------
def _typed_masses(rng, n, types=(0, 1, 4)):
    t = rng.choice(np.array(types), size=n).astype(np.int32)
    m = (rng.uniform(0.5, 2.0, n) * np.where(t == 0, 0.2, 1.0)).astype(np.float32) * 0.01
    return m, t


@pytest.mark.parametrize("nuns", [0, 7])
def test_upscale_global_conserves_mlast(nuns):
    # GetMassUpscaleFactor: one factor f for every type such that the unsampled
    # particles (true mass) plus f x the sample's mass equal Mlast
    rng = np.random.default_rng(1)
    n, mss = 5000, 300
    m = rng.uniform(0.5, 2.0, n).astype(np.float32) * 0.01
    mlast = np.float32(m.astype(np.float64).sum())
    for prm in (dict(dm_only=1), dict(dm_only=0, upscale_per_type=0)):
        ns, f = O.mass_upscale(m, None, n, mlast, mss, nuns, params=prm)
        assert ns == mss + nuns
        assert np.all(f == f[0])
        total = m[:nuns].astype(np.float64).sum() + f[0] * m[nuns:nuns + mss].astype(np.float64).sum()
        assert total == pytest.approx(float(mlast), rel=3e-6)
        # the factor is not the plain count ratio: masses vary
        assert f[0] != pytest.approx(n / mss, rel=1e-3)


def test_upscale_per_type_conserves_each_type():
    # GetMassUpscaleFactorPerParticleType: each type's sampled mass scales to that
    # type's mass over [Nunsample, Nlast); the front run of do-not-subsample
    # particles keeps its true mass and is not part of the sums
    rng = np.random.default_rng(2)
    n, mss, nuns = 6000, 400, 5
    m, t = _typed_masses(rng, n)
    t[:nuns] = 5
    m[:nuns] = 1.0
    ns, f = O.mass_upscale(m, t, n, 123.0, mss, nuns, params=dict(dm_only=0, upscale_per_type=1))
    assert ns == mss + nuns
    md = m.astype(np.float64)
    samp = slice(nuns, nuns + mss)
    for ty in (0, 1, 4):
        true = md[nuns:][t[nuns:] == ty].sum()
        scaled = f[ty] * md[samp][t[samp] == ty].sum()
        assert scaled == pytest.approx(true, rel=3e-6), ty
    assert f[5] == 0.0 and f[2] == 0.0  # absent from the sample
    # with every type present in the sample, Mlast plays no part
    ns2, f2 = O.mass_upscale(m, t, n, 999.0, mss, nuns, params=dict(dm_only=0, upscale_per_type=1))
    assert np.array_equal(f, f2)


def test_upscale_per_type_missing_type_conserves_total():
    # a type absent from the sample: the others scale up so that the total true
    # mass over [Nunsample, Nlast) is conserved
    rng = np.random.default_rng(3)
    n, mss = 4000, 200
    m, t = _typed_masses(rng, n, types=(1, 4))
    t[mss + 100: mss + 160] = 0  # gas only outside the sample
    ns, f = O.mass_upscale(m, t, n, 0.0, mss, 0, params=dict(dm_only=0, upscale_per_type=1))
    md = m.astype(np.float64)
    assert f[0] == 0.0
    scaled = sum(f[ty] * md[:mss][t[:mss] == ty].sum() for ty in (1, 4))
    assert scaled == pytest.approx(md.sum(), rel=3e-6)
    # the two present types keep the ratio of their own upscale factors
    r1 = f[1] / (md[t == 1].sum() / md[:mss][t[:mss] == 1].sum())
    r4 = f[4] / (md[t == 4].sum() / md[:mss][t[:mss] == 4].sum())
    assert r1 == pytest.approx(r4, rel=3e-6) and r1 > 1.0


def test_upscale_inactive_below_threshold():
    m = np.full(100, 0.01, np.float32)
    for nlast, mss, nuns in ((100, 0, 0), (100, 100, 1), (99, 90, 10)):
        ns, f = O.mass_upscale(m, None, nlast, 1.0, mss, nuns)
        assert ns == nlast and np.all(f == 1.0)
    ns, f = O.mass_upscale(m, None, 100, np.float32(1.0), 90, 10)
    assert ns == 100  # Nlast == MaxSampleSize + Nunsample: the sample is everything


# ---------------------------------------------------------------------------
# RefineBindingEnergyOrder
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mss", [100, 300])
def test_refine_order_is_the_exact_self_energy_sort(mss):
    """The first max(MaxSampleSize, 0.1 Nbound) particles are reordered by their
    binding energy to that set alone (true masses, final frame); the set itself
    and everything after it are those of the unrefined run.  Brute force: exact
    pair sums (theta = 0 makes the oracle's tree exact) sorted in numpy."""
    pop = synth.single_nfw(11, n_halo=1500, n_inter=300)
    rng = np.random.default_rng(5)
    perm = rng.permutation(pop.n)  # the caller's shuffle (:430-443)
    # unequal masses, so that the self term (true mass / eps) differs from particle to particle
    mass = (pop.mass * rng.uniform(0.3, 3.0, pop.n)).astype(np.float32)
    pop = synth.Population(pop.pos[perm], pop.vel[perm], mass, pop.offsets, pop.pid[perm], meta=pop.meta)
    base = dict(theta=0.0, max_sample_size=mss, bound_fraction_center_refinement=0.1)
    ref = O.unbind_batch(pop, O.default_params(refine_most_bound=1, **base), nthreads=1)
    raw = O.unbind_batch(pop, O.default_params(refine_most_bound=0, **base), nthreads=1)
    r = ref["result"][0]
    nb = int(r["nbound"])
    assert nb == int(raw["result"][0]["nbound"]) and nb > mss
    size = max(mss, int(0.1 * nb))
    top = ref["order"][:size]
    assert set(top.tolist()) == set(raw["order"][:size].tolist())
    assert np.array_equal(ref["order"][size:], raw["order"][size:])
    assert not np.array_equal(top, raw["order"][:size])  # the refinement did reorder
    x = pop.pos.astype(np.float64)
    self_m = pop.mass[top]
    phi = phi_exact(x, pop.mass, top, top, self_m)
    dv = pop.vel[top].astype(np.float32) - r["com_vel"].astype(np.float32)
    e = 0.5 * (dv.astype(np.float32) ** 2).sum(1).astype(np.float64) + phi
    # the refined order sorts these energies (ties within rounding allowed)
    assert np.all(np.diff(e) >= -1e-9 * np.abs(e[:-1])), np.min(np.diff(e))


# ---------------------------------------------------------------------------
# A rigidly rotating ring: closed forms for the energies, AverageKinematics and
# the Hubble term
# ---------------------------------------------------------------------------
RING_N, RING_R, RING_M = 64, 0.05, 0.01


def ring(omega, centre=(1.0, 1.0, 1.0), box=None):
    k = np.arange(RING_N)
    ang = 2 * np.pi * k / RING_N
    dx = np.stack([RING_R * np.cos(ang), RING_R * np.sin(ang), np.zeros(RING_N)], 1)
    pos = dx + np.asarray(centre)
    if box:
        pos = np.mod(pos, box)
    vel = omega * np.stack([-dx[:, 1], dx[:, 0], np.zeros(RING_N)], 1)  # omega z x dx
    pos, vel = pos.astype(np.float32), vel.astype(np.float32)
    mass = np.full(RING_N, RING_M, np.float32)
    pop = synth.Population(pos, vel, mass, np.array([0, RING_N], np.int64), np.arange(RING_N, dtype=np.int64))
    pop.ref_pos = np.asarray([centre], np.float32)
    pop.ref_vel = np.zeros((1, 3), np.float32)
    return pop


def ring_phi(a):
    # every particle sees the same sum over chords 2 R sin(pi k / N) (all beyond h here)
    k = np.arange(1, RING_N)
    chords = 2 * RING_R * np.sin(np.pi * k / RING_N)
    assert chords.min() > 2.8 * 0.001
    return -G / a * RING_M * np.sum(1.0 / chords)


@pytest.mark.parametrize("a,hz,omega", [(1.0, 0.0, 300.0), (0.5, 200.0, 300.0), (0.5, 200.0, 0.0),
                                        (0.25, 500.0, -150.0)])
def test_ring_energies_and_average_kinematics(a, hz, omega):
    pop = ring(omega)
    res = O.unbind_batch(pop, O.default_params(theta=0.0, scale_factor=a, hz=hz, bound_mass_precision=1.0),
                         nthreads=1)
    r = res["result"][0]
    assert int(r["nbound"]) == RING_N and not r["disrupted"]
    phi0 = ring_phi(a)
    # Snapshot_t::RelativeVelocity: v - V + H a (x - X); |omega x dx|^2 + (H a R)^2
    kin = 0.5 * RING_R ** 2 * (omega ** 2 + (hz * a) ** 2)
    assert kin + phi0 < 0
    np.testing.assert_allclose(res["energy"], kin + phi0, rtol=2e-5)
    # AverageKinematics: <E> - <K> = phi0, <K> = kin, <dx_phys x dv> = a omega R^2 z
    assert float(r["spec_pot"]) == pytest.approx(phi0, rel=2e-5)
    assert float(r["spec_kin"]) == pytest.approx(kin, rel=2e-5, abs=1e-3)
    am = np.asarray(r["spec_am"], np.float64)
    assert am[2] == pytest.approx(a * omega * RING_R ** 2, rel=2e-5, abs=1e-5)
    assert abs(am[0]) < 1e-5 and abs(am[1]) < 1e-5


def test_pure_hubble_flow_across_the_periodic_boundary():
    # no peculiar velocity: the kinetic term is (H a R)^2 / 2 for every particle,
    # which needs the periodic NEAREST offsets for the ring straddling the box corner
    a, hz, box = 0.5, 300.0, 10.0
    pop = ring(0.0, centre=(0.0, 0.0, 0.0), box=box)
    assert pop.pos.max() > box - 1.0 and pop.pos.min() < 1.0
    pop.ref_pos = np.zeros((1, 3), np.float32)
    prm = O.default_params(theta=0.0, scale_factor=a, hz=hz, periodic=1, box_size=box, bound_mass_precision=1.0)
    res = O.unbind_batch(pop, prm, nthreads=1)
    r = res["result"][0]
    assert int(r["nbound"]) == RING_N
    kin = 0.5 * (hz * a * RING_R) ** 2
    np.testing.assert_allclose(res["energy"], kin + ring_phi(a), rtol=2e-5)
    assert float(r["spec_kin"]) == pytest.approx(kin, rel=2e-5)
    assert float(r["spec_pot"]) == pytest.approx(ring_phi(a), rel=2e-5)
    c = np.asarray(r["com_pos"], np.float64)
    assert np.all(np.minimum(np.abs(c), np.abs(box - c)) < 1e-5)


# ---------------------------------------------------------------------------
# GetCurrentSoftening: min(SofteningHalo, MaxPhysicalSofteningHalo / a)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("a,soft,phys,eps", [(1.0, 1e-3, -1.0, 1e-3), (0.5, 1e-3, -1.0, 1e-3),
                                             (0.5, 1e-3, 4e-4, 8e-4), (0.5, 1e-3, 1e-3, 1e-3),
                                             (0.25, 1e-3, 1e-4, 4e-4)])
def test_softening_min_rule_through_the_centre_potential(a, soft, phys, eps):
    # one source evaluated at its own position: -G m / (a eps) (the spline's centre
    # value is -1/eps), so the potential measures the softening in use
    src = np.array([[1.0, 1.0, 1.0]], np.float32)
    m = np.float32(0.01)
    prm = dict(scale_factor=a, softening=soft, max_phys_softening=phys)
    for fn in (lambda: O.tree_potential(src, [m], src, None, params=prm)[0],
               lambda: O.direct_potential(src, [m], src, None, params=prm)):
        phi = fn()[0]
        eps_used = -G * float(m) / (a * phi)
        assert eps_used == pytest.approx(eps, rel=2e-6)
    # and outside 2.8 eps the softening plays no part: -G m / (a r)
    tgt = src + np.array([[3.0 * 2.8 * eps, 0, 0]], np.float32)
    phi = O.direct_potential(src, [m], tgt, None, params=prm)[0]
    r = float(tgt[0, 0] - src[0, 0])
    assert phi == pytest.approx(-G * float(m) / (a * r), rel=1e-6)
