"""Unbinding on the cosmological path and at the reference's default opening
angle, on the GPU against the oracle and against closed forms.

Every production HBT-HERONS run unbinds at a < 1 with H(z) != 0, and may cap the
physical softening (MaxPhysicalSofteningHalo).  These tests exercise:
  * the Hubble term of Snapshot_t::RelativeVelocity, v - V + H a (x - X)
    (src/snapshot.h:135-148), including periodic offsets;
  * G / a and GetCurrentSoftening at a < 1 with a physical cap
    (src/config_parser.cpp:641-649; src/gravity_tree.cpp:98-101,165);
  * AverageKinematics' physical offsets and Hubble term (src/subhalo_unbind.cpp:200-249);
  * the default TreeNodeOpenAngle 0.45 (src/config_parser.h:131);
and compare every per-subhalo output field with the oracle: spec_pot, spec_kin,
spec_am, the frame, the most-bound particle, and the core phase-space
properties (core_pos, core_vel, core_sigma_r, core_sigma_v).

Gates: as the plain parity tests (bound mass within 1e-3 summed and per subhalo
of >= 1000 bound particles, energies within 1e-5 of |potential| at q99.9); the
ring's closed forms to 2e-5 relative.
"""
import numpy as np
import pytest

import hbtx
import pyoracle as O
import synth
from parity import bound_mass_agreement, energy_errors, order_is_reference_layout
from test_oracle_pins import RING_N, RING_R, ring, ring_phi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("a,hz,omega,cutoff", [(1.0, 0.0, 300.0, 0), (0.5, 200.0, 300.0, 0),
                                               (0.5, 200.0, 0.0, 4096), (0.25, 500.0, -150.0, 4096)])
def test_ring_closed_forms(ctx, a, hz, omega, cutoff):
    # exact potentials: theta = 0 on the tree path, or the direct sum
    pop = ring(omega)
    prm = hbtx.workload_params(theta=0.0, scale_factor=a, hz=hz, bound_mass_precision=1.0, direct_cutoff=cutoff)
    out = ctx.unbind(pop, prm)
    r = out["result"][0]
    assert int(r["nbound"]) == RING_N and not r["disrupted"]
    phi0 = ring_phi(a)
    kin = 0.5 * RING_R ** 2 * (omega ** 2 + (hz * a) ** 2)
    np.testing.assert_allclose(out["energy"], kin + phi0, rtol=2e-5)
    assert float(r["spec_pot"]) == pytest.approx(phi0, rel=2e-5)
    assert float(r["spec_kin"]) == pytest.approx(kin, rel=2e-5, abs=1e-3)
    am = np.asarray(r["spec_am"], np.float64)
    assert am[2] == pytest.approx(a * omega * RING_R ** 2, rel=2e-5, abs=1e-5)
    assert abs(am[0]) < 1e-5 and abs(am[1]) < 1e-5


def test_ring_pure_hubble_flow_across_the_periodic_boundary(ctx):
    a, hz, box = 0.5, 300.0, 10.0
    pop = ring(0.0, centre=(0.0, 0.0, 0.0), box=box)
    pop.ref_pos = np.zeros((1, 3), np.float32)
    prm = hbtx.workload_params(theta=0.0, scale_factor=a, hz=hz, periodic=1, box_size=box,
                               bound_mass_precision=1.0)
    out = ctx.unbind(pop, prm)
    r = out["result"][0]
    assert int(r["nbound"]) == RING_N
    kin = 0.5 * (hz * a * RING_R) ** 2
    np.testing.assert_allclose(out["energy"], kin + ring_phi(a), rtol=2e-5)
    assert float(r["spec_kin"]) == pytest.approx(kin, rel=2e-5)


def test_softening_cap_through_the_centre_potential(ctx):
    # one source evaluated at its own position: -G m / (a eps), eps = min(eps, eps_phys / a)
    src = np.array([[1.0, 1.0, 1.0]], np.float32)
    for a, phys, eps in ((0.5, 4e-4, 8e-4), (0.5, 1e-3, 1e-3), (0.25, 1e-4, 4e-4), (1.0, -1.0, 1e-3)):
        prm = hbtx.workload_params(scale_factor=a, max_phys_softening=phys)
        phi = ctx.direct_potential(src, np.array([0.01], np.float32), src, None, params=prm)[0]
        assert -43.0071 * 0.01 / (a * phi) == pytest.approx(eps, rel=3e-6)
        tree = ctx.tree(src, np.array([0.01], np.float32), prm)
        phit = tree.potential(src)[0]
        tree.close()
        assert phit == pytest.approx(phi, rel=1e-6)


def oracle_params(prm):
    return O.default_params(**{k: v for k, v in prm.items() if k != "direct_cutoff"})


COSMO = [
    dict(scale_factor=0.5, hz=1000.0),
    dict(scale_factor=0.5, hz=1000.0, max_phys_softening=0.0004),
    dict(scale_factor=0.25, hz=2500.0, max_phys_softening=0.0002, theta=0.45),
    dict(theta=0.45),
    dict(scale_factor=0.5, hz=1000.0, theta=0.45, bound_mass_precision=1.0),
]


def compare_fields(gpu, ref, pop):
    """Per-subhalo outputs of the subhaloes both sides keep with the same bound set."""
    g, r = gpu["result"], ref["result"]
    same = []
    for s in range(pop.nsub):
        a = pop.offsets[s]
        nb = int(g[s]["nbound"])
        if g[s]["disrupted"] or r[s]["disrupted"] or nb != int(r[s]["nbound"]) or nb < 2:
            continue
        if set(gpu["order"][a:a + nb].tolist()) != set(ref["order"][a:a + nb].tolist()):
            continue
        same.append(s)
    same = np.array(same)
    assert len(same) >= 0.9 * np.sum(r["disrupted"] == 0)
    gs, rs = g[same], r[same]
    np.testing.assert_allclose(gs["com_pos"], rs["com_pos"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(gs["com_vel"], rs["com_vel"], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(gs["mbound"], rs["mbound"], rtol=1e-5)
    np.testing.assert_allclose(gs["spec_pot"], rs["spec_pot"], rtol=2e-5)
    np.testing.assert_allclose(gs["spec_kin"], rs["spec_kin"], rtol=2e-5)
    # angular momentum: relative to the scale r_half * sigma of the subhalo
    scale = np.sqrt(np.abs(rs["spec_kin"]))[:, None] * 0.05
    assert np.all(np.abs(gs["spec_am"] - rs["spec_am"]) <= 1e-5 * np.abs(rs["spec_am"]) + 1e-5 * scale)
    # core properties follow the most bound particles' order: compare where the
    # first tracer slots agree (energy ties at the top are unpinned)
    top = np.array([gpu["order"][pop.offsets[s]] == ref["order"][pop.offsets[s]] for s in same])
    assert top.mean() >= 0.99, top.mean()
    gc, rc = gs[top], rs[top]
    np.testing.assert_allclose(gc["mostbound_pos"], rc["mostbound_pos"], rtol=0, atol=0)
    np.testing.assert_allclose(gc["core_pos"], rc["core_pos"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(gc["core_vel"], rc["core_vel"], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(gc["core_sigma_r"], rc["core_sigma_r"], rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(gc["core_sigma_v"], rc["core_sigma_v"], rtol=1e-4, atol=1e-3)
    return len(same)


@pytest.mark.parametrize("cfg", COSMO, ids=["a0.5_H", "a0.5_H_physcap", "a0.25_H_physcap_theta0.45",
                                             "theta0.45", "a0.5_H_theta0.45_prec1"])
def test_population_cosmological_matches_oracle(ctx, cfg):
    pop = synth.population(seed=31, nsub=120, nmax=20000)
    prm = hbtx.workload_params(**cfg)
    gpu = ctx.unbind(pop, prm, hbtx.BATCHED | hbtx.SAVE_POTENTIAL)
    ref = O.unbind_batch(pop, oracle_params(prm), want_potential=True)
    frac, per = bound_mass_agreement(gpu, ref, pop)
    assert frac <= 1e-3, frac
    big = ref["result"]["nbound"] >= 1000
    assert np.all(per[big] <= 1e-3), per[big].max()
    err = energy_errors(gpu, ref, pop, ref["potential"])
    assert np.quantile(err, 0.999) < 1e-5, np.quantile(err, 0.999)
    assert order_is_reference_layout(gpu, pop)
    assert np.array_equal(gpu["result"]["disrupted"], ref["result"]["disrupted"])
    n = compare_fields(gpu, ref, pop)
    print(f"{cfg}: mismatch {frac:.2e}, energy q99.9 {np.quantile(err, 0.999):.2e}, {n} subhaloes compared")


def test_hydro_cosmological_matches_oracle(ctx):
    pop = synth.population(seed=32, nsub=60, nmax=8000, hydro=True)
    prm = hbtx.workload_params(dm_only=0, thermal=1, scale_factor=0.5, hz=1000.0, max_phys_softening=0.0004,
                               theta=0.45)
    gpu = ctx.unbind(pop, prm, hbtx.BATCHED | hbtx.SAVE_POTENTIAL)
    ref = O.unbind_batch(pop, oracle_params(prm), want_potential=True)
    frac, _ = bound_mass_agreement(gpu, ref, pop)
    assert frac <= 1e-3
    err = energy_errors(gpu, ref, pop, ref["potential"])
    assert np.quantile(err, 0.999) < 1e-5
    compare_fields(gpu, ref, pop)
    g, r = gpu["result"], ref["result"]
    alive = (g["disrupted"] == 0) & (r["disrupted"] == 0) & (g["nbound"] == r["nbound"])
    assert np.array_equal(g["nbound_type"][alive], r["nbound_type"][alive])
    assert np.array_equal(g["tracer_index"][alive], r["tracer_index"][alive])
    assert np.array_equal(g["mostbound_id"][alive], r["mostbound_id"][alive])


def test_subsampled_refined_cosmological_periodic_matches_oracle(ctx):
    # MaxSampleSizeOfPotentialEstimate with RefineMostBoundParticle on the periodic,
    # a < 1, H != 0 path, hydro types (upscale per type, do-not-subsample type 5)
    from test_gpu_unbind import prepare_for_subsampling
    pop = prepare_for_subsampling(synth.population(seed=34, nsub=40, nmax=20000, hydro=True))
    pos = pop.pos.copy()
    box = 50.0
    for s in range(pop.nsub):
        a, b = pop.offsets[s], pop.offsets[s + 1]
        c = pos[a:b].mean(0)
        pos[a:b] = np.mod(pos[a:b] - c + 0.01, box)
    p2 = synth.Population(pos, pop.vel, pop.mass, pop.offsets, pop.pid, pop.ptype, pop.u,
                          ref_pos=np.full((pop.nsub, 3), 0.01, np.float32), ref_vel=pop.ref_vel)
    prm = hbtx.workload_params(max_sample_size=1000, refine_most_bound=1, dm_only=0, thermal=1, periodic=1,
                               box_size=box, scale_factor=0.5, hz=1000.0, theta=0.45)
    gpu = ctx.unbind(p2, prm, hbtx.BATCHED | hbtx.SAVE_POTENTIAL)
    ref = O.unbind_batch(p2, oracle_params(prm), want_potential=True)
    frac, _ = bound_mass_agreement(gpu, ref, p2)
    assert frac <= 1e-3, frac
    err = energy_errors(gpu, ref, p2, ref["potential"])
    assert np.quantile(err, 0.999) < 1e-4, np.quantile(err, 0.999)
    assert np.array_equal(gpu["result"]["disrupted"], ref["result"]["disrupted"])
    assert order_is_reference_layout(gpu, p2, 1000)
    g, r = gpu["result"], ref["result"]
    same = (g["disrupted"] == 0) & (g["nbound"] == r["nbound"])
    assert same.mean() >= 0.9
    np.testing.assert_allclose(g["com_pos"][same], r["com_pos"][same], rtol=0, atol=2e-6)


def test_periodic_cosmological_matches_oracle(ctx):
    pop = synth.population(seed=33, nsub=40, nmax=6000)
    pos = pop.pos.copy()
    box = 50.0
    for s in range(pop.nsub):
        a, b = pop.offsets[s], pop.offsets[s + 1]
        c = pos[a:b].mean(0)
        pos[a:b] = np.mod(pos[a:b] - c + 0.01, box)
    p2 = synth.Population(pos, pop.vel, pop.mass, pop.offsets, pop.pid,
                          ref_pos=np.full((pop.nsub, 3), 0.01, np.float32), ref_vel=pop.ref_vel)
    prm = hbtx.workload_params(periodic=1, box_size=box, scale_factor=0.5, hz=1000.0, theta=0.45)
    gpu = ctx.unbind(p2, prm, hbtx.BATCHED | hbtx.SAVE_POTENTIAL)
    ref = O.unbind_batch(p2, oracle_params(prm), want_potential=True)
    frac, _ = bound_mass_agreement(gpu, ref, p2)
    assert frac <= 1e-3
    err = energy_errors(gpu, ref, p2, ref["potential"])
    assert np.quantile(err, 0.999) < 1e-5
