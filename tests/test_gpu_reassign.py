"""Particle reassignment on the GPU (hbtx_reassign_targets:
SubhaloSnapshot_t::ReassignParticles, src/subhalo_reassign_gas.cpp:65-197, with
GeoTree_t::NearestNeighbours, src/geometric_tree.cpp:108-141) against the
oracle's restatement, target for target.  Exact agreement is expected: the
neighbour lists are ordered by the same float distances; only two tracers at
exactly the same float distance could be ordered differently (not exercised)."""
import numpy as np
import pytest

import pyoracle as O
import synth

pytestmark = pytest.mark.gpu


def both(ctx, pts, tracer, sub, goff, k=10, rguess=None, **prm):
    print(f"reassign n={len(sub)} groups={len(goff) - 1} k={k} rguess={rguess} {prm}", flush=True)
    g, ms = ctx.reassign_targets(pts, tracer, sub, goff, k=k, rguess=rguess, params=prm)
    o = O.reassign_targets(pts, tracer, sub, goff, k=k, rguess=rguess, params=O.default_params(**prm))
    return g, o


@pytest.mark.parametrize("seed,k", [(1, 10), (2, 1), (3, 4), (4, 16), (5, 32)])
def test_host_haloes_match_oracle(ctx, seed, k):
    h = synth.host_haloes(seed=seed, ngroups=6, n_sub=(30, 2000))
    g, o = both(ctx, h.pos, h.is_tracer, h.sub, h.group_offsets, k=k)
    np.testing.assert_array_equal(g, o)
    assert (o >= 0).sum() > 100  # moves happen
    assert np.all(g[h.is_tracer == 1] == -1)


def test_periodic_host_haloes_match_oracle(ctx):
    h = synth.host_haloes(seed=11, ngroups=8, n_sub=(30, 1500), box=2.0, spread=0.08)
    g, o = both(ctx, h.pos, h.is_tracer, h.sub, h.group_offsets, periodic=1, box_size=2.0)
    np.testing.assert_array_equal(g, o)
    assert (o >= 0).sum() > 100


def test_halo_across_the_boundary(ctx):
    """A host halo centred on a box corner: offsets wrap on every axis."""
    h = synth.host_haloes(seed=12, ngroups=1, subs=(4, 4), n_sub=(500, 1500), spread=0.05)
    box = 10.0
    pos = np.mod(h.pos - h.pos[0] + 1e-3, box).astype(np.float32)  # first particle just inside the corner
    g, o = both(ctx, pos, h.is_tracer, h.sub, h.group_offsets, periodic=1, box_size=box)
    np.testing.assert_array_equal(g, o)
    assert (o >= 0).sum() > 10


def test_small_first_radius_and_many_growths(ctx):
    """rguess far below the spacing: tens of growths before the first tracer."""
    h = synth.host_haloes(seed=13, ngroups=3, n_sub=(30, 800))
    g, o = both(ctx, h.pos, h.is_tracer, h.sub, h.group_offsets, rguess=1e-7)
    np.testing.assert_array_equal(g, o)
    # and one that already holds every tracer: a single search
    g, o = both(ctx, h.pos, h.is_tracer, h.sub, h.group_offsets, rguess=100.0)
    np.testing.assert_array_equal(g, o)


def test_edge_cases(ctx):
    rng = np.random.default_rng(14)
    # empty host haloes, one without tracers, one with fewer tracers than k, one of tracers only
    parts = [
        (np.zeros((0, 3)), [], []),
        (rng.normal(0, 0.01, (40, 3)), [0] * 40, rng.integers(0, 3, 40)),
        (rng.normal(5, 0.01, (30, 3)), [1, 0, 1, 0, 0, 1] + [0] * 24, rng.integers(3, 6, 30)),
        (np.zeros((0, 3)), [], []),
        (rng.normal(9, 0.01, (50, 3)), [1] * 50, rng.integers(6, 8, 50)),
        (rng.normal(12, 0.02, (700, 3)), (rng.random(700) < 0.5).astype(int), rng.integers(8, 12, 700)),
    ]
    pos = np.concatenate([p for p, _, _ in parts]).astype(np.float32)
    tr = np.concatenate([np.asarray(t, np.uint8) for _, t, _ in parts]).astype(np.uint8)
    sub = np.concatenate([np.asarray(s, np.int32) for _, _, s in parts]).astype(np.int32)
    goff = np.cumsum([0] + [len(t) for _, t, _ in parts]).astype(np.int64)
    g, o = both(ctx, pos, tr, sub, goff)
    np.testing.assert_array_equal(g, o)
    assert np.all(g[goff[1]:goff[2]] == -1)  # no tracers: nobody moves
    assert np.all(g[goff[4]:goff[5]] == -1)  # tracers only
    # coincident tracers of one subhalo and a query on top of them
    pos2 = np.concatenate([np.full((20, 3), 1.0), np.full((20, 3), 1.5), [[1.0, 1.0, 1.0]]]).astype(np.float32)
    tr2 = np.r_[np.ones(40), 0].astype(np.uint8)
    sub2 = np.r_[np.zeros(20), np.ones(20), 1].astype(np.int32)
    g, o = both(ctx, pos2, tr2, sub2, [0, 41])
    np.testing.assert_array_equal(g, o)
    assert g[40] == 0


def test_oracle_pin_cases_on_gpu(ctx):
    """The radius-growth case of the oracle's pin tests: the reference collects the
    near tracers twice, so the particle moves although own tracers are among its
    true ten nearest."""
    eps = 0.001
    r0 = np.float32(eps * 0.1)
    q = np.array([[3.0, 3.0, 3.0]])
    near = q + np.array([[0.3, 0, 0], [-0.3, 0, 0], [0, 0.3, 0], [0, -0.3, 0], [0, 0, 0.3]]) * r0
    rng = np.random.default_rng(3)
    d = rng.normal(size=(20, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = np.concatenate([q, near, q + d * (1.1 * r0)]).astype(np.float32)
    tracer = np.r_[0, np.ones(25)].astype(np.uint8)
    sub = np.r_[0, np.ones(5), np.zeros(20)].astype(np.int32)
    g, o = both(ctx, pts, tracer, sub, [0, 26], softening=eps)
    assert g[0] == 1 and o[0] == 1


def test_invalid_arguments(ctx):
    import hbtx
    h = synth.host_haloes(seed=15, ngroups=1, subs=(2, 2), n_sub=(30, 40))
    with pytest.raises(hbtx.HbtxError):
        ctx.reassign_targets(h.pos, h.is_tracer, h.sub, h.group_offsets, k=33)
    with pytest.raises(hbtx.HbtxError):
        ctx.reassign_targets(h.pos, h.is_tracer, h.sub, [0, h.n - 1])
