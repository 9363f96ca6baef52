"""Pins the CPU oracle (the checker of every GPU parity test) with checks that do
not restate its own code: closed forms of the softened potential, the theta=0
limit, tree structure recomputed independently in numpy, a recursive walk of the
exported tree, a brute-force unbinding loop, and the reference's own core
phase-space test pattern (unit_tests/test_build_pos_vel.cpp:14-169)."""
import numpy as np
import pytest

import pyoracle as O
import synth

G = 43.0071
EPS = float(np.float32(0.001))  # SofteningHalo is a float in the reference
H = 2.8 * EPS


def pot_at(r, m=0.01, theta=0.7):
    """Potential of one source of mass m at distance r (tree and direct)."""
    src = np.array([[1.0, 1.0, 1.0]], np.float32)
    tgt = src + np.array([[r, 0.0, 0.0]], np.float32)
    t, _ = O.tree_potential(src, [m], tgt, None, params=dict(theta=theta))
    d = O.direct_potential(src, [m], tgt, None)
    return t[0], d[0], float(tgt[0, 0] - src[0, 0])


M32 = float(np.float32(0.01))


def test_kernel_is_newtonian_outside_h():
    for r in [H * 1.0001, 0.01, 0.3]:
        t, d, rr = pot_at(r)
        assert t == pytest.approx(-G * M32 / rr, rel=1e-12)
        assert d == pytest.approx(-G * M32 / rr, rel=1e-12)


def test_kernel_centre_value_is_minus_m_over_eps():
    # W(0) = -2.8/h = -1/eps (src/gravity_tree.cpp:101 relies on it)
    t, d, _ = pot_at(0.0)
    assert d == pytest.approx(-G * M32 / EPS, rel=1e-6)
    assert t == pytest.approx(-G * M32 / EPS, rel=1e-6)


def test_kernel_continuous_and_monotone():
    rs = np.linspace(1e-5, 1.5 * H, 400)
    vals = np.array([pot_at(r)[1] for r in rs])
    assert np.all(np.diff(vals) > 0), "potential must rise with r"
    for r0 in [0.5 * H, H]:
        a = pot_at(r0 * (1 - 1e-5))[1]
        b = pot_at(r0 * (1 + 1e-5))[1]
        assert abs(a - b) < 1e-3 * abs(a)
    # the softened potential is shallower than both the centre value and Newton
    for r in [0.2 * H, 0.7 * H]:
        d = pot_at(r)[1]
        assert d >= -G * M32 / EPS * (1 + 1e-9) and d >= -G * M32 / r and d < 0


def test_theta_zero_tree_is_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.normal(1.0, 0.05, (800, 3)).astype(np.float32)
    m = rng.uniform(0.5, 2.0, 800).astype(np.float32) * 0.01
    t, _ = O.tree_potential(x, m, x, None, params=dict(theta=0.0))
    d = O.direct_potential(x, m, x, None)
    assert np.max(np.abs(t - d) / np.abs(d)) < 1e-12


def test_self_term_cancels_self_interaction():
    rng = np.random.default_rng(1)
    x = rng.normal(1.0, 0.05, (300, 3)).astype(np.float32)
    m = np.full(300, 0.01, np.float32)
    with_self, _ = O.tree_potential(x, m, x, m, params=dict(theta=0.0))
    d = O.direct_potential(x[1:], m[1:], x[:1], None)
    # removing particle 0 from the sources == passing its mass as targetMass
    assert with_self[0] == pytest.approx(d[0], rel=1e-6)


def octree_keys_numpy(x, levels=21):
    """Child-index sequence of every point, recomputed from the cube geometry."""
    xd = x.astype(np.float64)
    lo, hi = xd.min(0), xd.max(0)
    L = (hi - lo).max()
    centre = 0.5 * (hi + lo)
    c = np.tile(centre, (len(x), 1))
    lh = 0.5 * L
    key = np.zeros(len(x), np.uint64)
    for _ in range(levels):
        lh *= 0.5
        b = xd > c
        c = c + np.where(b, lh, -lh)
        sub = b[:, 0].astype(np.uint64) + 2 * b[:, 1].astype(np.uint64) + 4 * b[:, 2].astype(np.uint64)
        key = key * np.uint64(8) + sub
    return key, L


def test_tree_structure_independent():
    pop = synth.single_nfw(4, n_halo=3000, n_inter=500)
    x, m = pop.pos, pop.mass
    t = O.tree_dfs(x, m)
    parts = t["kind"] == 0
    # particles appear in key order (depth-first order of the octree)
    key, L = octree_keys_numpy(x)
    order = np.argsort(key, kind="stable")
    assert np.array_equal(t["pidx"][parts], order)
    # every cell: mass and centre of mass of the particles of its subtree, cube size rule
    res = np.float32(np.float32(EPS) * np.float32(0.1))
    n = len(t["kind"])
    cs = np.concatenate([[0.0], np.cumsum(np.where(parts, t["mass"], 0.0))])
    for q in np.where(~parts)[0][::7]:
        a, b = q + 1, t["skip"][q]
        assert q < b <= n
        sub = np.arange(a, b)
        sp = sub[parts[sub]]
        mm = m[t["pidx"][sp]].astype(np.float64)
        assert t["mass"][q] == pytest.approx(mm.sum(), rel=1e-5)
        com = (x[t["pidx"][sp]].astype(np.float64) * mm[:, None]).sum(0) / mm.sum()
        np.testing.assert_allclose(t["com"][q], com, rtol=0, atol=1e-6 * L + 1e-7)
        d = t["depth"][q]
        Ld = L
        for _ in range(d):
            Ld = Ld / 2 if Ld >= res else Ld
        assert t["len"][q] == pytest.approx(np.float32(Ld), rel=1e-6)
    # subtrees nest: a skip never lands inside a sibling's subtree
    for q in np.where(~parts)[0][::11]:
        for c in range(q + 1, t["skip"][q]):
            assert t["skip"][c] <= t["skip"][q]


def recursive_walk(t, target, theta):
    """Barnes-Hut potential by explicit recursion over the exported preorder tree."""
    th2 = np.float32(theta) ** 2
    h, hinv = 2.8 * EPS, 1.0 / (2.8 * EPS)

    def kern(r2, mass):
        r = np.sqrt(r2)
        if r >= h:
            return -mass / r
        u = r * hinv
        if u < 0.5:
            wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6))
        else:
            wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)))
        return mass * hinv * wp

    def visit(q):
        d = t["com"][q].astype(np.float64) - target
        r2 = float(d @ d)
        if t["kind"][q] == 1:
            ln = np.float32(t["len"][q])
            if ln * ln > r2 * th2:
                tot, c = 0.0, q + 1
                while c < t["skip"][q]:
                    tot += visit(c)
                    c = t["skip"][c]
                return tot
        return kern(r2, float(t["mass"][q]))

    return visit(0) * G


def test_stackless_walk_equals_recursive_walk():
    pop = synth.single_nfw(5, n_halo=400, n_inter=100)
    x, m = pop.pos, pop.mass
    t = O.tree_dfs(x, m)
    tgt = x[::37].astype(np.float64)
    ref, _ = O.tree_potential(x, m, x[::37], None)
    got = np.array([recursive_walk(t, p, 0.7) for p in tgt])
    np.testing.assert_allclose(ref, got, rtol=1e-12)


def brute_force_unbind(pop, prec):
    """Subhalo_t::Unbind with exact potentials (theta -> 0), written with numpy."""
    x = pop.pos.astype(np.float64)
    v = pop.vel.astype(np.float32)
    m = pop.mass.astype(np.float32)
    rp = pop.ref_pos[0].copy()
    rv = pop.ref_vel[0].copy()
    n = len(x)
    h, hinv = 2.8 * EPS, 1.0 / (2.8 * EPS)

    def phi(targets, sources, selfm):
        d = x[targets][:, None, :] - x[sources][None, :, :]
        r = np.sqrt((d * d).sum(-1))
        u = r * hinv
        wp = np.where(u < 0.5, -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6)),
                      -3.2 + 0.066666666667 / np.maximum(u, 1e-30)
                      + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u))))
        k = np.where(r >= h, -1.0 / np.maximum(r, 1e-30), hinv * wp)
        return ((k * m[sources].astype(np.float64)[None, :]).sum(1) + selfm / np.float32(EPS)) * G

    cand = np.arange(n)
    E = np.zeros(n, np.float32)
    corr = False
    removed = np.array([], int)
    old = None
    while True:
        nlast = len(cand)
        if corr:
            dV = (old[1] - rv).astype(np.float32)
            dK = np.float32(0.5 * np.float32((dV * dV).sum()))
            ov = v[cand] - old[1]
            E[cand] = (E[cand].astype(np.float64) + ((ov * dV).sum(1).astype(np.float32) + dK)
                       - phi(cand, removed, 0.0)).astype(np.float32)
        else:
            dv = v[cand] - rv
            k2 = (dv * dv).sum(1).astype(np.float32)
            E[cand] = (k2.astype(np.float64) * 0.5 + phi(cand, cand, m[cand])).astype(np.float32)
        bound = cand[E[cand] < 0]
        removed = cand[E[cand] >= 0]
        if len(bound) < 20:
            return None
        if len(removed) < len(bound):
            corr = True
            old = (rp.copy(), rv.copy())
        mb = m[bound].astype(np.float64)
        rp = ((x[bound] * mb[:, None]).sum(0) / mb.sum()).astype(np.float32)
        rv = ((v[bound].astype(np.float64) * mb[:, None]).sum(0) / mb.sum()).astype(np.float32)
        cand = bound
        if np.float32(len(bound)) >= np.float32(nlast) * np.float32(prec):
            return set(bound.tolist()), E


@pytest.mark.parametrize("prec", [0.995, 1.0])
def test_unbinding_loop_matches_brute_force(prec):
    pop = synth.single_nfw(6, n_halo=300, n_inter=60)
    # exact potentials: theta = 0 makes the oracle's tree a direct sum
    r = O.unbind_batch(pop, O.default_params(theta=0.0, bound_mass_precision=prec), nthreads=1)
    bf = brute_force_unbind(pop, prec)
    nb = int(r["result"][0]["nbound"])
    got = set(r["order"][:nb].tolist())
    ref, E = bf
    assert got == ref
    e_or = r["energy"][:nb]
    e_bf = E[r["order"][:nb]]
    np.testing.assert_allclose(e_or, e_bf, rtol=2e-5, atol=1e-3)


def test_fixed_point_at_full_precision():
    # BoundMassPrecision 1: the final iteration removed nobody, so every bound
    # particle was bound with respect to the final set (up to the tree error)
    pop = synth.single_nfw(7, n_halo=2000, n_inter=400)
    r = O.unbind_batch(pop, O.default_params(bound_mass_precision=1.0), nthreads=1)
    nb = int(r["result"][0]["nbound"])
    assert np.all(r["energy"][:nb] < 0)
    assert np.all(np.diff(r["energy"][:nb]) >= 0)
    # removal is monotone: unbound particles never re-enter
    assert sorted(r["order"].tolist()) == list(range(pop.n))


def test_core_phase_space_reference_pattern():
    """unit_tests/test_build_pos_vel.cpp: uniform spheres at several positions of a
    periodic box of 100, tracer fraction 10%, checks at 1e-5."""
    rng = np.random.default_rng(0)
    box = 100.0
    prm = dict(periodic=1, box_size=box, dm_only=0, tracer_mask=(1 << 1) | (1 << 4))
    for x0 in [-3.0, -0.1, 0.1, 50.0, 99.9, 100.1, 103.0]:
        for dim in range(3):
            centre = np.array([10.0, 10.0, 10.0])
            centre[dim] = x0
            n = 100
            pts = []
            while len(pts) < n:
                p = rng.uniform(-2.0, 2.0, 3)
                if p @ p <= 4.0:
                    pts.append(p)
            pos = np.mod(np.array(pts) + centre, box).astype(np.float32)
            vel = (20.0 + rng.uniform(-30.0, 30.0, (n, 3))).astype(np.float32)
            typ = np.where(rng.uniform(0, 1, n) < 0.1, 1, 0).astype(np.int32)
            mass = np.ones(n, np.float32)
            cp, cv, sr, sv = O.core_phase_space(pos, vel, mass, typ, params=prm)
            tr = np.where(typ == 1)[0][:20]
            sel = np.concatenate([tr, np.where(typ != 1)[0][:20 - len(tr)]])
            ref = pos[0].astype(np.float64)
            w = pos[sel].astype(np.float64)
            w = np.where(w > ref + box / 2, w - box, np.where(w < ref - box / 2, w + box, w))
            check_pos = w.mean(0)
            d = np.abs(np.mod(cp - check_pos + box / 2, box) - box / 2)
            assert np.sqrt((d ** 2).sum()) < 1e-5 * 10
            np.testing.assert_allclose(cv, vel[sel].astype(np.float64).mean(0), atol=1e-5 * 100)
            srr = np.sqrt(((w - check_pos) ** 2).mean(0).sum())
            assert sr == pytest.approx(srr, rel=1e-5)
            svv = np.sqrt(((vel[sel] - vel[sel].astype(np.float64).mean(0)) ** 2).mean(0).sum())
            assert sv == pytest.approx(svv, rel=1e-5)


def test_disruption_below_min_num_part():
    rng = np.random.default_rng(2)
    pos = rng.normal(5.0, 0.01, (19, 3)).astype(np.float32)
    vel = np.zeros((19, 3), np.float32)
    pop = synth.Population(pos, vel, np.full(19, 0.01, np.float32), [0, 19], np.arange(19))
    r = O.unbind_batch(pop, O.default_params(), nthreads=1)
    assert r["result"][0]["disrupted"] == 1 and r["result"][0]["nbound"] == 0
