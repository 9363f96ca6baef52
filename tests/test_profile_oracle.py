"""Pins the oracle's restatement of Subhalo_t::CalculateProfileProperties and
CalculateShape (src/subhalo.cpp:175-309) on inputs whose answers are known in
closed form."""
import numpy as np

import pyoracle as O


def run(pos, mass, **kw):
    pos = np.asarray(pos, np.float32)
    mass = np.asarray(mass, np.float32)
    n = len(mass)
    return O.profiles(pos, mass, [0], [n], [mass.sum()], O.default_params(**kw))[0]


def shells(radii, per=6, m=0.5, centre=(1.0, 2.0, 3.0)):
    """A centre particle plus `per` particles on the +-axes of each radius."""
    c = np.asarray(centre, np.float64)
    pts = [c]
    for r in radii:
        for ax in range(3):
            for sgn in (1.0, -1.0):
                d = np.zeros(3)
                d[ax] = sgn * r
                pts.append(c + d)
    return np.array(pts), np.full(len(pts), m)


def test_profile_of_shells():
    # softening 0.05 < every shell radius; the centre sits at r = 0, clamped to the softening.
    # Cumulative mass at the last particle of shell k = m (1 + 6k)
    radii = [0.1, 0.2, 0.4, 0.5]
    pos, mass = shells(radii)
    p = run(pos, mass, softening=0.05, G=43.0071, scale_factor=1.0)
    m, eps = 0.5, 0.05
    r = [eps] + radii
    v = [m / eps] + [m * (1 + 6 * (k + 1)) / rr for k, rr in enumerate(radii)]
    k = int(np.argmax(v))
    assert k > 0
    assert np.isclose(p["rmax"], r[k], rtol=1e-6)
    assert np.isclose(p["vmax"], np.sqrt(v[k] * 43.0071), rtol=1e-6)
    # with a larger softening the clamped centre dominates: Rmax = eps
    p2 = run(pos, mass, softening=0.001)
    assert np.isclose(p2["rmax"], 0.001) and np.isclose(p2["vmax"], np.sqrt(m / 0.001 * 43.0071), rtol=1e-6)
    # 25 particles: index 12 is on the second shell; the last is on the outermost
    assert np.isclose(p["r_half"], 0.2, rtol=1e-6) and np.isclose(p["r_enclose"], 0.5, rtol=1e-6)
    # H = 0: the threshold density is 0, so the outermost particle sets M200 and R200 is infinite
    assert p["has_r200"] == 1 and np.isclose(p["m200"], mass.sum(), rtol=1e-6) and np.isinf(p["r200"])


def test_softening_floor_and_overdensity():
    radii = [0.1, 0.2, 0.4]
    pos, mass = shells(radii)
    # every radius below the softening: all clamped to eps, v grows with m, Rmax = eps
    p = run(pos, mass, softening=1.0)
    assert np.isclose(p["rmax"], 1.0) and np.isclose(p["r_half"], 1.0) and np.isclose(p["r_enclose"], 1.0)
    # overdensity threshold: rho = 200 H^2 / (2 G) a^3; pick H so that only the two inner shells qualify
    G, a = 43.0071, 1.0
    m = 0.5
    msh = [m * (1 + 6 * (k + 1)) for k in range(3)]
    rho_target = 0.5 * (msh[1] / 0.2 ** 3 + msh[2] / 0.4 ** 3)  # between shell 2 and shell 3 densities
    H = np.sqrt(rho_target * 2 * G / 200.0 / a ** 3)
    p = run(pos, mass, softening=0.001, hz=H, G=G)
    rho = np.float32(200.0 * np.float32(H) * np.float32(H) / 2.0 / np.float32(G))
    assert p["has_r200"] == 1 and np.isclose(p["m200"], msh[1], rtol=1e-6)
    assert np.isclose(p["r200"], (msh[1] / rho) ** (1 / 3), rtol=1e-5)


def test_shape_closed_form():
    # centre + (+-a,0,0), (0,+-b,0), (0,0,+-c): diagonal tensor 2 m a^2 / M etc., weighted 2 m / M per axis
    a, b, c, m = 0.3, 0.2, 0.1, 0.25
    cen = np.array([5.0, 5.0, 5.0])
    pos = [cen] + [cen + s * np.array(v) for v in ((a, 0, 0), (0, b, 0), (0, 0, c)) for s in (1, -1)]
    mass = np.full(7, m)
    p = run(np.array(pos), mass)
    M = mass.sum()
    np.testing.assert_allclose(p["inertia"], [2 * m * a * a / M, 0, 0, 2 * m * b * b / M, 0, 2 * m * c * c / M],
                               rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(p["inertia_w"], [2 * m / M, 0, 0, 2 * m / M, 0, 2 * m / M], rtol=1e-5, atol=1e-9)
    # an off-diagonal pair: points on the x=y diagonal
    pos2 = [cen, cen + np.array([a, a, 0]), cen - np.array([a, a, 0])]
    p2 = run(np.array(pos2), np.full(3, m))
    assert np.isclose(p2["inertia"][1], 2 * m * a * a / (3 * m), rtol=1e-5)
    assert np.isclose(p2["inertia_w"][1], 2 * m * 0.5 / (3 * m), rtol=1e-5)


def test_periodic_wrap_and_tiny_sets():
    box = 10.0
    cen = np.array([9.95, 5.0, 5.0])
    pos = np.array([cen, [0.05, 5.0, 5.0]])  # 0.1 apart across the boundary
    p = O.profiles(pos.astype(np.float32), np.full(2, 0.5, np.float32), [0], [2], [1.0],
                   O.default_params(periodic=1, box_size=box, softening=0.001))[0]
    assert np.isclose(p["r_enclose"], 0.1, rtol=1e-4)
    # Nbound <= 1: everything zero
    z = O.profiles(pos.astype(np.float32), np.full(2, 0.5, np.float32), [0], [1], [0.5], O.default_params())[0]
    assert z["rmax"] == 0 and z["vmax"] == 0 and z["has_r200"] == 0 and not z["inertia"].any()
