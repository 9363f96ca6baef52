"""Post-unbinding subhalo properties on the GPU (hbtx_batch_profiles:
Subhalo_t::CalculateProfileProperties and CalculateShape, src/subhalo.cpp:175-309)
against the oracle's restatement over the same final particle order."""
import numpy as np
import pytest

import hbtx
import pyoracle as O
import synth

pytestmark = pytest.mark.gpu


def oracle_on_gpu_order(pop, out, prm):
    res = out["result"]
    idx = np.concatenate([pop.offsets[s] + out["order"][pop.offsets[s]:pop.offsets[s + 1]]
                          for s in range(pop.nsub)])
    nb = np.where(res["disrupted"] == 0, res["nbound"], 0)
    return O.profiles(pop.pos[idx], pop.mass[idx], pop.offsets[:-1], nb, res["mbound"],
                      O.default_params(**{k: v for k, v in prm.items() if k != "direct_cutoff"}))


def compare(g, r, nb):
    live = nb > 1
    assert np.all(g[~live]["rmax"] == 0) and np.all(g[~live]["has_r200"] == 0)
    assert not g[~live]["inertia"].any()
    np.testing.assert_array_equal(g["r_half"], r["r_half"])
    np.testing.assert_array_equal(g["r_enclose"], r["r_enclose"])
    np.testing.assert_array_equal(g["has_r200"], r["has_r200"])
    # cumulative masses: double sums grouped differently, stored as float
    np.testing.assert_allclose(g["vmax"], r["vmax"], rtol=1e-6)
    np.testing.assert_allclose(g["m200"], r["m200"], rtol=1e-6)
    np.testing.assert_allclose(g["r200"], r["r200"], rtol=1e-6)
    # the radius of the maximum moves only where two values of M/r agree to rounding
    assert np.mean(g["rmax"] != r["rmax"]) <= 0.01
    np.testing.assert_allclose(g["inertia"], r["inertia"], rtol=2e-6, atol=1e-12)
    np.testing.assert_allclose(g["inertia_w"], r["inertia_w"], rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("kind", ["dm", "hydro", "periodic"])
def test_profiles_match_oracle(ctx, kind):
    pop = synth.population(seed=61, nsub=160, nmax=30000, hydro=(kind == "hydro"))
    kw = dict(hz=100.0)
    if kind == "hydro":
        kw.update(dm_only=0, thermal=1)
    if kind == "periodic":
        kw.update(periodic=1, box_size=50.0)
    prm = hbtx.workload_params(**kw)
    out = ctx.unbind(pop, prm)
    g = ctx.profiles(prm)
    r = oracle_on_gpu_order(pop, out, prm)
    res = out["result"]
    nb = np.where(res["disrupted"] == 0, res["nbound"], 0)
    compare(g, r, nb)
    assert np.sum(g["has_r200"]) > 0.5 * np.sum(nb > 1)
    # profiles follow the last run: a second run with other parameters changes them
    out2 = ctx.unbind(pop, dict(prm, theta=0.5))
    g2 = ctx.profiles(prm)
    compare(g2, oracle_on_gpu_order(pop, out2, prm), np.where(out2["result"]["disrupted"] == 0,
                                                              out2["result"]["nbound"], 0))


def test_caller_given_bound_sets(ctx):
    """hbtx_profiles on bound sets handed over by the caller equals the batch path."""
    pop = synth.population(seed=62, nsub=90, nmax=20000)
    prm = hbtx.workload_params(hz=70.0)
    out = ctx.unbind(pop, prm)
    g = ctx.profiles(prm)
    res = out["result"]
    nb = np.where(res["disrupted"] == 0, res["nbound"], 0)
    idx = np.concatenate([pop.offsets[s] + out["order"][pop.offsets[s]:pop.offsets[s] + nb[s]]
                          for s in range(pop.nsub)])
    off = np.concatenate([[0], np.cumsum(nb)])
    h = ctx.profiles_of(pop.pos[idx], pop.mass[idx], off, nb, res["mbound"], prm)
    for f in ("rmax", "vmax", "r_half", "r_enclose", "r200", "m200", "has_r200", "inertia", "inertia_w"):
        np.testing.assert_array_equal(h[f], g[f])


def test_profiles_need_a_run():
    c = hbtx.Context(0)
    with pytest.raises(hbtx.HbtxError):
        c._nsub = 1
        c.profiles(hbtx.workload_params())
    c.close()
