"""The batched unbinding at the headline size (a power-law population drawn to
5x10^7 particles), checked through properties that do not depend on size:
graph replay and direct launches agree bit for bit, every subhalo's output is a
permutation of its particles, the bound part is ascending in energy with every
energy negative, every removed particle was unbound when it was removed, and
the per-subhalo sums match the bound sets."""
import numpy as np
import pytest

import hbtx
import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    return synth.population(seed=2, target_particles=50_000_000)


def test_fullsize_invariants(ctx, big):
    pop = big
    assert pop.n >= 50_000_000
    prm = hbtx.workload_params(direct_cutoff=-1)
    a = ctx.unbind(pop, prm)
    b = ctx.unbind(pop, prm, hbtx.NO_GRAPH)
    assert np.array_equal(a["order"], b["order"])
    assert np.array_equal(a["energy"], b["energy"])
    assert np.array_equal(a["result"], b["result"])
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
    assert res["nbound"][ok].sum() > 0.3 * pop.n
