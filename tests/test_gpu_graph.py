"""Batched rounds replayed from captured HIP graphs give the same results as
rounds launched kernel by kernel, across repeated replays with direct work in
between, for the plain, subsampled, periodic and warm-started iterations."""
import numpy as np
import pytest

import hbtx
import synth

pytestmark = pytest.mark.gpu


def same(a, b):
    assert np.array_equal(a["order"], b["order"])
    assert np.array_equal(a["energy"], b["energy"])
    assert np.array_equal(a["result"], b["result"])


def run(ctx, pop, prm, flags):
    out = ctx.unbind(pop, prm, flags)
    return out, ctx.stats()


def test_graph_replay_matches_direct(ctx):
    pop = synth.population(seed=51, nsub=400, nmax=60000)
    prm = hbtx.workload_params(direct_cutoff=512)
    direct, sd = run(ctx, pop, prm, hbtx.BATCHED | hbtx.NO_GRAPH)
    assert sd["graph_rounds"] == 0 and sd["graph_captures"] == 0
    assert sd["rounds"] >= 3
    first, s1 = run(ctx, pop, prm, hbtx.BATCHED)
    same(first, direct)
    assert s1["rounds"] == sd["rounds"]
    assert s1["graph_rounds"] == s1["rounds"] - 1
    # replays of the same batch: the graphs are reused, direct runs in between
    for _ in range(3):
        again, s = run(ctx, pop, prm, hbtx.BATCHED)
        same(again, direct)
        assert s["graph_rounds"] == s1["graph_rounds"]
        d, _ = run(ctx, pop, prm, hbtx.BATCHED | hbtx.NO_GRAPH)
        same(d, direct)


def test_graph_recaptured_for_new_batch_and_params(ctx):
    pa = synth.population(seed=52, nsub=200, nmax=30000)
    pb = synth.population(seed=53, nsub=150, nmax=50000)
    prm = hbtx.workload_params(direct_cutoff=256)
    for pop in (pa, pb, pa):
        ref, _ = run(ctx, pop, prm, hbtx.BATCHED | hbtx.NO_GRAPH)
        got, s = run(ctx, pop, prm, hbtx.BATCHED)
        same(got, ref)
        assert s["graph_rounds"] > 0
    # another opening angle changes the captured parameters
    prm2 = hbtx.workload_params(direct_cutoff=256, theta=0.5)
    ref, _ = run(ctx, pa, prm2, hbtx.BATCHED | hbtx.NO_GRAPH)
    got, s = run(ctx, pa, prm2, hbtx.BATCHED)
    same(got, ref)
    assert s["graph_captures"] >= 1


def test_graph_subsampled_and_periodic(ctx):
    pop = synth.population(seed=54, nsub=120, nmax=40000)
    for kw in (dict(max_sample_size=2000), dict(periodic=1, box_size=100.0)):
        prm = hbtx.workload_params(direct_cutoff=256, **kw)
        ref, _ = run(ctx, pop, prm, hbtx.BATCHED | hbtx.NO_GRAPH)
        got, s = run(ctx, pop, prm, hbtx.BATCHED)
        same(got, ref)
        assert s["graph_rounds"] > 0


def test_graph_warm_start(ctx):
    pop = synth.population(seed=55, nsub=150, nmax=40000)
    prm = hbtx.workload_params(direct_cutoff=256)
    plain = ctx.unbind(pop, prm, hbtx.BATCHED | hbtx.NO_GRAPH)
    # proposal: the plain bound set with a few flips
    bound = np.zeros(pop.n, np.uint8)
    for s in range(pop.nsub):
        a = pop.offsets[s]
        nb = int(plain["result"]["nbound"][s])
        bound[a + plain["order"][a:a + nb]] = 1
    rng = np.random.default_rng(5)
    flip = rng.random(pop.n) < 0.01
    bound[flip] ^= 1
    outs = []
    for flags in (hbtx.NO_GRAPH, 0, 0):
        ctx.load(pop)
        ctx.set_proposal(bound)
        ctx.run(prm, hbtx.BATCHED | hbtx.WARMSTART | flags)
        outs.append((ctx.fetch(), ctx.stats()))
    for o, s in outs[1:]:
        same(o, outs[0][0])
        assert s["graph_rounds"] == s["rounds"] - 1
