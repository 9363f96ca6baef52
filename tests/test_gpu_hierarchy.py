"""Hierarchical unbinding (Subhalo_t::RecursiveUnbind, src/subhalo_unbind.cpp:810-840,
with the phase-space mergers of src/subhalo_merge.cpp:45-91) through the C++
host layer, one GPU batch per level of the hierarchies.  The driver
(tests/cpp/hierarchy_test.cpp) first checks the level-batched result against
the reference's depth-first recursion over single-subhalo GPU unbindings
(identical subhaloes); here the result is compared with the same recursion over
the CPU oracle (tests/hierarchy_oracle.py)."""
import os
import subprocess

import numpy as np
import pytest

import hbtx
import hierarchy_oracle as HO
import pyoracle as O
import synth

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(hbtx.LIB_PATH), "hierarchy_test")
REC = np.dtype([("nbound", np.int64), ("sink", np.int64), ("np", np.int64), ("mbid", np.int64),
                ("snap_sink", np.int32), ("snap_death", np.int32), ("core", np.float32, 8)])


def read_subs(path, nsub):
    raw = open(path, "rb").read()
    subs = np.frombuffer(raw, REC, count=nsub)
    ids = np.frombuffer(raw, np.int64, offset=REC.itemsize * nsub)
    return subs, np.split(ids, np.cumsum(subs["np"])[:-1])


def run_driver(H, tmp_path, merge_trapped, num_sub_old=None):
    pop = H.pop
    coff = np.concatenate([[0], np.cumsum([len(c) for c in H.children])]).astype(np.int64)
    child = np.array([c for cs in H.children for c in cs], np.int64)
    inp, outp = tmp_path / f"h{int(merge_trapped)}.bin", tmp_path / f"o{int(merge_trapped)}.bin"
    with open(inp, "wb") as f:
        np.array([pop.nsub, pop.n, len(H.roots), len(child), int(merge_trapped), 0], np.int64).tofile(f)
        pop.offsets.astype(np.int64).tofile(f)
        for a in (pop.pos, pop.vel, pop.mass):
            np.ascontiguousarray(a, np.float32).tofile(f)
        pop.pid.astype(np.int64).tofile(f)
        for a in (pop.ref_pos, pop.ref_vel):
            np.ascontiguousarray(a, np.float32).tofile(f)
        for a in (H.track, H.sink, coff, child, np.asarray(H.roots, np.int64)):
            np.asarray(a, np.int64).tofile(f)
        if num_sub_old is not None:
            np.array([num_sub_old], np.int64).tofile(f)
    snapp = tmp_path / f"s{int(merge_trapped)}.bin"
    r = subprocess.run([BIN, str(inp), str(outp), str(snapp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    info = dict(zip(words[::2], map(int, words[1::2])))
    subs, lists = read_subs(outp, pop.nsub)
    info["snapshot_path"] = str(snapp)
    return info, subs, lists


@pytest.mark.parametrize("merge_trapped", [False, True])
def test_recursive_unbind_matches_oracle_recursion(tmp_path, merge_trapped):
    assert os.path.exists(BIN), "build the C++ test driver first (make -C hbt-herons_amd)"
    H = synth.hierarchy(seed=5)
    info, gpu, lists = run_driver(H, tmp_path, merge_trapped)
    # the level-batched result equals the depth-first recursion, and the case is not trivial
    assert info["identical"] == 1 and info["truncate_ok"] == 1 and info["stats_ok"] == 1
    # the snapshot-level driver (RefineSnapshot: heads appended to the central, field
    # subhaloes on their own, truncation) equals SubhaloSnapshot_t::RefineParticles' flow
    assert info["snapshot_identical"] == 1 and info["snapshot_groups"] >= 1 and info["field"] >= 1
    assert info["sunk"] >= 2
    assert info["levels"] >= 4 and info["merged"] >= 2
    ref = HO.run(H, O.default_params(), 7, merge_trapped)
    # the same mergers (sink track, snapshots) in every subhalo
    assert np.array_equal(gpu["sink"], [nd.sink for nd in ref])
    assert np.array_equal(gpu["snap_sink"], [nd.snap_sink for nd in ref])
    assert np.array_equal(gpu["snap_death"], [nd.snap_death for nd in ref])
    # bound sets: FP32 walk vs the oracle's FP64 walk, <= 1e-3 of the bound mass (equal masses)
    diff = tot = 0
    for s, nd in enumerate(ref):
        g = set(lists[s][:gpu["nbound"][s]].tolist())
        r = set(nd.pid[:nd.nbound].tolist())
        diff += len(g ^ r)
        tot += len(r)
        assert abs(int(gpu["nbound"][s]) - nd.nbound) <= max(2, nd.nbound // 200), s
        # sources gathered from the children: the same particles
        assert set(lists[s].tolist()) == set(nd.pid.tolist()), s
    assert diff <= 1e-3 * tot
    # merged subhaloes hand their bound particles over and keep none
    if merge_trapped:
        sunk = [s for s, nd in enumerate(ref) if nd.sink != -1 and H.sink[s] == -1]
        assert sunk and all(gpu["np"][s] == 0 for s in sunk)


def test_repeated_child_is_rejected(tmp_path):
    """A child listed twice is rejected before anything runs."""
    assert os.path.exists(BIN)
    H = synth.hierarchy(seed=6, nhost=1)
    H.children[H.roots[0]].append(H.children[H.roots[0]][0])
    coff = np.concatenate([[0], np.cumsum([len(c) for c in H.children])]).astype(np.int64)
    child = np.array([c for cs in H.children for c in cs], np.int64)
    pop = H.pop
    inp, outp = tmp_path / "bad.bin", tmp_path / "bad_out.bin"
    with open(inp, "wb") as f:
        np.array([pop.nsub, pop.n, len(H.roots), len(child), 0, 0], np.int64).tofile(f)
        pop.offsets.astype(np.int64).tofile(f)
        for a in (pop.pos, pop.vel, pop.mass):
            np.ascontiguousarray(a, np.float32).tofile(f)
        pop.pid.astype(np.int64).tofile(f)
        for a in (pop.ref_pos, pop.ref_vel):
            np.ascontiguousarray(a, np.float32).tofile(f)
        for a in (H.track, H.sink, coff, child, np.asarray(H.roots, np.int64)):
            np.asarray(a, np.int64).tofile(f)
    r = subprocess.run([BIN, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 4 and "reached twice" in r.stderr


def with_newborn(H, seed=9, n=1500):
    """H plus one subhalo born this snapshot (after every old one, outside every
    hierarchy): a compact NFW clump with 10% interlopers."""
    extra = synth.single_nfw(seed, n_halo=n, n_inter=n // 10, rs=0.01, centre=(5.0, 40.0, 40.0))
    pop = H.pop
    npos = np.concatenate([pop.pos, extra.pos])
    nvel = np.concatenate([pop.vel, extra.vel])
    nmass = np.concatenate([pop.mass, extra.mass])
    noff = np.concatenate([pop.offsets, [pop.offsets[-1] + extra.n]])
    npid = np.concatenate([pop.pid, np.arange(extra.n) + pop.pid.max() + 1])
    p2 = synth.Population(npos, nvel, nmass, noff, npid, ref_pos=np.concatenate([pop.ref_pos, extra.ref_pos]),
                          ref_vel=np.concatenate([pop.ref_vel, extra.ref_vel]))
    return synth.Hierarchy(p2, [list(c) for c in H.children] + [[]], list(H.roots),
                           np.concatenate([H.track, [len(H.track)]]), np.concatenate([H.sink, [-1]]))


@pytest.mark.parametrize("merge_trapped", [False, True])
def test_refine_snapshot_matches_oracle_flow(tmp_path, merge_trapped):
    """hbtx::RefineSnapshot against SubhaloSnapshot_t::RefineParticles restated over
    the CPU oracle (tests/hierarchy_oracle.py::refine_snapshot): the heads appended
    to the central of each host halo, the field subhaloes and the subhalo born this
    snapshot unbound on their own, every source truncated (SourceSubRelaxFactor 3),
    and PrintSubhaloStatistics' counts (src/subhalo_unbind.cpp:855-919, :940-980)."""
    H0 = synth.hierarchy(seed=5)
    num_old = H0.pop.nsub
    H = with_newborn(H0)
    info, _, _ = run_driver(H, tmp_path, merge_trapped, num_sub_old=num_old)
    assert info["snapshot_identical"] == 1
    gpu, lists = read_subs(info["snapshot_path"], H.pop.nsub)
    roots = list(H.roots)
    # the driver's host haloes: roots in pairs (the first the central, both heads);
    # an odd last root and its hierarchy are field subhaloes
    groups = [[roots[r], roots[r + 1]] for r in range(0, len(roots) - 1, 2)]
    field = []
    if len(roots) % 2:
        st = [roots[-1]]
        while st:
            x = st.pop()
            field.append(x)
            st.extend(H.children[x])
    ref = HO.refine_snapshot(H, O.default_params(), 7, merge_trapped, groups, groups, field, num_old)
    assert np.array_equal(gpu["sink"], [nd.sink for nd in ref])
    assert np.array_equal(gpu["snap_sink"], [nd.snap_sink for nd in ref])
    assert np.array_equal(gpu["snap_death"], [nd.snap_death for nd in ref])
    diff = tot = 0
    for s, nd in enumerate(ref):
        assert abs(int(gpu["nbound"][s]) - nd.nbound) <= max(2, nd.nbound // 200), s
        g = set(lists[s][:gpu["nbound"][s]].tolist())
        r = set(nd.pid[:nd.nbound].tolist())
        diff += len(g ^ r)
        tot += len(r)
        # truncated sources: Nbound x 3 particles, bound first, then the latest removals
        assert abs(len(lists[s]) - len(nd.pid)) <= 3 * max(2, nd.nbound // 200), s
        common = len(set(lists[s].tolist()) & set(nd.pid.tolist()))
        assert common >= len(nd.pid) - 3 * max(2, nd.nbound // 200), s
    assert diff <= 1e-3 * tot
    # the subhalo born this snapshot was unbound: bound, with its own source truncated
    assert int(gpu["nbound"][-1]) > 0 and len(lists[-1]) <= 3 * int(gpu["nbound"][-1])
    births = np.zeros(H.pop.nsub, np.int64)
    births[num_old:] = 7
    c_gpu = HO.subhalo_statistics(gpu["snap_death"], gpu["snap_sink"], births, 7)
    c_ref = HO.subhalo_statistics([nd.snap_death for nd in ref], [nd.snap_sink for nd in ref], births, 7)
    assert c_gpu == c_ref
    assert c_gpu[0] >= 2 and c_gpu[2] == 1  # mergers happened; one new subhalo
