"""Pins the CPU restatement of the hierarchical unbinding (tests/hierarchy_oracle.py)
on the CPU: invariants of Subhalo_t::RecursiveUnbind (src/subhalo_unbind.cpp:810-840)
and closed forms of the merger test (src/subhalo_merge.cpp:15-91, :183-239)."""
import numpy as np

import hierarchy_oracle as HO
import pyoracle as O
import synth


def test_flat_hierarchy_is_independent_unbinding():
    H = synth.hierarchy(seed=8, nhost=2)
    for cs in H.children:
        cs.clear()
    H.roots = list(range(H.pop.nsub))
    nodes = HO.run(H, O.default_params(), 7, False)
    ref = O.unbind_batch(H.pop, O.default_params(), nthreads=1)
    for s, nd in enumerate(nodes):
        a = H.pop.offsets[s]
        assert nd.nbound == ref["result"]["nbound"][s]
        assert np.array_equal(nd.pid, H.pop.pid[a + ref["order"][a:H.pop.offsets[s + 1]]])


def test_parent_sources_gain_the_children_tails():
    H = synth.hierarchy(seed=9, nhost=1)
    nodes = HO.run(H, O.default_params(), 7, False)
    pop = H.pop
    for s, nd in enumerate(nodes):
        own = set(pop.pid[pop.offsets[s]:pop.offsets[s + 1]].tolist())
        tails = set()
        for c in H.children[s]:
            tails |= set(nodes[c].pid[nodes[c].nbound:].tolist())
        assert set(nd.pid.tolist()) == own | tails
        assert len(nd.pid) == len(own) + sum(len(nodes[c].pid) - nodes[c].nbound for c in H.children[s])


def _node(core_pos, core_vel, sr=1.0, sv=1.0, nbound=10, track=0):
    nd = HO.Node(np.zeros((nbound, 3), np.float32), np.zeros((nbound, 3), np.float32), np.ones(nbound, np.float32),
                 np.arange(nbound), np.zeros(3), np.zeros(3), track, -1)
    nd.core_pos, nd.core_vel = np.array(core_pos, np.float32), np.array(core_vel, np.float32)
    nd.core_sr, nd.core_sv, nd.nbound = sr, sv, nbound
    return nd


def test_phase_space_distance_closed_form():
    ref = _node([0, 0, 0], [0, 0, 0], sr=0.5, sv=10.0)
    sub = _node([0.3, 0.4, 0.0], [6.0, 0.0, 8.0])
    # |dx| / sigma_r + |dv| / sigma_v = 0.5 / 0.5 + 10 / 10
    assert abs(HO.phase_space_distance(sub, ref) - 2.0) < 1e-6


def test_merge_rules():
    ref = _node([0, 0, 0], [0, 0, 0], track=3)
    near = _node([0.1, 0, 0], [0.5, 0, 0], nbound=6, track=4)
    far = _node([5, 0, 0], [0, 0, 0], track=5)
    trapped = _node([0, 0, 0], [0, 0, 0], track=6)
    trapped.sink = 9
    orphan = _node([0, 0, 0], [0, 0, 0], nbound=1, track=7)
    nodes = [ref, near, far, trapped, orphan]
    ref.children = [1, 2, 3, 4]
    n0 = len(ref.pid)
    merged = HO.merge_recursively(nodes, 0, ref, 11, True)
    assert merged
    assert near.sink == 3 and near.snap_sink == 11 and near.snap_death == 11 and near.nbound == 0
    assert far.sink == -1 and trapped.sink == 9
    # an unresolved subhalo sinks but passes nothing and does not force a second unbinding
    assert orphan.sink == 3 and orphan.nbound == 1
    assert len(ref.pid) == n0 + 6
    # a reference with at most one bound particle takes nothing
    tiny = _node([0, 0, 0], [0, 0, 0], nbound=1)
    other = _node([0, 0, 0], [0, 0, 0])
    tiny.children = [1]
    assert not HO.merge_recursively([tiny, other], 0, tiny, 11, True) and other.sink == -1
