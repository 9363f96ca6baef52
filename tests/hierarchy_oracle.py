"""CPU restatement of the hierarchical unbinding for the tests: the reference's
sequential recursion Subhalo_t::RecursiveUnbind (src/subhalo_unbind.cpp:810-840)
with MergeRecursively / SetMergerInformation / MergeTo / PhaseSpaceDistance
(src/subhalo_merge.cpp:15-91, :183-239) over the oracle's Subhalo_t::Unbind
(oracle/oracle.cpp, one subhalo per call).  Test infrastructure only."""
import numpy as np

import pyoracle as O
import synth


class Node:
    """The Subhalo_t fields the recursion reads and writes."""

    def __init__(self, pos, vel, mass, pid, ref_pos, ref_vel, track, sink):
        self.pos, self.vel, self.mass, self.pid = pos, vel, mass, pid
        self.com_pos, self.com_vel = np.array(ref_pos, np.float32), np.array(ref_vel, np.float32)
        self.mb_pos, self.mb_vel = self.com_pos.copy(), self.com_vel.copy()
        self.track, self.sink = int(track), int(sink)
        self.snap_sink = -1
        self.snap_death = -1
        self.nbound = 0
        self.mbound = 0.0
        self.tracer = 0
        self.mostbound_id = -1
        self.core_pos = np.zeros(3, np.float32)
        self.core_vel = np.zeros(3, np.float32)
        self.core_sr = 0.0
        self.core_sv = 0.0
        self.children = []


def nodes_from(H):
    pop = H.pop
    out = []
    for s in range(pop.nsub):
        a, b = pop.offsets[s], pop.offsets[s + 1]
        nd = Node(pop.pos[a:b].copy(), pop.vel[a:b].copy(), pop.mass[a:b].copy(), pop.pid[a:b].copy(),
                  pop.ref_pos[s], pop.ref_vel[s], H.track[s], H.sink[s])
        nd.children = list(H.children[s])
        out.append(nd)
    return out


def unbind(nd, params, snapshot_id):
    """Subhalo_t::Unbind (:463-808) of one node by the oracle, with the
    SnapshotOfDeath / Sink bookkeeping of :646-707."""
    n = len(nd.mass)
    if n == 0:
        nd.nbound, nd.mbound = 0, 0.0
        nd.core_pos, nd.core_vel = nd.mb_pos.copy(), nd.mb_vel.copy()
        nd.core_sr = nd.core_sv = 0.0
        return
    pop = synth.Population(nd.pos, nd.vel, nd.mass, [0, n], nd.pid, ref_pos=nd.com_pos[None],
                           ref_vel=nd.com_vel[None])
    out = O.unbind_batch(pop, params, nthreads=1, mb_pos=nd.mb_pos[None], mb_vel=nd.mb_vel[None])
    order = out["order"]
    r = out["result"][0]
    nd.pos, nd.vel, nd.mass, nd.pid = nd.pos[order], nd.vel[order], nd.mass[order], nd.pid[order]
    nd.nbound, nd.mbound = int(r["nbound"]), float(r["mbound"])
    nd.com_pos, nd.com_vel = r["com_pos"].copy(), r["com_vel"].copy()
    nd.mb_pos, nd.mb_vel = r["mostbound_pos"].copy(), r["mostbound_vel"].copy()
    nd.tracer = int(r["tracer_index"])
    nd.core_pos, nd.core_vel = r["core_pos"].copy(), r["core_vel"].copy()
    nd.core_sr, nd.core_sv = float(r["core_sigma_r"]), float(r["core_sigma_v"])
    if r["disrupted"]:
        if nd.snap_death == -1:
            nd.snap_death = snapshot_id
    else:
        if nd.snap_death != -1:
            nd.snap_death = -1
        if nd.sink != -1:
            nd.snap_sink, nd.sink = -1, -1
        nd.mostbound_id = int(nd.pid[nd.tracer])


def phase_space_distance(sub, ref):
    dx = np.float32(np.sqrt(np.sum((ref.core_pos.astype(np.float32) - sub.core_pos.astype(np.float32)) ** 2,
                                   dtype=np.float32)))
    dv = np.float32(np.sqrt(np.sum((ref.core_vel.astype(np.float32) - sub.core_vel.astype(np.float32)) ** 2,
                                   dtype=np.float32)))
    return np.float32(dx / np.float32(ref.core_sr)) + np.float32(dv / np.float32(ref.core_sv))


def merge_to(sub, host):
    if sub.nbound <= 1:
        return
    nb = sub.nbound
    host.pos = np.concatenate([host.pos, sub.pos[:nb]])
    host.vel = np.concatenate([host.vel, sub.vel[:nb]])
    host.mass = np.concatenate([host.mass, sub.mass[:nb]])
    host.pid = np.concatenate([host.pid, sub.pid[:nb]])
    t = sub.tracer
    sub.mostbound_id = int(sub.pid[t])
    sub.mb_pos, sub.mb_vel = sub.pos[t].copy(), sub.vel[t].copy()
    sub.com_pos, sub.com_vel = sub.mb_pos.copy(), sub.mb_vel.copy()
    sub.nbound = 0
    sub.pos, sub.vel = sub.pos[:0], sub.vel[:0]
    sub.mass, sub.pid = sub.mass[:0], sub.pid[:0]
    sub.mbound = 0.0


def merge_recursively(nodes, s, ref, snapshot_id, merge_trapped):
    merged = False
    if ref.nbound <= 1:
        return merged
    S = nodes[s]
    for c in S.children:
        merged = merge_recursively(nodes, c, ref, snapshot_id, merge_trapped) | merged
    if S.sink == -1 and S is not ref:
        if phase_space_distance(S, ref) < 2.0:
            S.snap_sink = snapshot_id
            S.sink = ref.track
            if S.snap_death == -1:
                S.snap_death = snapshot_id
            merged = (S.nbound > 1) | merged
            if merge_trapped:
                merge_to(S, ref)
    return merged


def recursive_unbind(nodes, s, params, snapshot_id, merge_trapped):
    S = nodes[s]
    for c in S.children:
        recursive_unbind(nodes, c, params, snapshot_id, merge_trapped)
        C = nodes[c]
        nb = C.nbound
        S.pos = np.concatenate([S.pos, C.pos[nb:]])
        S.vel = np.concatenate([S.vel, C.vel[nb:]])
        S.mass = np.concatenate([S.mass, C.mass[nb:]])
        S.pid = np.concatenate([S.pid, C.pid[nb:]])
    unbind(S, params, snapshot_id)
    merged = merge_recursively(nodes, s, S, snapshot_id, merge_trapped)
    if merge_trapped and merged:
        unbind(S, params, snapshot_id)


def run(H, params, snapshot_id, merge_trapped):
    nodes = nodes_from(H)
    for r in H.roots:
        recursive_unbind(nodes, r, params, snapshot_id, merge_trapped)
    return nodes


def truncate_source(nd, relax=3.0):
    """Subhalo_t::TruncateSource (src/subhalo_unbind.cpp:842-852)."""
    ns = nd.nbound if nd.nbound <= 1 else int(nd.nbound * relax)
    ns = min(ns, len(nd.mass))
    nd.pos, nd.vel, nd.mass, nd.pid = nd.pos[:ns], nd.vel[:ns], nd.mass[:ns], nd.pid[:ns]


def refine_snapshot(H, params, snapshot_id, merge_trapped, groups, heads, field, num_sub_old, relax=3.0):
    """SubhaloSnapshot_t::RefineParticles (src/subhalo_unbind.cpp:855-919, the
    exclusive-mass build) restated sequentially: per host halo with members, the
    central's nested list gets the other heads for its RecursiveUnbind; then every
    field subhalo and every subhalo born this snapshot (index >= num_sub_old) is
    unbound on its own; then every source is truncated."""
    nodes = nodes_from(H)
    for h, members in enumerate(groups):
        if len(members) == 0:
            continue
        central = nodes[members[0]]
        old = len(central.children)
        central.children = central.children + list(heads[h][1:])
        recursive_unbind(nodes, members[0], params, snapshot_id, merge_trapped)
        central.children = central.children[:old]
    for s in field:
        unbind(nodes[s], params, snapshot_id)
    for s in range(num_sub_old, len(nodes)):
        unbind(nodes[s], params, snapshot_id)
    for nd in nodes:
        truncate_source(nd, relax)
    return nodes


def subhalo_statistics(snap_death, snap_sink, snap_birth, snapshot_id):
    """SubhaloSnapshot_t::PrintSubhaloStatistics' counts (src/subhalo_unbind.cpp:940-980):
    (sunk, disrupted, new, fake)."""
    d, k, b = np.asarray(snap_death), np.asarray(snap_sink), np.asarray(snap_birth)
    sunk = int(np.sum((d == snapshot_id) & (k == snapshot_id)))
    disrupted = int(np.sum((b < snapshot_id) & (d == snapshot_id) & (k == -1)))
    fake = int(np.sum((b == snapshot_id) & (d == snapshot_id)))
    new = int(np.sum((b == snapshot_id) & (d == -1)))
    return sunk, disrupted, new, fake
