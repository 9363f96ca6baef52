"""CPU model of the tree walk's batched stack (hbt-herons_amd/csrc/walk.h,
tree_walk_stack) over the oracle's depth-first trees: the batches' entries
(sibling chain, reached lanes) reproduce every lane's own walk exactly, and the
stack stays within the bound its capacity rests on (at most
WSTACK_CAP - WSTACK_HEADROOM + 24 entries) on the deepest trees the keys allow,
with every cell opened (theta = 0) and at theta = 0.7.  The GPU kernel itself is
checked against the FP64 direct sum on the same kind of trees in
tests/test_gpu_tree.py."""
import os
import re

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _walk_constants():
    src = open(os.path.join(ROOT, "hbt-herons_amd", "csrc", "walk.h")).read()
    cap = int(re.search(r"constexpr int WSTACK_CAP = (\d+);", src).group(1))
    head = int(re.search(r"constexpr int WSTACK_HEADROOM = (\d+);", src).group(1))
    return cap, head


def _per_lane(kind, com, crit2, skip, P, valid):
    """Each lane's own stackless preorder walk: the records it accepts."""
    tot = len(kind)
    out = []
    for t in range(len(P)):
        acc = set()
        if valid[t]:
            o = 0
            while o < tot:
                r2 = ((com[o] - P[t]) ** 2).sum()
                if kind[o] == 1 and r2 < crit2[o]:
                    o += 1
                else:
                    acc.add(o)
                    o = skip[o] if kind[o] == 1 else o + 1
        out.append(acc)
    return out


def _stack_walk(kind, com, crit2, skip, P, valid, cap, head):
    """The kernel's batches: pop up to 64 entries (fewer when the room above the
    headroom is short, at least one), visit their records, push each entry's
    chain rest and, for an opened cell, its children's chain."""
    tot = len(kind)
    acc = [set() for _ in range(len(P))]
    stk = [(0, tot, valid.copy())]
    most = 1
    while stk:
        top = len(stk)
        nb = min(64, top, max(1, cap - head - top))
        batch = stk[top - nb:]
        del stk[top - nb:]
        for o, e, mk in batch:
            r2 = ((com[o] - P) ** 2).sum(1)
            opened = mk & (r2 < crit2[o])
            for t in np.nonzero(mk & ~opened)[0]:
                acc[t].add(o)
            nx = skip[o] if kind[o] == 1 else o + 1
            if nx < e:
                stk.append((nx, e, mk))
            if kind[o] == 1 and opened.any():
                stk.append((o + 1, nx, opened))
        most = max(most, len(stk))
    return acc, most


def _deep_population(rng):
    u = rng.uniform(0, 1, (1500, 3)).astype(np.float32)
    chains = (u[:200] + np.float32(3e-6)).astype(np.float32)
    clump = np.repeat(np.array([[0.25, 0.5, 0.75]], np.float32), 100, axis=0)
    return np.concatenate([u, chains, clump])


@pytest.mark.parametrize("theta", [0.0, 0.7])
def test_stack_walk_model_equals_lane_walks_within_capacity(theta):
    cap, head = _walk_constants()
    rng = np.random.default_rng(21)
    pos = _deep_population(rng)
    mass = rng.uniform(0.5, 1.5, len(pos)).astype(np.float32)
    t = O.tree_dfs(pos, mass, dict(theta=max(theta, 1e-30), softening=0.001))
    kind, com, skip = t["kind"], t["com"].astype(np.float64), t["skip"]
    crit2 = np.where(kind == 1, np.inf if theta == 0 else (t["len"].astype(np.float64) ** 2) / theta ** 2, 0.0)
    assert t["depth"].max() >= 19  # the chains reach near the key depth limit
    tgt = com[kind == 0]
    most_all = 0
    for g in rng.choice(len(tgt) // 64, 6, replace=False):
        P = tgt[64 * g: 64 * g + 64]
        valid = np.ones(64, bool)
        valid[rng.integers(0, 64, 4)] = False
        got, most = _stack_walk(kind, com, crit2, skip, P, valid, cap, head)
        ref = _per_lane(kind, com, crit2, skip, P, valid)
        assert got == ref
        most_all = max(most_all, most)
    assert most_all <= cap - head + 24, most_all
