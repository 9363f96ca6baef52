"""Segmented sort primitive on the GPU vs numpy (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_sorted(keys, vals, k2, v2, off):
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        exp = np.sort(keys[a:b], kind="stable")
        assert np.array_equal(k2[a:b], exp), f"segment {s} keys differ"
        # values travel with their keys: (key, val) multiset preserved
        pairs_in = sorted(zip(keys[a:b].tolist(), vals[a:b].tolist()))
        pairs_out = sorted(zip(k2[a:b].tolist(), v2[a:b].tolist()))
        assert pairs_in == pairs_out, f"segment {s} payload lost"


@pytest.mark.parametrize("sizes", [[0, 1, 2, 3, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 10000, 70001],
                                   [100000], [5] * 200, [20, 5000, 0, 3, 131073]])
def test_segmented_sort_matches_numpy(ctx, sizes):
    rng = np.random.default_rng(len(sizes))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    # many duplicates in some segments
    keys[::7] = keys[::7] % 16
    vals = np.arange(n, dtype=np.int32)
    k2, v2 = ctx.segmented_sort(keys, vals, off)
    _check_sorted(keys, vals, k2, v2, off)


@pytest.mark.parametrize("bits", [18, 30, 36, 63])
def test_segmented_sort_key_widths(ctx, bits):
    """Narrow keys take fewer digit passes (9-bit digits at 18 and 36 bits)."""
    rng = np.random.default_rng(bits)
    sizes = [3000, 9000, 1, 40000, 2048, 2049]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    keys = rng.integers(0, 2 ** bits, n, dtype=np.uint64)
    keys[::5] = keys[::5] & np.uint64(0xff)  # shared low digits
    keys[1::9] = keys[0:n - 1:9]  # duplicates
    vals = np.arange(n, dtype=np.int32)
    k2, v2 = ctx.segmented_sort(keys, vals, off)
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        order = np.argsort(keys[a:b], kind="stable")
        assert np.array_equal(k2[a:b], keys[a:b][order])
        if b - a > 2048:  # radix path: stable
            assert np.array_equal(v2[a:b], vals[a:b][order])


def test_radix_sort_is_stable_for_large_segments(ctx):
    rng = np.random.default_rng(7)
    n = 50000
    keys = rng.integers(0, 50, n).astype(np.uint64)
    vals = np.arange(n, dtype=np.int32)
    k2, v2 = ctx.segmented_sort(keys, vals, np.array([0, n], np.int64))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v2, vals[order])


def test_many_segments_just_over_the_lds_size_fresh_context():
    """Regression: every segment longer than the LDS sort takes a partial radix tile,
    so segments just over 2048 elements need almost two tiles per 2048 elements; the
    tile list and its histograms are sized for that (a fresh context, so no earlier
    call has grown the buffers)."""
    import hbtx
    c = hbtx.Context(0)
    try:
        rng = np.random.default_rng(21)
        sizes = rng.integers(2049, 2300, 1000)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        n = int(off[-1])
        print(f"segmented sort: {len(sizes)} segments, {n} elements", flush=True)
        keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        vals = np.arange(n, dtype=np.int32)
        k2, v2 = c.segmented_sort(keys, vals, off)
        _check_sorted(keys, vals, k2, v2, off)
    finally:
        c.close()


def test_direct_sums_of_subhaloes_just_over_the_lds_size_fresh_context():
    """Regression: 300 subhaloes of 2048 members plus interlopers (about 2250
    sources each) in direct-sum mode; batched and per-subhalo runs agree."""
    import hbtx
    import synth
    c = hbtx.Context(0)
    try:
        pop = synth.population(seed=33, nsub=300, nmin=2048, nmax=2048)
        print(f"direct-sum population: {pop.nsub} subhaloes, {pop.n} particles", flush=True)
        prm = hbtx.workload_params(direct_cutoff=1 << 40)
        a = c.unbind(pop, prm, hbtx.BATCHED)
        b = c.unbind(pop, prm, hbtx.PLAIN)
        assert np.array_equal(a["result"]["nbound"], b["result"]["nbound"])
        assert np.array_equal(a["order"], b["order"])
    finally:
        c.close()


def _bitonic_reference(keys, vals):
    """The sort network of the in-lane and LDS sorts, run in numpy: the segment is
    padded with all-ones keys to the next power of two P, and for k = 2..P,
    j = k/2..1 the pair (i, i^j), i < i^j, is exchanged when (a > c) equals
    (i & k) == 0."""
    n = len(keys)
    P = 1
    while P < n:
        P <<= 1
    k_ = np.full(P, np.iinfo(np.uint64).max, np.uint64)
    v_ = np.zeros(P, np.int32)
    k_[:n], v_[:n] = keys, vals
    i = np.arange(P)
    k = 2
    while k <= P:
        j = k >> 1
        while j > 0:
            ixj = i ^ j
            lo = i[ixj > i]
            hi = lo ^ j
            a, c = k_[lo], k_[hi]
            sw = (a > c) == ((lo & k) == 0)
            ls, hs = lo[sw], hi[sw]
            k_[ls], k_[hs] = k_[hs].copy(), k_[ls].copy()
            v_[ls], v_[hs] = v_[hs].copy(), v_[ls].copy()
            j >>= 1
        k <<= 1
    return k_[:n], v_[:n]


def test_small_sorts_follow_the_bitonic_network(ctx):
    """Segments of up to 2048 keys: keys and the order of tied keys' values equal
    the bitonic network's, whichever barriers separate its stages on the GPU."""
    rng = np.random.default_rng(11)
    sizes = [33, 64, 65, 100, 127, 128, 129, 192, 255, 256, 257, 300, 512, 513, 700, 1024, 1500, 2047, 2048]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    keys = rng.integers(0, 8, n).astype(np.uint64)  # many ties
    vals = np.arange(n, dtype=np.int32)
    k2, v2 = ctx.segmented_sort(keys, vals, off)
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        ek, ev = _bitonic_reference(keys[a:b], vals[a:b])
        assert np.array_equal(k2[a:b], ek), sizes[s]
        assert np.array_equal(v2[a:b], ev), sizes[s]
