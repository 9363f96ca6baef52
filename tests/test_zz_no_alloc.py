"""hbtx_direct_potential_noalloc: the exact direct sum with its staging in the
library's file-scope device array (no allocation), checked against
hbtx_direct_potential (the same records, the same kernel: identical results)
on two calls in a row, the second smaller, plus the size limit.

Device buffers come from the HIP runtime the library itself is linked to
(hipMalloc / hipMemcpy through ctypes), so the inputs and the output are device
pointers as the entry point requires."""
import ctypes as C

import numpy as np
import pytest

import hbtx

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevBuf:
    def __init__(self, L, host):
        self.L = L
        self.nbytes = host.nbytes
        self.p = C.c_void_p()
        assert L.hipMalloc(C.byref(self.p), max(self.nbytes, 16)) == 0
        if self.nbytes:
            assert L.hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def points(rng, n, spread=0.05):
    pos = (rng.normal(0, spread, (n, 3)) + 2.0).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32) * np.float32(0.01)
    return pos, mass


def noalloc(ctx, L, prm, sp, sm, tp, tm):
    bufs = [DevBuf(L, a) for a in (sp, sm, tp)] + ([DevBuf(L, tm)] if tm is not None else [])
    out = DevBuf(L, np.zeros(len(tp), np.float64))
    free0 = C.c_size_t()
    total = C.c_size_t()
    assert L.hipDeviceSynchronize() == 0
    assert L.hipMemGetInfo(C.byref(free0), C.byref(total)) == 0
    rc = hbtx.lib().hbtx_direct_potential_noalloc(ctx._h, C.byref(hbtx.make_params(**prm)), len(sp), bufs[0].p,
                                                  bufs[1].p, len(tp), bufs[2].p,
                                                  bufs[3].p if tm is not None else None, out.p)
    assert rc == hbtx.OK, hbtx.lib().hbtx_last_error(ctx._h)
    assert hbtx.lib().hbtx_device_sync(ctx._h) == hbtx.OK
    free1 = C.c_size_t()
    assert L.hipMemGetInfo(C.byref(free1), C.byref(total)) == 0
    res = out.get(np.zeros(len(tp), np.float64))
    for b in bufs + [out]:
        b.free()
    return res, free0.value - free1.value


@pytest.mark.parametrize("periodic", [False, True])
def test_noalloc_direct_potential_matches_the_allocating_entry_point(ctx, periodic):
    L = hip()
    rng = np.random.default_rng(7)
    prm = dict(scale_factor=0.5, hz=100.0, max_phys_softening=0.0004, theta=0.7)
    if periodic:
        prm.update(periodic=1, box_size=4.0)
    # a first call loads this process's kernels (the runtime may set up launch
    # resources once); then two calls in a row, the second smaller: the array keeps
    # the previous call's contents, which must not leak into the next, and device
    # memory in use does not change during either call
    for k, (ns, nt, with_mass) in enumerate(((300, 200, True), (6000, 4000, True), (1500, 900, False))):
        sp, sm = points(rng, ns)
        if with_mass:
            tp, tm = sp[:nt].copy(), sm[:nt].copy()  # targets that are sources: self term removed
        else:
            tp, tm = points(rng, nt, 0.08)[0], None
        ref = ctx.direct_potential(sp, sm, tp, tm, params=hbtx.workload_params(**prm))
        got, used = noalloc(ctx, L, hbtx.workload_params(**prm), sp, sm, tp, tm)
        assert np.array_equal(got, ref), np.max(np.abs(got - ref) / np.abs(ref))
        if k > 0:
            assert used == 0, f"device memory in use grew by {used} bytes during the call"


def test_noalloc_direct_potential_rejects_sizes_beyond_the_array(ctx):
    nmax = 1 << 20
    txt = open(hbtx.HEADER).read()
    assert "#define HBTX_NOALLOC_MAX_POINTS (1 << 20)" in txt
    prm = C.byref(hbtx.make_params())
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    for ns, nt in ((nmax + 1, 10), (10, nmax + 1)):
        rc = hbtx.lib().hbtx_direct_potential_noalloc(ctx._h, prm, ns, dummy, dummy, nt, dummy, None, dummy)
        assert rc == hbtx.E_CAPACITY
        assert b"HBTX_NOALLOC_MAX_POINTS" in hbtx.lib().hbtx_last_error(ctx._h)
    assert hbtx.lib().hbtx_direct_potential_noalloc(ctx._h, prm, 0, dummy, dummy, 1, dummy, None,
                                                    dummy) == hbtx.E_INVALID
