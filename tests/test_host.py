"""Host-side checks that need no GPU: the C-ABI library, parameters, generators."""
import ctypes as C

import numpy as np

import hbtx
import synth


def test_library_exports_every_header_symbol():
    L = hbtx.lib()
    syms = hbtx.header_symbols()
    assert len(syms) >= 19
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_default_params_follow_reference_defaults():
    # src/config_parser.h:98-172 and derived G (src/config_parser.cpp:200)
    p = hbtx.Params()
    hbtx.lib().hbtx_default_params(C.byref(p))
    assert p.G == 43.0071
    assert p.theta == np.float32(0.45)
    assert p.bound_mass_precision == np.float32(0.995)
    assert p.max_sample_size == 1000
    assert p.refine_most_bound == 1
    assert p.bound_fraction_center_refinement == np.float32(0.1)
    assert p.min_num_part == 20 and p.min_num_tracer == 10
    assert p.tracer_mask == (1 << 1) | (1 << 4)
    assert p.donot_subsample_mask == (1 << 5)
    assert p.max_phys_softening == -1.0
    assert p.periodic == 1
    assert p.upscale_per_type == 1
    assert p.sub_core_size_min == 20


def test_context_creation_fails_cleanly_or_succeeds():
    h = C.c_void_p()
    rc = hbtx.lib().hbtx_create(C.byref(h), 0)
    if rc == hbtx.OK:
        hbtx.lib().hbtx_destroy(h)
    else:
        assert rc == hbtx.E_HIP
    # invalid device index is an error, never a crash
    rc2 = hbtx.lib().hbtx_create(C.byref(h), 1 << 20)
    assert rc2 in (hbtx.E_INVALID, hbtx.E_HIP)


def test_null_arguments_rejected():
    assert hbtx.lib().hbtx_batch_run(None, None, 0, None) == hbtx.E_INVALID
    assert hbtx.lib().hbtx_create(None, 0) == hbtx.E_INVALID


def test_power_law_counts():
    rng = np.random.default_rng(0)
    n = synth.draw_counts(rng, 200000)
    assert n.min() >= 20 and n.max() <= 1_000_000
    # dN/dN ~ N^-1.9: the fraction above 1000 is (1000^-0.9 - 1e6^-0.9)/(20^-0.9 - 1e6^-0.9)
    frac = np.mean(n > 1000)
    exp = (1000 ** -0.9 - 1e6 ** -0.9) / (20 ** -0.9 - 1e6 ** -0.9)
    assert abs(frac - exp) < 0.1 * exp


def test_nfw_sampling_follows_mass_profile():
    pop = synth.single_nfw(0, n_halo=20000, n_inter=0)
    r = np.sqrt(((pop.pos - pop.pos.mean(0)) ** 2).sum(1))
    rs, c = 0.02, 10.0
    for x in [0.5, 1.0, 3.0]:
        frac = np.mean(r < x * rs)
        assert abs(frac - synth._mu(x) / synth._mu(c)) < 0.02


def test_population_target_size():
    pop = synth.population(seed=3, target_particles=300000, nmax=20000)
    assert 300000 <= pop.n < 330000
    assert np.all(np.diff(pop.offsets) >= 20)


def test_layout_check_rejects_subhaloes_beyond_the_walk_span():
    # the walk keeps a tree's position as an int32 byte offset (16 bytes a particle
    # record): a subhalo above HBTX_MAX_SUBHALO_PARTICLES is a capacity error before
    # any upload, the largest accepted one passes
    txt = open(hbtx.HEADER).read()
    assert "HBTX_MAX_SUBHALO_PARTICLES ((int64_t)(1 << 27) - 64)" in txt
    nmax = (1 << 27) - 64
    assert 16 * nmax < 2 ** 31
    rc, msg = hbtx.check_layout(nmax, [0, nmax])
    assert rc == hbtx.OK and msg == ""
    rc, msg = hbtx.check_layout(nmax + 11, [0, 10, nmax + 11])
    assert rc == hbtx.E_CAPACITY and "subhalo 1" in msg and "2 GiB" in msg
    rc, _ = hbtx.check_layout(2 * nmax, [0, nmax, 2 * nmax])
    assert rc == hbtx.OK


def test_layout_check_rejects_malformed_offsets():
    assert hbtx.check_layout(10, [0, 5, 10])[0] == hbtx.OK
    assert hbtx.check_layout(10, [0, 6, 5, 10])[0] == hbtx.E_INVALID   # decreasing
    assert hbtx.check_layout(10, [1, 5, 10])[0] == hbtx.E_INVALID      # not starting at 0
    assert hbtx.check_layout(10, [0, 5, 9])[0] == hbtx.E_INVALID       # not ending at n
    assert hbtx.check_layout(1 << 29, [0, 1 << 29])[0] == hbtx.E_INVALID  # batch limit
    assert hbtx.check_layout(0, [0])[0] == hbtx.E_INVALID              # no subhalo
