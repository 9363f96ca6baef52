"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker or the reported CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")


class Params(C.Structure):
    _fields_ = [
        ("G", C.c_double),
        ("scale_factor", C.c_float),
        ("hz", C.c_float),
        ("softening", C.c_float),
        ("max_phys_softening", C.c_float),
        ("theta", C.c_float),
        ("bound_mass_precision", C.c_float),
        ("max_sample_size", C.c_int64),
        ("refine_most_bound", C.c_int32),
        ("bound_fraction_center_refinement", C.c_float),
        ("min_num_part", C.c_int64),
        ("min_num_tracer", C.c_int64),
        ("tracer_mask", C.c_int32),
        ("donot_subsample_mask", C.c_int32),
        ("periodic", C.c_int32),
        ("box_size", C.c_float),
        ("thermal", C.c_int32),
        ("dm_only", C.c_int32),
        ("upscale_per_type", C.c_int32),
        ("sub_core_size_factor", C.c_float),
        ("sub_core_size_min", C.c_int64),
    ]


class Result(C.Structure):
    _fields_ = [
        ("nbound", C.c_int64),
        ("mbound", C.c_float),
        ("nbound_type", C.c_int64 * 6),
        ("mbound_type", C.c_float * 6),
        ("com_pos", C.c_float * 3),
        ("com_vel", C.c_float * 3),
        ("mostbound_pos", C.c_float * 3),
        ("mostbound_vel", C.c_float * 3),
        ("spec_pot", C.c_float),
        ("spec_kin", C.c_float),
        ("spec_am", C.c_float * 3),
        ("tracer_index", C.c_int64),
        ("mostbound_id", C.c_int64),
        ("core_pos", C.c_float * 3),
        ("core_vel", C.c_float * 3),
        ("core_sigma_r", C.c_float),
        ("core_sigma_v", C.c_float),
        ("iterations", C.c_int32),
        ("disrupted", C.c_int32),
    ]


RESULT_DTYPE = np.dtype([
    ("nbound", np.int64), ("mbound", np.float32), ("nbound_type", np.int64, 6),
    ("mbound_type", np.float32, 6), ("com_pos", np.float32, 3), ("com_vel", np.float32, 3),
    ("mostbound_pos", np.float32, 3), ("mostbound_vel", np.float32, 3), ("spec_pot", np.float32),
    ("spec_kin", np.float32), ("spec_am", np.float32, 3), ("tracer_index", np.int64),
    ("mostbound_id", np.int64), ("core_pos", np.float32, 3), ("core_vel", np.float32, 3),
    ("core_sigma_r", np.float32), ("core_sigma_v", np.float32), ("iterations", np.int32),
    ("disrupted", np.int32)], align=True)
assert RESULT_DTYPE.itemsize == C.sizeof(Result)


def default_params(**kw):
    """Reference defaults (src/config_parser.h:98-172) with the workloads' settings."""
    p = dict(G=43.0071, scale_factor=1.0, hz=0.0, softening=0.001, max_phys_softening=-1.0,
             theta=0.7, bound_mass_precision=0.995, max_sample_size=0, refine_most_bound=1,
             bound_fraction_center_refinement=0.1, min_num_part=20, min_num_tracer=10,
             tracer_mask=(1 << 1) | (1 << 4), donot_subsample_mask=(1 << 5), periodic=0,
             box_size=100.0, thermal=0, dm_only=1, upscale_per_type=1, sub_core_size_factor=0.0,
             sub_core_size_min=20)
    p.update(kw)
    return p


def make_params(**kw):
    return Params(**default_params(**kw))


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_unbind_batch.restype = C.c_double
        L.oracle_unbind_batch.argtypes = [C.POINTER(Params), C.c_int64, P, P, P, P, P, P, C.c_int64, P, P, P,
                                          P, P, C.c_int, P, P, P, P]
        L.oracle_unbind_batch_warm.restype = C.c_double
        L.oracle_unbind_batch_warm.argtypes = [C.POINTER(Params), C.c_int64, P, P, P, P, P, P, C.c_int64, P, P,
                                               P, P, P, P, C.c_int, P, P, P, P]
        L.oracle_tree_potential.restype = C.c_int64
        L.oracle_tree_potential.argtypes = [C.POINTER(Params), C.c_int64, P, P, C.c_int64, P, P, C.c_int, P]
        L.oracle_direct_potential.restype = None
        L.oracle_direct_potential.argtypes = [C.POINTER(Params), C.c_int64, P, P, C.c_int64, P, P, C.c_int, P]
        L.oracle_tree_dfs.restype = C.c_int64
        L.oracle_tree_dfs.argtypes = [C.POINTER(Params), C.c_int64, P, P, C.c_int64, P, P, P, P, P, P, P]
        L.oracle_core_phase_space.restype = None
        L.oracle_core_phase_space.argtypes = [C.POINTER(Params), C.c_int64, P, P, P, P, P, P, P, P]
        L.oracle_reassign_targets.restype = None
        L.oracle_reassign_targets.argtypes = [C.POINTER(Params), C.c_int64, P, P, P, C.c_int64, P, C.c_int32,
                                              C.c_float, P]
        L.oracle_mass_upscale.restype = C.c_int64
        L.oracle_mass_upscale.argtypes = [C.POINTER(Params), C.c_int64, P, P, C.c_int64, C.c_float, C.c_int64,
                                          C.c_int64, P]
        L.oracle_profiles.restype = None
        L.oracle_profiles.argtypes = [C.POINTER(Params), C.c_int64, P, P, C.c_int64, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def unbind_batch(pop, params=None, nthreads=0, want_potential=False, mb_pos=None, mb_vel=None, proposal=None):
    """Run the oracle unbinding over a synth.Population. Returns dict of outputs.
    proposal (uint8 per particle) runs the warm-start restatement."""
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    n = pop.n
    order = np.zeros(n, np.int64)
    energy = np.zeros(n, np.float32)
    pot = np.zeros(n, np.float32) if want_potential else None
    res = np.zeros(pop.nsub, RESULT_DTYPE)
    ptype = None if pop.ptype is None else np.ascontiguousarray(pop.ptype, dtype=np.int32)
    common = (C.byref(prm), n, _p(pop.pos), _p(pop.vel), _p(pop.mass), _p(pop.u), _p(ptype), _p(pop.pid),
              pop.nsub, _p(pop.offsets), _p(pop.ref_pos), _p(pop.ref_vel), _p(_f32(mb_pos)), _p(_f32(mb_vel)))
    outs = (_p(order), _p(energy), _p(pot), _p(res))
    if proposal is None:
        secs = lib().oracle_unbind_batch(*common, nthreads, *outs)
    else:
        pr = np.ascontiguousarray(proposal, dtype=np.uint8)
        assert len(pr) == n
        secs = lib().oracle_unbind_batch_warm(*common, _p(pr), nthreads, *outs)
    return dict(order=order, energy=energy, potential=pot, result=res, seconds=secs)


def tree_potential(src_pos, src_mass, tgt_pos, tgt_mass=None, params=None, nthreads=0):
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    src_pos = _f32(src_pos)
    src_mass = _f32(src_mass)
    tgt_pos = _f32(tgt_pos)
    tgt_mass = _f32(tgt_mass)
    out = np.zeros(len(tgt_pos), np.float64)
    ncell = lib().oracle_tree_potential(C.byref(prm), len(src_pos), _p(src_pos), _p(src_mass), len(tgt_pos),
                                        _p(tgt_pos), _p(tgt_mass), nthreads, _p(out))
    return out, ncell


def direct_potential(src_pos, src_mass, tgt_pos, tgt_mass=None, params=None, nthreads=0):
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    src_pos = _f32(src_pos)
    src_mass = _f32(src_mass)
    tgt_pos = _f32(tgt_pos)
    tgt_mass = _f32(tgt_mass)
    out = np.zeros(len(tgt_pos), np.float64)
    lib().oracle_direct_potential(C.byref(prm), len(src_pos), _p(src_pos), _p(src_mass), len(tgt_pos),
                                  _p(tgt_pos), _p(tgt_mass), nthreads, _p(out))
    return out


def tree_dfs(src_pos, src_mass, params=None):
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    src_pos = _f32(src_pos)
    src_mass = _f32(src_mass)
    n = len(src_pos)
    cap = 24 * n + 16
    kind = np.zeros(cap, np.int32)
    pidx = np.zeros(cap, np.int64)
    mass = np.zeros(cap, np.float32)
    com = np.zeros((cap, 3), np.float32)
    ln = np.zeros(cap, np.float32)
    skip = np.zeros(cap, np.int64)
    depth = np.zeros(cap, np.int32)
    tot = lib().oracle_tree_dfs(C.byref(prm), n, _p(src_pos), _p(src_mass), cap, _p(kind), _p(pidx), _p(mass),
                                _p(com), _p(ln), _p(skip), _p(depth))
    assert tot >= 0
    s = slice(0, tot)
    return dict(kind=kind[s], pidx=pidx[s], mass=mass[s], com=com[s], len=ln[s], skip=skip[s], depth=depth[s])


def core_phase_space(pos, vel, mass, ptype=None, params=None):
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    pos = _f32(pos)
    vel = _f32(vel)
    mass = _f32(mass)
    pt = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.int32)
    cp = np.zeros(3, np.float32)
    cv = np.zeros(3, np.float32)
    sr = C.c_float()
    sv = C.c_float()
    lib().oracle_core_phase_space(C.byref(prm), len(pos), _p(pos), _p(vel), _p(mass), _p(pt), _p(cp), _p(cv),
                                  C.byref(sr), C.byref(sv))
    return cp, cv, sr.value, sv.value


PROFILE_DTYPE = np.dtype([
    ("rmax", np.float32), ("vmax", np.float32), ("r_half", np.float32), ("r_enclose", np.float32),
    ("r200", np.float32), ("m200", np.float32), ("inertia", np.float32, 6), ("inertia_w", np.float32, 6),
    ("has_r200", np.int32), ("reserved", np.int32)])


def profiles(pos, mass, offsets, nbound, mbound, params=None):
    """Subhalo_t::CalculateProfileProperties + CalculateShape of every subhalo's
    bound set, particles given in final order from offsets[s] (first = centre)."""
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    pos, mass = _f32(pos), _f32(mass)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nb = np.ascontiguousarray(nbound, dtype=np.int64)
    mb = _f32(mbound)
    out = np.zeros(len(nb), PROFILE_DTYPE)
    lib().oracle_profiles(C.byref(prm), len(mass), _p(pos), _p(mass), len(nb), _p(off), _p(nb), _p(mb), _p(out))
    return out


def reassign_targets(pos, is_tracer, sub, group_offsets, k=10, rguess=None, params=None):
    """SubhaloSnapshot_t::ReassignParticles' selection: target subhalo per particle, -1 to stay."""
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    pos = _f32(pos)
    tr = np.ascontiguousarray(is_tracer, dtype=np.uint8)
    sb = np.ascontiguousarray(sub, dtype=np.int32)
    go = np.ascontiguousarray(group_offsets, dtype=np.int64)
    if rguess is None:
        rguess = np.float32(prm.softening * 0.1)
    out = np.zeros(len(tr), np.int32)
    lib().oracle_reassign_targets(C.byref(prm), len(tr), _p(pos), _p(tr), _p(sb), len(go) - 1, _p(go), int(k),
                                  float(rguess), _p(out))
    return out


def mass_upscale(mass, ptype, nlast, mlast, mss, nuns, params=None):
    """SetMassUpscaleFactor over particles in E-list order: (sources, factors[6])."""
    prm = params if isinstance(params, Params) else make_params(**(params or {}))
    mass = _f32(mass)
    pt = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.int32)
    out = np.zeros(6, np.float32)
    ns = lib().oracle_mass_upscale(C.byref(prm), len(mass), _p(mass), _p(pt), int(nlast), float(mlast), int(mss),
                                   int(nuns), _p(out))
    return int(ns), out
