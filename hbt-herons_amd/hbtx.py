"""ctypes binding of libhbtx (the C-ABI in include/hbtx.h).

The HIP library is the product: this module only marshals numpy arrays across
the C ABI.  It raises if the library is missing; there is no CPU fallback.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libhbtx.so")
HEADER = os.path.join(ROOT, "include", "hbtx.h")

OK = 0
E_INVALID = -1
E_ALLOC = -2
E_HIP = -3
E_CAPACITY = -4
E_NOT_IMPLEMENTED = -5

BATCHED = 0
PLAIN = 1
WARMSTART = 2
SAVE_POTENTIAL = 4
NO_GRAPH = 8
COUNT = 16


class HbtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hbtx error {code}: {msg}")
        self.code = code


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
        ("direct_cutoff", C.c_int64),
        ("reserved", C.c_int32 * 2),
    ]


# hbtx_sub_profile (CalculateProfileProperties + CalculateShape)
PROFILE_DTYPE = np.dtype([
    ("rmax", np.float32), ("vmax", np.float32), ("r_half", np.float32), ("r_enclose", np.float32),
    ("r200", np.float32), ("m200", np.float32), ("inertia", np.float32, 6), ("inertia_w", np.float32, 6),
    ("has_r200", np.int32), ("reserved", np.int32)])

RESULT_DTYPE = np.dtype([
    ("nbound", np.int64), ("mbound", np.float32), ("nbound_type", np.int64, 6),
    ("mbound_type", np.float32, 6), ("com_pos", np.float32, 3), ("com_vel", np.float32, 3),
    ("mostbound_pos", np.float32, 3), ("mostbound_vel", np.float32, 3), ("spec_pot", np.float32),
    ("spec_kin", np.float32), ("spec_am", np.float32, 3), ("tracer_index", np.int64),
    ("mostbound_id", np.int64), ("core_pos", np.float32, 3), ("core_vel", np.float32, 3),
    ("core_sigma_r", np.float32), ("core_sigma_v", np.float32), ("iterations", np.int32),
    ("disrupted", np.int32)], align=True)

STAT_NAMES = ["tree_wave_steps", "tree_lane_visits", "tree_lane_accepts", "direct_lane_pairs", "targets",
              "work_items", "near_interactions", "tree_entries", "tree_entries_first", "unused9", "unused10", "unused11",
              "walk_ms", "rounds", "graph_rounds", "graph_captures"]


def workload_params(**kw):
    """The benchmark workloads' parameters: theta=0.7, eps=1 kpc/h, a=1, H=0,
    non-periodic, no subsampling, reference defaults elsewhere."""
    p = dict(G=43.0071, scale_factor=1.0, hz=0.0, softening=0.001, max_phys_softening=-1.0, theta=0.7,
             bound_mass_precision=0.995, max_sample_size=0, refine_most_bound=1,
             bound_fraction_center_refinement=0.1, min_num_part=20, min_num_tracer=10,
             tracer_mask=(1 << 1) | (1 << 4), donot_subsample_mask=(1 << 5), periodic=0, box_size=100.0,
             thermal=0, dm_only=1, upscale_per_type=1, sub_core_size_factor=0.0, sub_core_size_min=20,
             direct_cutoff=0)
    p.update(kw)
    return p


def make_params(**kw):
    d = workload_params(**kw)
    return Params(**d)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HbtxError(E_INVALID, f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64 = C.c_int64
        L.hbtx_default_params.argtypes = [C.POINTER(Params)]
        L.hbtx_default_params.restype = None
        L.hbtx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        L.hbtx_destroy.argtypes = [P]
        L.hbtx_destroy.restype = None
        L.hbtx_last_error.argtypes = [P]
        L.hbtx_last_error.restype = C.c_char_p
        L.hbtx_device_sync.argtypes = [P]
        L.hbtx_unbind_batch.argtypes = [P, C.POINTER(Params), i64, P, P, P, P, P, P, i64, P, P, P, P, P,
                                        C.c_uint32, P, P, P, P]
        L.hbtx_batch_load.argtypes = [P, i64, P, P, P, P, P, P, i64, P, P, P, P, P]
        L.hbtx_batch_set_proposal.argtypes = [P, P]
        L.hbtx_batch_run.argtypes = [P, C.POINTER(Params), C.c_uint32, C.POINTER(C.c_double)]
        L.hbtx_batch_fetch.argtypes = [P, P, P, P, P]
        L.hbtx_batch_stats.argtypes = [P, P, C.c_int]
        L.hbtx_device_bytes.argtypes = [C.POINTER(C.c_int64)]
        L.hbtx_stat_name.argtypes = [C.c_int]
        L.hbtx_stat_name.restype = C.c_char_p
        L.hbtx_tree_build.argtypes = [P, C.POINTER(Params), i64, P, P, C.POINTER(C.c_void_p), C.POINTER(i64)]
        L.hbtx_tree_potential.argtypes = [P, P, i64, P, P, P]
        L.hbtx_tree_binding_energy.argtypes = [P, P, i64, P, P, P, P, P, P]
        L.hbtx_tree_export.argtypes = [P, P, i64, P, P, P, P, P, P, P, C.POINTER(i64)]
        L.hbtx_tree_free.argtypes = [P]
        L.hbtx_tree_free.restype = None
        L.hbtx_direct_potential.argtypes = [P, C.POINTER(Params), i64, P, P, i64, P, P, P]
        L.hbtx_direct_potential_noalloc.argtypes = [P, C.POINTER(Params), i64, P, P, i64, P, P, P]
        L.hbtx_segmented_sort.argtypes = [P, i64, P, P, i64, P]
        L.hbtx_check_layout.argtypes = [i64, i64, P, C.c_char_p, C.c_int32]
        L.hbtx_batch_profiles.argtypes = [P, C.POINTER(Params), P]
        L.hbtx_profiles.argtypes = [P, C.POINTER(Params), i64, P, P, i64, P, P, P, P]
        L.hbtx_reassign_targets.argtypes = [P, C.POINTER(Params), i64, P, P, P, i64, P, C.c_int32, C.c_float, P,
                                            C.POINTER(C.c_double)]
        L.hbtx_mlp_create.argtypes = [P, C.c_uint64, C.POINTER(C.c_void_p)]
        L.hbtx_mlp_free.argtypes = [P]
        L.hbtx_mlp_free.restype = None
        L.hbtx_mlp_get_weights.argtypes = [P, P, P, P]
        L.hbtx_mlp_set_weights.argtypes = [P, P, P]
        L.hbtx_mlp_gradient.argtypes = [P, P, i64, P, P, P, C.POINTER(C.c_double)]
        L.hbtx_mlp_train.argtypes = [P, P, i64, P, P, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                     C.POINTER(C.c_double)]
        L.hbtx_mlp_predict.argtypes = [P, P, i64, P, P]
        L.hbtx_batch_set_previous.argtypes = [P, P, P, P, P, P, P, P]
        L.hbtx_batch_features.argtypes = [P, C.POINTER(Params), P]
        L.hbtx_batch_propose.argtypes = [P, P, C.c_float, i64, C.POINTER(i64), C.POINTER(C.c_double)]
        L.hbtx_batch_warm_proposal.argtypes = [P, C.POINTER(Params), P, C.c_float, i64, C.POINTER(i64),
                                               C.POINTER(C.c_double)]
        L.hbtx_batch_get_proposal.argtypes = [P, P]
        L.hbtx_get_tuning.argtypes = [P, C.POINTER(i64), C.POINTER(C.c_int32)]
        L.hbtx_set_tuning.argtypes = [P, i64, C.c_int32]
        _lib = L
    return _lib


def device_bytes():
    """Device memory the library's buffers hold now (all contexts of the process)."""
    b = C.c_int64(0)
    rc = lib().hbtx_device_bytes(C.byref(b))
    if rc != 0:
        raise HbtxError(rc, "hbtx_device_bytes failed")
    return b.value


def check_layout(n, offsets):
    """hbtx_check_layout: (status, message) for a batch of n particles in
    len(offsets) - 1 subhaloes; needs no device."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    buf = C.create_string_buffer(512)
    rc = lib().hbtx_check_layout(int(n), len(off) - 1, off.ctypes.data_as(C.c_void_p), buf, 512)
    return rc, buf.value.decode()


def header_symbols():
    """Function names declared in include/hbtx.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(hbtx_[a-z_0-9]+)\s*\(", txt)))


class Context:
    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().hbtx_create(C.byref(self._h), device)
        if rc != OK:
            raise HbtxError(rc, "hbtx_create failed (no HIP device?)")

    def close(self):
        if self._h:
            lib().hbtx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise HbtxError(rc, lib().hbtx_last_error(self._h).decode())

    def sync(self):
        self._check(lib().hbtx_device_sync(self._h))

    # --- batch API -----------------------------------------------------------
    def load(self, pop, mb_pos=None, mb_vel=None):
        ptype = None if pop.ptype is None else np.ascontiguousarray(pop.ptype, dtype=np.int32)
        self._n = pop.n
        self._nsub = pop.nsub
        self._check(lib().hbtx_batch_load(self._h, pop.n, _p(pop.pos), _p(pop.vel), _p(pop.mass), _p(pop.u),
                                          _p(ptype), _p(pop.pid), pop.nsub, _p(pop.offsets), _p(pop.ref_pos),
                                          _p(pop.ref_vel), _p(_f32(mb_pos)), _p(_f32(mb_vel))))

    def set_proposal(self, proposal):
        pr = np.ascontiguousarray(proposal, dtype=np.uint8)
        self._check(lib().hbtx_batch_set_proposal(self._h, _p(pr)))

    def run(self, params=None, flags=BATCHED):
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        ms = C.c_double(0.0)
        self._check(lib().hbtx_batch_run(self._h, C.byref(prm), flags, C.byref(ms)))
        return ms.value

    def fetch(self, want_potential=False):
        order = np.zeros(self._n, np.int64)
        energy = np.zeros(self._n, np.float32)
        pot = np.zeros(self._n, np.float32) if want_potential else None
        res = np.zeros(self._nsub, RESULT_DTYPE)
        self._check(lib().hbtx_batch_fetch(self._h, _p(order), _p(energy), _p(pot), _p(res)))
        return dict(order=order, energy=energy, potential=pot, result=res)

    def stats(self):
        out = np.zeros(16, np.float64)
        self._check(lib().hbtx_batch_stats(self._h, _p(out), 16))
        return {STAT_NAMES[i]: out[i] for i in range(16)}

    def unbind(self, pop, params=None, flags=BATCHED, mb_pos=None, mb_vel=None):
        """One-call Subhalo_t::Unbind over every subhalo of a Population."""
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        n = pop.n
        order = np.zeros(n, np.int64)
        energy = np.zeros(n, np.float32)
        want_pot = bool(flags & SAVE_POTENTIAL)
        pot = np.zeros(n, np.float32) if want_pot else None
        res = np.zeros(pop.nsub, RESULT_DTYPE)
        ptype = None if pop.ptype is None else np.ascontiguousarray(pop.ptype, dtype=np.int32)
        self._n = n
        self._nsub = pop.nsub
        self._check(lib().hbtx_unbind_batch(self._h, C.byref(prm), n, _p(pop.pos), _p(pop.vel), _p(pop.mass),
                                            _p(pop.u), _p(ptype), _p(pop.pid), pop.nsub, _p(pop.offsets),
                                            _p(pop.ref_pos), _p(pop.ref_vel), _p(_f32(mb_pos)), _p(_f32(mb_vel)),
                                            flags, _p(order), _p(energy), _p(pot), _p(res)))
        return dict(order=order, energy=energy, potential=pot, result=res)

    # --- tuning ------------------------------------------------------------------
    def tuning(self):
        dc = C.c_int64(0)
        wc = C.c_int32(0)
        self._check(lib().hbtx_get_tuning(self._h, C.byref(dc), C.byref(wc)))
        return dict(direct_cutoff=dc.value, walk_chunks=wc.value)

    def set_tuning(self, direct_cutoff, walk_chunks):
        self._check(lib().hbtx_set_tuning(self._h, int(direct_cutoff), int(walk_chunks)))

    # --- learned warm start ----------------------------------------------------
    def set_previous(self, prev):
        """Upload the previous snapshot's outputs (a dict from learned.previous_state)."""
        rank = np.ascontiguousarray(prev["rank"], dtype=np.int32)
        bound = np.ascontiguousarray(prev["bound"], dtype=np.uint8)
        energy = _f32(prev["energy"])
        cp, cv = _f32(prev["com_pos"]), _f32(prev["com_vel"])
        mb, rs = _f32(prev["mbound"]), _f32(prev["rscale"])
        self._check(lib().hbtx_batch_set_previous(self._h, _p(rank), _p(bound), _p(energy), _p(cp), _p(cv),
                                                  _p(mb), _p(rs)))

    def features(self, params=None, fetch=True):
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        out = np.zeros((self._n, MLP_NIN), np.float32) if fetch else None
        self._check(lib().hbtx_batch_features(self._h, C.byref(prm), _p(out)))
        return out

    def propose(self, mlp, threshold=0.5, count=True, min_particles=0):
        """Proposal = P(bound) > threshold from the batch features (none for
        subhaloes of fewer than min_particles particles); returns (number
        proposed or None, inference milliseconds)."""
        npro = C.c_int64(0)
        ms = C.c_double(0.0)
        self._check(lib().hbtx_batch_propose(self._h, mlp._h, threshold, int(min_particles),
                                             C.byref(npro) if count else None, C.byref(ms)))
        return (npro.value if count else None), ms.value

    def warm_proposal(self, mlp, params=None, threshold=0.5, min_particles=0, count=True):
        """Features + inference + proposal in one call, evaluated only on the
        particles of subhaloes of at least min_particles particles (the proposal
        of features() followed by propose()); returns (number proposed or None,
        milliseconds of the feature and inference kernels)."""
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        npro = C.c_int64(0)
        ms = C.c_double(0.0)
        self._check(lib().hbtx_batch_warm_proposal(self._h, C.byref(prm), mlp._h, threshold, int(min_particles),
                                                   C.byref(npro) if count else None, C.byref(ms)))
        return (npro.value if count else None), ms.value

    def get_proposal(self):
        out = np.zeros(self._n, np.uint8)
        self._check(lib().hbtx_batch_get_proposal(self._h, _p(out)))
        return out

    # --- GravityTree_t API -----------------------------------------------------
    def tree(self, pos, mass, params=None):
        return Tree(self, pos, mass, params)

    def direct_potential(self, src_pos, src_mass, tgt_pos, tgt_mass=None, params=None):
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        src_pos = _f32(src_pos)
        src_mass = _f32(src_mass)
        tgt_pos = _f32(tgt_pos)
        tgt_mass = _f32(tgt_mass)
        out = np.zeros(len(tgt_pos), np.float64)
        self._check(lib().hbtx_direct_potential(self._h, C.byref(prm), len(src_pos), _p(src_pos), _p(src_mass),
                                                len(tgt_pos), _p(tgt_pos), _p(tgt_mass), _p(out)))
        return out

    def profiles(self, params=None):
        """Subhalo_t::CalculateProfileProperties + CalculateShape of the last run's
        subhaloes (PROFILE_DTYPE per subhalo)."""
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        out = np.zeros(self._nsub, PROFILE_DTYPE)
        self._check(lib().hbtx_batch_profiles(self._h, C.byref(prm), _p(out)))
        return out

    def profiles_of(self, pos, mass, offsets, nbound, mbound, params=None):
        """Profiles and shapes of caller-given bound sets (final order, first = centre)."""
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        pos, mass = _f32(pos), _f32(mass)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        nb = np.ascontiguousarray(nbound, dtype=np.int64)
        mb = _f32(mbound)
        out = np.zeros(len(off) - 1, PROFILE_DTYPE)
        self._check(lib().hbtx_profiles(self._h, C.byref(prm), len(mass), _p(pos), _p(mass), len(off) - 1, _p(off),
                                        _p(nb), _p(mb), _p(out)))
        return out

    def reassign_targets(self, pos, is_tracer, sub, group_offsets, k=10, rguess=None, params=None):
        """SubhaloSnapshot_t::ReassignParticles' choice per particle: the subhalo to
        move it to, or -1.  Particles are grouped by host halo (group_offsets).
        Returns (targets int32[n], device milliseconds)."""
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        pos = _f32(pos)
        tr = np.ascontiguousarray(is_tracer, dtype=np.uint8)
        sb = np.ascontiguousarray(sub, dtype=np.int32)
        go = np.ascontiguousarray(group_offsets, dtype=np.int64)
        out = np.zeros(len(tr), np.int32)
        ms = C.c_double(0.0)
        self._check(lib().hbtx_reassign_targets(self._h, C.byref(prm), len(tr), _p(pos), _p(tr), _p(sb), len(go) - 1,
                                                _p(go), int(k), float(rguess or 0.0), _p(out), C.byref(ms)))
        return out, ms.value

    def segmented_sort(self, keys, vals, offsets):
        k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
        v = np.ascontiguousarray(vals, dtype=np.int32).copy()
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._check(lib().hbtx_segmented_sort(self._h, len(k), _p(k), _p(v), len(off) - 1, _p(off)))
        return k, v


MLP_NIN = 8
MLP_NPAR = 4801
MLP_LAYOUT = [("W1", (64, 8)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,))]


def mlp_unpack(flat):
    """Split a flat parameter vector into the named arrays of the 8-64-64-1 network."""
    out, k = {}, 0
    for name, shape in MLP_LAYOUT:
        size = int(np.prod(shape))
        out[name] = np.asarray(flat[k:k + size]).reshape(shape)
        k += size
    return out


class MLP:
    """The warm-start classifier: FP32 master weights, BF16 working copy on the device."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx._check(lib().hbtx_mlp_create(ctx._h, seed, C.byref(self._h)))

    def weights(self):
        master = np.zeros(MLP_NPAR, np.float32)
        bf16 = np.zeros(MLP_NPAR, np.uint16)
        self.ctx._check(lib().hbtx_mlp_get_weights(self.ctx._h, self._h, _p(master), _p(bf16)))
        return master, bf16

    def set_weights(self, master):
        w = np.ascontiguousarray(master, dtype=np.float32)
        assert w.shape == (MLP_NPAR,)
        self.ctx._check(lib().hbtx_mlp_set_weights(self.ctx._h, self._h, _p(w)))

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        g = np.zeros(MLP_NPAR, np.float32)
        loss = C.c_double(0.0)
        self.ctx._check(lib().hbtx_mlp_gradient(self.ctx._h, self._h, len(X), _p(X), _p(y), _p(g), C.byref(loss)))
        return g, loss.value

    def train(self, X, y, epochs=3, batch=4096, lr=2e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        loss = C.c_double(0.0)
        self.ctx._check(lib().hbtx_mlp_train(self.ctx._h, self._h, len(X), _p(X), _p(y), epochs, batch, lr, seed,
                                             C.byref(loss)))
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), np.float32)
        self.ctx._check(lib().hbtx_mlp_predict(self.ctx._h, self._h, len(X), _p(X), _p(out)))
        return out

    def close(self):
        if self._h:
            lib().hbtx_mlp_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tree:
    """GravityTree_t counterpart: Build on construction, EvaluatePotential / BindingEnergy."""

    def __init__(self, ctx, pos, mass, params=None):
        self.ctx = ctx
        prm = params if isinstance(params, Params) else make_params(**(params or {}))
        self._pos = _f32(pos)
        self._mass = _f32(mass)
        self._h = C.c_void_p()
        nc = C.c_int64(0)
        ctx._check(lib().hbtx_tree_build(ctx._h, C.byref(prm), len(self._pos), _p(self._pos), _p(self._mass),
                                         C.byref(self._h), C.byref(nc)))
        self.ncells = nc.value

    def potential(self, tgt_pos, tgt_mass=None):
        tgt_pos = _f32(tgt_pos)
        tgt_mass = _f32(tgt_mass)
        out = np.zeros(len(tgt_pos), np.float64)
        self.ctx._check(lib().hbtx_tree_potential(self.ctx._h, self._h, len(tgt_pos), _p(tgt_pos), _p(tgt_mass),
                                                  _p(out)))
        return out

    def binding_energy(self, tgt_pos, tgt_vel, ref_pos, ref_vel, tgt_mass=None):
        tgt_pos = _f32(tgt_pos)
        tgt_vel = _f32(tgt_vel)
        rp = _f32(ref_pos)
        rv = _f32(ref_vel)
        tgt_mass = _f32(tgt_mass)
        out = np.zeros(len(tgt_pos), np.float64)
        self.ctx._check(lib().hbtx_tree_binding_energy(self.ctx._h, self._h, len(tgt_pos), _p(tgt_pos), _p(tgt_vel),
                                                       _p(rp), _p(rv), _p(tgt_mass), _p(out)))
        return out

    def export(self):
        cap = len(self._pos) + self.ncells
        kind = np.zeros(cap, np.int32)
        pidx = np.zeros(cap, np.int64)
        mass = np.zeros(cap, np.float32)
        com = np.zeros((cap, 3), np.float32)
        ln = np.zeros(cap, np.float32)
        skip = np.zeros(cap, np.int64)
        depth = np.zeros(cap, np.int32)
        ne = C.c_int64(0)
        self.ctx._check(lib().hbtx_tree_export(self.ctx._h, self._h, cap, _p(kind), _p(pidx), _p(mass), _p(com),
                                               _p(ln), _p(skip), _p(depth), C.byref(ne)))
        s = slice(0, ne.value)
        return dict(kind=kind[s], pidx=pidx[s], mass=mass[s], com=com[s], len=ln[s], skip=skip[s], depth=depth[s])

    def close(self):
        if self._h:
            lib().hbtx_tree_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
