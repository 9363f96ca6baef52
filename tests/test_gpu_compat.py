"""The C++ host layer (include/hbtx_compat.hpp: GravityTree, Unbind,
RefineParticles over types with the reference's member names) against the C ABI
called directly: same bound sets, permutations, frames and potentials."""
import os
import subprocess

import numpy as np
import pytest

import hbtx
import synth

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(hbtx.LIB_PATH), "compat_test")


def test_cpp_host_layer_matches_c_abi(ctx, tmp_path):
    assert os.path.exists(BIN), "build the C++ test driver first (make -C hbt-herons_amd)"
    pop = synth.population(seed=31, nsub=40, nmax=5000)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([pop.n, pop.nsub], np.int64).tofile(f)
        pop.offsets.astype(np.int64).tofile(f)
        for a in (pop.pos, pop.vel, pop.mass, pop.ref_pos, pop.ref_vel):
            np.ascontiguousarray(a, np.float32).tofile(f)
    r = subprocess.run([BIN, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(outp, "rb").read()
    rec = np.dtype([("nbound", np.int64), ("mbid", np.int64), ("mbound", np.float32), ("com", np.float32, 6),
                    ("death", np.int32)])
    subs = np.frombuffer(raw, rec, count=pop.nsub)
    k = rec.itemsize * pop.nsub
    ids = np.frombuffer(raw, np.int64, count=pop.n, offset=k)
    k += 8 * pop.n
    prm = hbtx.workload_params()
    ref = ctx.unbind(pop, prm, hbtx.BATCHED)
    res = ref["result"]
    assert np.array_equal(subs["nbound"], res["nbound"])
    assert np.array_equal(ids, ref["order"])
    np.testing.assert_array_equal(subs["mbound"], res["mbound"])
    np.testing.assert_array_equal(subs["com"][:, :3], res["com_pos"])
    np.testing.assert_array_equal(subs["com"][:, 3:], res["com_vel"])
    alive = res["disrupted"] == 0
    # Particle ids in the driver are local input indices
    expect = np.array([ref["order"][pop.offsets[s] + res["tracer_index"][s]] for s in range(pop.nsub)])
    assert np.array_equal(subs["mbid"][alive], expect[alive])
    assert np.all(subs["mbid"][~alive] == -1)  # kept its previous value
    assert np.all(subs["death"][~alive] == 7) and np.all(subs["death"][alive] == -1)
    # Unbind() of subhalo 0 (plain path)
    n0 = int(pop.offsets[1])
    nb0 = int(np.frombuffer(raw, np.int64, count=1, offset=k)[0])
    k += 8
    ids0 = np.frombuffer(raw, np.int64, count=n0, offset=k)
    k += 8 * n0
    assert nb0 == res["nbound"][0]
    assert np.array_equal(ids0, ref["order"][:n0])
    # GravityTree over subhalo 0
    m = min(16, n0)
    pot = np.frombuffer(raw, np.float64, count=m, offset=k)
    be = np.frombuffer(raw, np.float64, count=m, offset=k + 8 * m)
    t = ctx.tree(pop.pos[:n0], pop.mass[:n0], prm)
    # one target per call vs 64 per wave: the compensated sums group differently (FP32 rounding)
    np.testing.assert_allclose(pot, t.potential(pop.pos[:m], pop.mass[:m]), rtol=1e-6)
    np.testing.assert_allclose(be, t.binding_energy(pop.pos[:m], pop.vel[:m], pop.ref_pos[0], pop.ref_vel[0],
                                                    pop.mass[:m]), rtol=1e-6)
    # CalculateProfileProperties / CalculateShape written back by the host layer
    k += 16 * m
    prec = np.dtype([("f", np.float32, 6), ("snap", np.int32)])
    pr = np.frombuffer(raw, prec, count=pop.nsub, offset=k)
    ctx.unbind(pop, prm, hbtx.BATCHED)
    g = ctx.profiles(prm)
    np.testing.assert_array_equal(pr["f"][:, 0], g["rmax"])
    np.testing.assert_array_equal(pr["f"][:, 1], g["vmax"])
    np.testing.assert_array_equal(pr["f"][:, 2], g["r_half"])
    np.testing.assert_array_equal(pr["f"][:, 3], g["r_enclose"])
    np.testing.assert_array_equal(pr["f"][:, 4], g["inertia"][:, 0])
    np.testing.assert_array_equal(pr["f"][:, 5], g["inertia_w"][:, 5])
    live = res["nbound"] > 1
    assert np.all(pr["snap"][live] == 7) and np.all(pr["snap"][~live] == -1)


def test_cpp_host_layer_hydro_types_and_thermal(ctx, tmp_path):
    """The host layer with particle types and internal energy (the hydro build:
    Type, InternalEnergy; NboundType / MboundType written back): the same bound
    sets, permutations, per-type counts and masses and tracer index as the C ABI
    with the same types and u."""
    assert os.path.exists(BIN), "build the C++ test driver first (make -C hbt-herons_amd)"
    pop = synth.population(seed=32, nsub=40, nmax=5000, hydro=True)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([pop.n, pop.nsub], np.int64).tofile(f)
        pop.offsets.astype(np.int64).tofile(f)
        for a in (pop.pos, pop.vel, pop.mass, pop.ref_pos, pop.ref_vel):
            np.ascontiguousarray(a, np.float32).tofile(f)
        np.ascontiguousarray(pop.ptype, np.int32).tofile(f)
        np.ascontiguousarray(pop.u, np.float32).tofile(f)
    r = subprocess.run([BIN, str(inp), str(outp), "hydro"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = np.dtype([("nbound", np.int64), ("nbt", np.int64, 6), ("mbt", np.float32, 6), ("mbound", np.float32),
                    ("tracer", np.int64)])
    raw = open(outp, "rb").read()
    subs = np.frombuffer(raw, rec, count=pop.nsub)
    ids = np.frombuffer(raw, np.int64, count=pop.n, offset=rec.itemsize * pop.nsub)
    ref = ctx.unbind(pop, hbtx.workload_params(dm_only=0, thermal=1), hbtx.BATCHED)
    res = ref["result"]
    assert np.array_equal(subs["nbound"], res["nbound"])
    assert np.array_equal(ids, ref["order"])
    assert np.array_equal(subs["nbt"], res["nbound_type"])
    assert np.array_equal(subs["mbt"], res["mbound_type"])
    assert np.array_equal(subs["mbound"], res["mbound"])
    assert np.array_equal(subs["tracer"], res["tracer_index"])
    assert res["nbound_type"][:, 0].sum() > 0  # gas is bound somewhere
