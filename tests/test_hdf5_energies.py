"""The BindingEnergies / PotentialEnergies datasets of a subhalo snapshot file
(include/hbtx_hdf5.hpp; src/io/subhalo_io.cpp:425-451), written and read back
through the HDF5 C library by tests/cpp/energies_hdf5_test.cpp:

* host side: one dataset over the subhaloes (empty ones included) of
  variable-length native floats, every value bit for bit, the vectors cleared as
  the reference clears them, a group holding only the binding energies, and the
  error for a vector shorter than Nbound;
* the header compiles as C++11 with -pedantic -Werror;
* GPU: hbtx::RefineParticles fills both vectors and the file holds, per subhalo,
  exactly the C ABI's energies and potentials of its bound particles in final
  order.

Skipped where the HDF5 headers and library are absent (the driver is then not
built)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import hbtx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(os.path.dirname(hbtx.LIB_PATH), "energies_hdf5_test")
HDF5_INCLUDE = "/opt/conda/include"
CXX = shutil.which("g++")

needs_driver = pytest.mark.skipif(not os.path.exists(BIN), reason="HDF5 driver not built (no HDF5 library)")


@needs_driver
def test_datasets_round_trip(tmp_path):
    r = subprocess.run([BIN, "selftest", str(tmp_path / "energies.h5")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


@pytest.mark.skipif(CXX is None or not os.path.exists(os.path.join(HDF5_INCLUDE, "hdf5.h")),
                    reason="no g++ or no HDF5 headers")
def test_header_compiles_as_cpp11(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "hbtx_hdf5.hpp"\n#include <vector>\n'
                   'struct S { long long Nbound; std::vector<float> ParticleBindingEnergies, '
                   'ParticlePotentialEnergies; };\n'
                   'void f(hid_t h, std::vector<S> &v) { hbtx::SaveParticleEnergies(h, v, true, true); }\n')
    cmd = [CXX, "-std=c++11", "-fsyntax-only", "-Wall", "-Wextra", "-pedantic", "-Werror",
           "-I", os.path.join(ROOT, "include"), "-isystem", HDF5_INCLUDE, str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.gpu
@needs_driver
def test_gpu_energies_in_the_file(ctx, tmp_path):
    import synth
    pop = synth.population(seed=31, nsub=40, nmax=5000)
    inp, h5, outp = tmp_path / "in.bin", tmp_path / "e.h5", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([pop.n, pop.nsub], np.int64).tofile(f)
        pop.offsets.astype(np.int64).tofile(f)
        for a in (pop.pos, pop.vel, pop.mass, pop.ref_pos, pop.ref_vel):
            np.ascontiguousarray(a, np.float32).tofile(f)
    r = subprocess.run([BIN, "gpu", str(inp), str(h5), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    raw = open(outp, "rb").read()
    # the driver's parameters: theta 0.7, no subsampling, trees for every subhalo
    ref = ctx.unbind(pop, hbtx.workload_params(direct_cutoff=0), hbtx.BATCHED | hbtx.SAVE_POTENTIAL)
    nb = ref["result"]["nbound"]
    k = 0
    for key in ("energy", "potential"):
        for s in range(pop.nsub):
            ln = int(np.frombuffer(raw, np.int64, count=1, offset=k)[0])
            k += 8
            got = np.frombuffer(raw, np.float32, count=ln, offset=k)
            k += 4 * ln
            a = pop.offsets[s]
            assert ln == nb[s], (key, s)
            assert np.array_equal(got, ref[key][a:a + ln].astype(np.float32)), (key, s)
    assert k == len(raw)
    assert nb.sum() > 0
