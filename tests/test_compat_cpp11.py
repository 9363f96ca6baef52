"""The C++ host layer (include/hbtx_compat.hpp) compiles as C++11, the standard
the reference builds with (CMakeLists.txt:38-39: CMAKE_CXX_STANDARD 11).

* the three C++ drivers under tests/cpp (stand-in subhalo types) pass
  `g++ -std=c++11 -pedantic -Werror`;
* tests/cpp/compat_reference_types.cpp instantiates every template with the
  reference's own Subhalo_t / Snapshot_t / MemberShipTable_t types, as a
  maintainer would call them from SubhaloSnapshot_t::RefineParticles.  It reads
  the reference's headers with `-fsyntax-only` (nothing is built or run), under
  the macro sets of the default, DM-only and hydro builds; it is skipped where
  the reference tree or its MPI/HDF5 headers are absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
MPI_HDF5_INCLUDE = "/opt/conda/include"
CXX = shutil.which("g++")


def _compile(src, extra):
    cmd = [CXX, "-std=c++11", "-fsyntax-only", "-Wall", "-Wextra", "-pedantic", "-Werror",
           "-I", os.path.join(ROOT, "include")] + extra + [src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stderr[-4000:]}"


@pytest.mark.skipif(CXX is None, reason="no g++")
@pytest.mark.parametrize("name", ["compat_test.cpp", "hierarchy_test.cpp", "reassign_test.cpp"])
def test_drivers_compile_as_cpp11(name):
    _compile(os.path.join(ROOT, "tests", "cpp", name), [])


@pytest.mark.skipif(CXX is None or not os.path.isdir(REF_SRC)
                    or not os.path.exists(os.path.join(MPI_HDF5_INCLUDE, "hdf5.h")),
                    reason="reference headers (and their MPI/HDF5 headers) not present")
@pytest.mark.parametrize("macros", [
    ["-DHBT_INT8", "-DINPUT_INT8"],
    ["-DHBT_INT8", "-DINPUT_INT8", "-DDM_ONLY"],
    ["-DHBT_INT8", "-DINPUT_INT8", "-DUNBIND_WITH_THERMAL_ENERGY", "-DMEASURE_UNBINDING_TIME"],
    [],
], ids=["default", "dm_only", "hydro_thermal", "int4"])
def test_host_layer_instantiates_with_reference_types(macros):
    _compile(os.path.join(ROOT, "tests", "cpp", "compat_reference_types.cpp"),
             macros + ["-isystem", REF_SRC, "-isystem", MPI_HDF5_INCLUDE])
