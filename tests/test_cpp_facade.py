"""The C++ drop-in facade (include/swe_fvm.hpp): compiles with g++ against
the C ABI; on the GPU, a testGaussWave driver written against it matches
the oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from tests import _common as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "gauss_wave.cpp")
LIBDIR = os.path.join(ROOT, "swe_fvm_amd")


def compile_driver(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), SRC,
           "-o", out, "-L", LIBDIR, "-lswe_fvm_amd", f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_facade_compiles(tmp_path):
    compile_driver(str(tmp_path / "gauss_wave"))


@pytest.mark.gpu
def test_facade_gauss_wave_matches_oracle(tmp_path, golden):
    from oracle import oracle as O
    exe = compile_driver(str(tmp_path / "gauss_wave"))
    res = subprocess.run([exe, golden.bowl_path(), str(tmp_path) + "/"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout)
    out0 = np.loadtxt(tmp_path / "out0.dat")
    out1 = np.loadtxt(tmp_path / "out1.dat")
    xyz, tri, bnd = golden.bowl()
    topo = O.build_topology(xyz.shape[0], tri, bnd)
    xyz = xyz.copy()
    xyz[:, 2] = 0.0
    # the driver evaluates exp with the C library; numpy's exp may differ in
    # the last bit, so the oracle starts from the driver's own initial state
    prim = C.gauss_wave_state(xyz, topo["tp"])
    assert np.max(np.abs(out0[:, 0] - prim[:, 0])) <= 4e-16
    assert np.all(out0[:, 1:] == 0.0)
    prim = np.stack([out0[:, 0], np.zeros(len(out0)), np.zeros(len(out0))], axis=1)
    s = O.OracleSolver(xyz, topo, "HLL", "Einfeldt")
    s.set_state(prim)
    s.run("Euler", 1, dt=1e-3)
    st = s.state()
    ref = np.stack([st[:, 0], st[:, 0] * st[:, 1], st[:, 0] * st[:, 2]], axis=1)
    assert np.array_equal(out1, ref)
