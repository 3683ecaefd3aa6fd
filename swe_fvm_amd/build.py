"""Builds libswe_fvm_amd.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libswe_fvm_amd.so")
SOURCES = ["swe_kernels.hip", "swe_fused.hip", "swe_adjoint.hip", "swe_tangent.hip", "swe_api.cpp",
           "swe_mesh.cpp", "swe_tiles.cpp"]
HEADERS = ["swe_device.h", "swe_dual.h", "swe_common.h", "swe_launch.h", "swe_internal.h", "swe_tiles.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every expression is evaluated as written (no FMA contraction), so the
    # device results reproduce a scalar CPU evaluation bit for bit
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False, out=None, defines=()):
    """Compiles every source into LIB (or `out`); `defines` are extra -D
    tuning macros (work items per thread).  Raises CalledProcessError."""
    target = out or LIB
    cmd = [HIPCC] + FLAGS + [f"-D{d}" for d in defines] + ["-o", target + ".tmp"] + \
        [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    os.replace(target + ".tmp", target)
    return target


def build_pybind(verbose=False):
    """Compiles the Python module SWE_FVM (pybind11, host code) next to the
    library it links, with an $ORIGIN rpath.  Needs LIB built first."""
    import sysconfig

    import pybind11
    out = os.path.join(HERE, "SWE_FVM" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-Wall", "-Wextra",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"], os.path.join(CSRC, "swe_pybind.cpp"),
           "-L", HERE, "-lswe_fvm_amd", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(verbose=True, out=a.out, defines=a.defines))
