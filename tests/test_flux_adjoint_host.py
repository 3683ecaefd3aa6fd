"""Reverse mode of the Riemann fluxes (swe_fvm_amd/csrc/swe_flux_adj.h, used
by the adjoint's edge kernel) against the dual-number derivative of the
forward flux templates (swe_device.h), on the host: HLL and HLLC with the
Rusanov, Davis and Einfeldt wave speeds (include/Fluxes.h:14-111,
src/Fluxes.cpp:5-26), 100,000 random edge states each, dry sides, films,
supersonic flows and tied wave speeds included.  tests/cpp/flux_vjp_check.cpp
holds the cases and the tolerances."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "flux_vjp_check.cpp")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_flux_reverse_mode_is_the_transposed_dual_derivative(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "flux_vjp_check")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "swe_fvm_amd", "csrc"), SRC, "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    res = subprocess.run([exe, "100000"], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    lines = [ln for ln in res.stdout.splitlines() if "cases=" in ln]
    assert len(lines) == 6 and all("bad=0" in ln for ln in lines)
