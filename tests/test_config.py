"""The .ini Parser and DimensionManager (include/swe_config.hpp and
swe_fvm_amd/config.py): parsing rules, scale identities, the Classic Thacker
configuration's closed-form constants, and the C++ and Python twins agree."""
import math
import os
import subprocess

import pytest

from swe_fvm_amd import config as CF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TEXT = """
; the classic Thacker bowl in SI units: h0 = 0.1 m deep, a = 1 m wide
[Dimensions]
h0 = 0.1
l0 = 1.0
g  = 9.81

[Common]
  cor = 0.0   
tau=0.0
delta = 1.0
# Thacker: H0 for r0 = 0.8 (A = 0.2195)
[Thacker]
H0 = 0.125
p0 = 0
q0 = 0
[Ball]
x0 = 0.3
y0 = -0.1
u0 = 0.05
v0 = 0
label = left
"""


def test_parser_rules():
    p = CF.Parser.from_string(TEXT)
    assert p.Get("Common", "cor") == 0.0 and p.Get("Common", "tau") == 0.0
    assert p.Get("Dimensions", "g") == 9.81
    assert p.Get("Ball", "y0") == -0.1
    with pytest.raises(CF.ParserError):
        p.Get("Common", "missing")
    with pytest.raises(CF.ParserError):
        p.Get("Nowhere", "cor")
    with pytest.raises(CF.ParserError):
        p.Get("Ball", "label")
    with pytest.raises(CF.ParserError):
        CF.Parser.from_string("[A]\nno equals sign\n")


def test_scales_invert_and_match_the_classic_thacker_constants():
    p = CF.Parser.from_string(TEXT)
    d = CF.DimensionManager(p)
    for k in CF.Scales:
        for x in (0.3, 2.5):
            assert math.isclose(d.Unscale(k, d.Scale(k, x)), x, rel_tol=1e-15)
    # U = sqrt(g h0), T = l0 / U: one period 2 pi / sqrt(8) in solver units is
    # 2.2429 s at g = 9.81, h0 = 0.1, a = 1
    th = CF.make_test("ClassicThacker", p, d)
    assert math.isclose(th.period, 2 * math.pi / math.sqrt(8), rel_tol=1e-15)
    assert abs(d.Scale(CF.Scales.time, th.period) - 2.2429) < 1e-4
    # H0 = 0.125 m -> 1.25 in units of h0; its amplitude parameter is
    # A = (a^2 - r0^2) / (a^2 + r0^2) for r0 = 0.8
    assert math.isclose(th.H0, 1.25, rel_tol=1e-15)
    assert math.isclose(th.a, (1 - 0.64) / (1 + 0.64), rel_tol=1e-12)
    ball = CF.make_test("Ball", p, d, mid=(0.1, 0.0))
    assert math.isclose(ball.x0, 0.3 - 0.1) and math.isclose(ball.u0, 0.05 / math.sqrt(0.981))
    with pytest.raises(CF.DimensionError):
        CF.DimensionManager(CF.Parser.from_string("[Dimensions]\nh0 = 0\nl0 = 1\n"))


def test_cpp_header_agrees(tmp_path):
    exe = str(tmp_path / "config_check")
    res = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "tests", "cpp", "config_check.cpp"), "-o", exe],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    cfg = tmp_path / "config.ini"
    cfg.write_text(TEXT)
    res = subprocess.run([exe, str(cfg)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, (res.returncode, res.stderr)
    got = [float(x) for x in res.stdout.split()]
    p = CF.Parser.from_string(TEXT)
    d = CF.DimensionManager(p)
    want = [d.Unscale(CF.Scales.source, p.Get("Common", "cor")), d.Unscale(CF.Scales.height, p.Get("Thacker", "H0")),
            d.Scale(CF.Scales.time, 2.0), d.Scale(CF.Scales.velocity, 1.5), d.Unscale(CF.Scales.length, p.Get("Ball", "x0"))]
    assert got == want
