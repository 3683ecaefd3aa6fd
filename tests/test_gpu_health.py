"""Failure detection (swe_state_stats): exact statistics of a normal state,
and detection of host-injected non-finite values and negative depths; the
step runs through them without faulting (every loop bound on the path comes
from the mesh or the bathymetry, not from the state)."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from tests import _common as C

pytestmark = pytest.mark.gpu

FL = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]


def _numpy_stats(xyz, tp, prim):
    h = prim[:, 0] - C.cell_bath(xyz, tp)
    wet = h > 1e-12
    sp = np.sqrt(prim[:, 1] ** 2 + prim[:, 2] ** 2)
    return h.min(), h.max(), sp[wet].max(), (sp + np.sqrt(np.where(wet, h, 0.0)))[wet].max()


def test_stats_exact_on_a_normal_state():
    xyz, topo, prim = C.wet_dry_mesh(40, seed=2)
    sd = swe.SpaceDisc(FL, swe.Domain(xyz, topo), prim)
    st = sd.Stats()
    hmin, hmax, umax, cmax = _numpy_stats(xyz, topo.tables()["tp"], prim)
    print("stats", st, flush=True)
    assert st["nonfinite"] == 0
    assert (st["hmin"], st["hmax"], st["umax"], st["cmax"]) == (hmin, hmax, umax, cmax)
    sd.run("SSPRK2", 10, dt=0.0, cfl=0.15)
    a, b = sd.Stats(), sd.Stats()
    assert a == b and a["nonfinite"] == 0 and a["hmin"] >= 0.0


def test_injected_nan_and_negative_depth_are_detected():
    xyz, topo, prim = C.wet_dry_mesh(30, seed=4)
    tp = topo.tables()["tp"]
    bad = prim.copy()
    wet = np.nonzero(prim[:, 0] - C.cell_bath(xyz, tp) > 0.05)[0]
    bad[wet[10], 1] = np.nan
    bad[wet[20], 0] = C.cell_bath(xyz, tp)[wet[20]] - 0.01  # h = -0.01
    sd = swe.SpaceDisc(FL, swe.Domain(xyz, topo), bad)
    st = sd.Stats()
    print("injected", st, flush=True)
    assert st["nonfinite"] == 1 and st["hmin"] == pytest.approx(-0.01, abs=1e-15)
    sd.run("SSPRK2", 3, dt=1e-3)
    st2 = sd.Stats()
    print("after 3 steps", st2, flush=True)
    assert st2["nonfinite"] >= 1  # the NaN spreads; nothing faulted
