"""The allocation-free integrals (swe_integrals_noalloc, static per-GPU
scratch) against swe_integrals, through the C ABI: bitwise equal totals and
per-cell terms, on two calls in a row, the second on a smaller mesh (the
scratch keeps the first call's partials and must not leak into it)."""
import ctypes

import numpy as np
import pytest

import swe_fvm_amd as swe
from swe_fvm_amd import _lib
from tests import _common as C

pytestmark = pytest.mark.gpu


def _solver(n, seed):
    xyz, topo, prim = C.wet_dry_mesh(n, seed=seed)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, swe.Domain(xyz, topo), prim, device=0)
    sd.run("SSPRK2", 3, dt=0.0, cfl=0.15)
    return sd


def _noalloc(sd, torch):
    L = _lib.load()
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda:0")
    pc = torch.full((sd.T, 3), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = L.swe_integrals_noalloc(sd._h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(pc.data_ptr()))
    assert rc == 0, L.swe_last_error()
    assert L.swe_synchronize(sd._h) == 0
    totals = out.cpu().numpy()
    rc = L.swe_integrals_noalloc(sd._h, ctypes.c_void_p(out.data_ptr()), None)  # per-cell terms optional
    assert rc == 0, L.swe_last_error()
    assert L.swe_synchronize(sd._h) == 0
    assert np.array_equal(out.cpu().numpy(), totals)
    return totals, pc.cpu().numpy()


def test_noalloc_integrals_equal_the_allocating_entry_point():
    torch = pytest.importorskip("torch")
    for n, seed in ((160, 3), (40, 5)):  # 51,200 then 3,200 cells
        sd = _solver(n, seed)
        ref, ref_pc = sd.Integrals(per_cell=True)
        got, got_pc = _noalloc(sd, torch)
        print(f"{sd.T} cells: totals {got}, bitwise={np.array_equal(got, ref)}", flush=True)
        assert np.array_equal(got, ref)
        assert np.array_equal(got_pc, ref_pc)
        del sd


def test_noalloc_integrals_refuse_null_output():
    sd = _solver(8, 1)
    L = _lib.load()
    assert L.swe_integrals_noalloc(sd._h, None, None) != 0
