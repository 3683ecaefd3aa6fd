"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a scalar C++ restatement of SWE_FVM's explicit step
(``oracle/swe_oracle.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this module; the product
library (``swe_fvm_amd``) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_swe.so")

FLUX = {"HLL": 0, "HLLC": 1}
WAVE = {"Rusanov": 0, "Davis": 1, "Einfeldt": 2}
SCHEME = {"Euler": 0, "SSPRK2": 1, "SSPRK3": 2}

OPT_HEAD_GRADIENT = 1
OPT_HEAD_PARTWET2 = 2
OPT_SEQ_DRAINING = 4
OPT_SEQ_NODEMAX = 8
OPT_FRICTION = 16
OPT_DRAIN_LENGTH = 32
OPT_EXACT_BALANCE = 64

_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_int = ctypes.c_int
_p = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    L.orc_topology_build.restype = _p
    L.orc_topology_build.argtypes = [_i64, _i64, _p, _i64, _p]
    L.orc_topology_num_edges.restype = _i64
    L.orc_topology_num_edges.argtypes = [_p]
    L.orc_topology_copy.argtypes = [_p, _p, _p, _p, _p, _p]
    L.orc_topology_free.argtypes = [_p]
    L.orc_create.restype = _p
    L.orc_create.argtypes = [_i64, _p, _i64, _p, _p, _i64, _p, _p, _p, _int, _int, _d, _d, _int, _p]
    L.orc_destroy.argtypes = [_p]
    for n in ("orc_set_state", "orc_get_state", "orc_get_edge_field", "orc_get_src_field",
              "orc_get_fluxes", "orc_get_max_wp", "orc_classify"):
        getattr(L, n).argtypes = [_p, _p]
    L.orc_compute_interface_values.argtypes = [_p]
    L.orc_compute_fluxes.restype = _d
    L.orc_compute_fluxes.argtypes = [_p]
    L.orc_min_len_to_wavespeed.restype = _d
    L.orc_min_len_to_wavespeed.argtypes = [_p]
    L.orc_reconstruct.argtypes = [_p, _i64, _int, _p]
    L.orc_rhs.argtypes = [_p, _i64, _d, _p]
    L.orc_integrals.argtypes = [_p, _p, _p]
    L.orc_run.argtypes = [_p, _int, _d, _d, _i64, _p]
    L.orc_time_run.restype = _d
    L.orc_time_run.argtypes = [_p, _int, _d, _d, _i64]
    L.orc_cbrt.restype = _d
    L.orc_cbrt.argtypes = [_d]
    L.orc_ilog2_trunc.restype = _int
    L.orc_ilog2_trunc.argtypes = [_d]
    L.orc_bisection_cubic.restype = _d
    L.orc_bisection_cubic.argtypes = [_d, _d, _d, _d, _d]
    L.orc_gradient.argtypes = [_p, _p]
    L.orc_flux_edge.restype = _int
    L.orc_flux_edge.argtypes = [_int, _int, _d, _p, _p, _p, _int, _p, _p, _p]
    L.orc_wavespeed.argtypes = [_int, _d, _d, _d, _d, _p]
    L.orc_edge_flux.argtypes = [_p, _int, _int, _i64, _i64, _i64, _p, _p]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def build_topology(num_nodes, tri, bnd):
    """Reference-convention tables from triangles (T x 3) and boundary lines (B x 2)."""
    L = lib()
    tri = np.ascontiguousarray(tri, dtype=np.int64)
    bnd = np.ascontiguousarray(bnd, dtype=np.int64).reshape(-1, 2)
    h = L.orc_topology_build(num_nodes, tri.shape[0], _ptr(tri), bnd.shape[0], _ptr(bnd))
    E = L.orc_topology_num_edges(h)
    T = tri.shape[0]
    ep = np.empty((E, 2), np.int64)
    et = np.empty((E, 2), np.int64)
    tp = np.empty((T, 3), np.int64)
    te = np.empty((T, 3), np.int64)
    tt = np.empty((T, 3), np.int64)
    L.orc_topology_copy(h, _ptr(ep), _ptr(et), _ptr(tp), _ptr(te), _ptr(tt))
    L.orc_topology_free(h)
    return dict(ep=ep, et=et, tp=tp, te=te, tt=tt)


class OracleSolver:
    """SpaceDisc + TimeDisc + Solvers restated on the CPU for one mesh."""

    def __init__(self, xyz, topo, flux="HLLC", wave="Einfeldt", cor=0.0, tau=0.0, opts=0):
        L = lib()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.topo = {k: np.ascontiguousarray(v, dtype=np.int64) for k, v in topo.items() if v is not None}
        t = self.topo
        self.V = self.xyz.shape[0]
        self.E = t["ep"].shape[0]
        self.T = t["tp"].shape[0]
        pe = t.get("pe")
        self.h = L.orc_create(self.V, _ptr(self.xyz), self.E, _ptr(t["ep"]), _ptr(t["et"]),
                              self.T, _ptr(t["tp"]), _ptr(t["te"]), _ptr(t["tt"]),
                              FLUX[flux], WAVE[wave], cor, tau, opts, _ptr(pe) if pe is not None else None)
        if not self.h:
            raise ValueError("unsupported boundary tag (SOLID_WALL = -1, FREE_FLOW = -2 and paired "
                             "PERIODIC = -3 only)")

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:  # lib is None at interpreter exit
            lib().orc_destroy(self.h)
            self.h = None

    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.T, 3)
        lib().orc_set_state(self.h, _ptr(prim))

    def state(self):
        out = np.empty((self.T, 3), np.float64)
        lib().orc_get_state(self.h, _ptr(out))
        return out

    def compute_interface_values(self):
        lib().orc_compute_interface_values(self.h)

    def compute_fluxes(self):
        return lib().orc_compute_fluxes(self.h)

    def edge_field(self):
        out = np.empty((2 * self.E, 3), np.float64)
        lib().orc_get_edge_field(self.h, _ptr(out))
        return out

    def src_field(self):
        out = np.empty((2 * self.E, 3), np.float64)
        lib().orc_get_src_field(self.h, _ptr(out))
        return out

    def fluxes(self):
        out = np.empty((self.E, 3), np.float64)
        lib().orc_get_fluxes(self.h, _ptr(out))
        return out

    def max_wp(self):
        out = np.empty(self.V, np.float64)
        lib().orc_get_max_wp(self.h, _ptr(out))
        return out

    def classify(self):
        out = np.empty(self.T, np.int8)
        lib().orc_classify(self.h, _ptr(out))
        return out

    def reconstruct(self, i, kind):
        out = np.empty(5, np.float64)
        lib().orc_reconstruct(self.h, i, kind, _ptr(out))
        return out

    def integrals(self, per_cell=False):
        """(mass, kinetic, potential) of the current state's reconstruction,
        summed in cell order; with per_cell also the (T, 3) cell terms."""
        out = np.zeros(3)
        pc = np.zeros((self.T, 3)) if per_cell else None
        lib().orc_integrals(self.h, _ptr(out), _ptr(pc) if per_cell else None)
        return (out, pc) if per_cell else out

    def rhs(self, i, dt):
        out = np.empty(3, np.float64)
        lib().orc_rhs(self.h, i, dt, _ptr(out))
        return out

    def run(self, scheme="SSPRK2", nsteps=1, dt=1e-3, cfl=0.15):
        """dt > 0: fixed step; dt <= 0: CFL mode dt = cfl * r. Returns the dts."""
        dts = np.empty(nsteps, np.float64)
        lib().orc_run(self.h, SCHEME[scheme], dt, cfl, nsteps, _ptr(dts))
        return dts

    def edge_flux(self, flux, wave, e, frm, to, r=1.0):
        """Fluxes::HLL / HLLC of edge e from cell frm towards to on the
        current edge slots: (F, min(r, candidate))."""
        F = np.empty(3, np.float64)
        rr = np.array([r], np.float64)
        lib().orc_edge_flux(self.h, FLUX[flux], WAVE[wave], int(e), int(frm), int(to), _ptr(rr), _ptr(F))
        return F, float(rr[0])

    def time_run(self, scheme="SSPRK2", nsteps=1, dt=1e-3, cfl=0.15):
        return lib().orc_time_run(self.h, SCHEME[scheme], dt, cfl, nsteps)


def cbrt(x):
    return lib().orc_cbrt(x)


def ilog2_trunc(x):
    return lib().orc_ilog2_trunc(x)


def bisection_cubic(b, c, d, xmin, xmax):
    return lib().orc_bisection_cubic(b, c, d, xmin, xmax)


def gradient(p):
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(3, 3)
    g = np.empty(2, np.float64)
    lib().orc_gradient(_ptr(p), _ptr(g))
    return g


def flux_edge(flux, wave, xyz, left, right=None, to_tag=0, cor=0.0):
    """Flux of the edge A B of a two-cell mesh through the oracle's
    ComputeFluxes (Solver::HLL / HLLC, walls, free flow).  xyz: 4 x 3 nodes
    A, B, C, D (cells A B C = `from` and B A D = `to`, counter-clockwise);
    left / right: primitive (w, u, v) edge states; to_tag < 0: boundary
    edge of that tag with no `to` cell (xyz then needs 3 rows).
    Returns (F, n, r): flux, Norm(e, from) and the pass's CFL length."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    Lp = np.ascontiguousarray(left, dtype=np.float64)
    Rp = np.ascontiguousarray(left if right is None else right, dtype=np.float64)
    F, n, r = np.empty(3), np.empty(2), np.empty(1)
    rc = lib().orc_flux_edge(FLUX[flux], WAVE[wave], cor, _ptr(xyz), _ptr(Lp), _ptr(Rp), to_tag,
                             _ptr(F), _ptr(n), _ptr(r))
    if rc != 0:
        raise ValueError("orc_flux_edge: bad two-cell mesh")
    return F, n, float(r[0])


def wavespeed(wave, ul, hl, ur, hr):
    a = np.empty(2, np.float64)
    lib().orc_wavespeed(WAVE[wave], ul, hl, ur, hr, _ptr(a))
    return a
