"""Integrals and error norms on the CPU: the oracle's ComputeIntegrals
restatement (src/SpaceDisc.cpp:77-103) pinned by closed forms and by a fine
quadrature, and the numpy Compare / norms pinned by exact solutions."""
import numpy as np
import pytest

from oracle import oracle as O
from swe_fvm_amd import analytic as A
from tests import _common as C


def _setup(n=20, seed=5):
    xyz, topo, prim = C.wet_dry_mesh(n, seed=seed)
    t = topo.tables()
    orc = O.OracleSolver(xyz, t, "HLLC", "Einfeldt")
    orc.set_state(prim)
    return xyz, t, prim, orc


def test_integrals_full_wet_cells_closed_form():
    """Full-wet cells: area * (h, 1/2 h |u|^2, 1/2 h^2 + h b) of the cell
    average; dry cells contribute nothing; the total is the cell-order sum."""
    xyz, t, prim, orc = _setup()
    tot, pc = orc.integrals(per_cell=True)
    orc.compute_interface_values()
    cls = orc.classify()
    tp = t["tp"]
    Ar, b = C.cell_area(xyz, tp), C.cell_bath(xyz, tp)
    h = prim[:, 0] - b
    u, v = prim[:, 1], prim[:, 2]
    full, dry, part = cls == 2, cls == 0, cls == 1
    assert full.sum() > 50 and dry.sum() > 10 and part.sum() > 10
    assert np.array_equal(pc[full, 0], Ar[full] * h[full])
    assert np.array_equal(pc[full, 1], Ar[full] * (0.5 * h[full] * (u[full] ** 2 + v[full] ** 2)))
    assert np.array_equal(pc[full, 2], Ar[full] * (0.5 * h[full] * h[full] + h[full] * b[full]))
    assert np.all(pc[dry] == 0.0)
    assert np.array_equal(pc[part, 0], Ar[part] * h[part])
    s = np.zeros(3)
    for row in pc:
        s += row
    assert np.array_equal(tot, s)


def test_integrals_partially_wet_cells_against_fine_quadrature():
    """Partially wet cells: TriangAverage<2,10> of the energies of the
    pass-2 reconstruction, against the same integrand at 200^2 points, and
    the mass term against the volume the reconstruction holds."""
    xyz, t, prim, orc = _setup()
    _, pc = orc.integrals(per_cell=True)
    cls = orc.classify()
    tp = t["tp"]
    Ar, b = C.cell_area(xyz, tp), C.cell_bath(xyz, tp)
    part = np.nonzero(cls == 1)[0]
    P = [xyz[tp[part, k]] for k in range(3)]
    cen = (P[0] * (1 / 3) + P[1] * (1 / 3)) + P[2] * (1 / 3)
    mu = np.array([orc.reconstruct(int(i), 4) for i in part])  # w0, u, v, gx, gy

    def integrand(q):
        a = mu[:, 0] + (mu[:, 3] * (q[:, 0] - cen[:, 0]) + mu[:, 4] * (q[:, 1] - cen[:, 1]))
        wet = (a - q[:, 2]) >= 0
        a = np.where(wet, a, q[:, 2])
        hq = a - q[:, 2]
        uu, vv = np.where(wet, mu[:, 1], 0.0), np.where(wet, mu[:, 2], 0.0)
        return np.stack([hq, 0.5 * hq * (uu * uu + vv * vv), 0.5 * hq * hq + hq * q[:, 2]])

    fine = A.triang_average(P[0], P[1], P[2], integrand, 200) * Ar[part]
    coarse = A.triang_average(P[0], P[1], P[2], integrand, 10) * Ar[part]
    scale = np.abs(pc[part, 1:]).max(0)
    # the oracle's sum is this quadrature; both are close to the fine one
    assert np.max(np.abs(coarse[1:].T - pc[part, 1:]) / scale) < 1e-12
    assert np.max(np.abs(fine[1:].T - pc[part, 1:]) / scale) < 2e-2
    # the reconstruction holds the cell's water (volume identity)
    assert np.max(np.abs(fine[0] - Ar[part] * (prim[part, 0] - b[part])) / Ar[part]) < 2e-3


def test_integrals_lake_at_rest_are_steady():
    """Lake at rest: mass and energies do not change over steps; kinetic 0."""
    from swe_fvm_amd import workloads as W
    dom, prim = W.lake_at_rest(n=30, seed=2)
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
    orc.set_state(prim)
    i0 = orc.integrals()
    orc.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    i1 = orc.integrals()
    assert i0[1] == 0.0 and abs(i1[1]) < 1e-24
    assert abs(i1[0] - i0[0]) <= 1e-13 * i0[0]
    assert abs(i1[2] - i0[2]) <= 1e-12 * abs(i0[2])


class _FakeSD:
    """The part of SpaceDisc that compare() reads, fed from the oracle."""

    def __init__(self, dom, orc, prim):
        self.dom, self.orc, self.prim = dom, orc, prim

    def ComputeInterfaceValues(self):
        self.orc.compute_interface_values()

    def CellClass(self):
        return self.orc.classify()

    def GetMuscl(self):
        cls = self.orc.classify()
        kind = np.array([0, 4, 2])[cls]
        mu = np.array([self.orc.reconstruct(i, int(k)) for i, k in enumerate(kind)])
        return mu[:, 0], mu[:, 3], mu[:, 4]

    def GetVolField(self):
        return self.prim

    def GetDomain(self):
        return self.dom


def test_error_norms_vanish_for_the_exact_lake():
    """compare() of a flat lake against LakeAtRest-like exact data: the
    reconstruction is exact, so every norm is at rounding level; a wrong
    exact solution gives an O(1) error."""
    from swe_fvm_amd import workloads as W
    dom, prim = W.lake_at_rest(n=16, seed=4)
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
    orc.set_state(prim)
    level = float(np.median(prim[:, 0]))

    class Flat(A.Analytic):
        def __init__(self, lvl):
            super().__init__()
            self.lvl = lvl

        def b(self, x, y):  # unused by compare: the point's z is the bathymetry
            return np.zeros_like(x)

        def h(self, x, y, t):
            return np.zeros_like(x)

        def w(self, x, y, t):
            return np.full_like(x, self.lvl)

        def u(self, x, y, t):
            return np.zeros_like(x)

        v = u

    sd = _FakeSD(dom, orc, prim)
    cls = orc.classify()
    wet = cls == 2
    e = A.compare(sd, Flat(level), 0.0, n_partial=20)
    assert np.max(e[wet]) < 1e-12
    e2 = A.compare(sd, Flat(level + 0.1), 0.0, n_partial=20)
    assert np.allclose(e2[wet, 0], 0.1, atol=1e-12)


def test_classic_thacker_initial_state_error_is_small_and_converges():
    """The TriangAverage'd Thacker state (examples/Main.cpp:213-222) against
    its own analytic solution at t = 0: the limited linear reconstruction of
    the smooth paraboloid converges at second order (L1 error of w / 4 when
    the mesh is refined twice)."""
    from swe_fvm_amd import workloads as W
    errs = []
    for n in (20, 40):
        dom, prim, th = W.thacker(n=n, quad=10)
        orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
        orc.set_state(prim)
        nrm = A.error_norms(_FakeSD(dom, orc, prim), th, 0.0, n_partial=10)
        errs.append(nrm["L1"][0])
    assert errs[0] < 2e-2 and errs[1] < 0.35 * errs[0], errs


def test_stack_domains_tables():
    """The stacked domain of a batch: sizes add up, member k's tables are
    its own shifted by the offsets, boundary tags are kept."""
    import swe_fvm_amd as swe
    a = swe.Domain(*swe.jittered_mesh(6, 0.2, 1))
    xyz, topo = swe.jittered_mesh(5, 0.2, 2)
    b = swe.Domain(xyz, topo.with_boundary_tag(np.nonzero(topo.tables()["et"][:, 1] < 0)[0][:3], swe.FREE_FLOW))
    big, off = swe.stack_domains([a, b, a])
    t, ta, tb = big.topology.tables(), a.topology.tables(), b.topology.tables()
    assert list(off["cells"]) == [0, 72, 122, 194] and list(off["nodes"]) == [0, 49, 85, 134]
    assert big.topology.NumEdges() == 2 * a.topology.NumEdges() + b.topology.NumEdges()
    e0, e1 = off["edges"][1], off["edges"][2]
    assert np.array_equal(t["tp"][72:122], tb["tp"] + 49)
    assert np.array_equal(t["te"][72:122], tb["te"] + e0)
    et = t["et"][e0:e1]
    assert np.array_equal(et[:, 0], tb["et"][:, 0] + 72)
    assert np.array_equal(et[:, 1], np.where(tb["et"][:, 1] >= 0, tb["et"][:, 1] + 72, tb["et"][:, 1]))
    assert np.sum(et[:, 1] == swe.FREE_FLOW) == 3
    assert np.array_equal(t["tt"][122:], np.where(ta["tt"] >= 0, ta["tt"] + 122, ta["tt"]))
    assert np.array_equal(big.geometry[85:], a.geometry)
