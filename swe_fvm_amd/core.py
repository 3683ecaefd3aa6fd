"""Host-side mirror of SWE_FVM's operator interface over the C ABI.

Names, argument meaning and error behaviour follow the reference headers:
``Topology`` (include/TriangMesh.h:14-65), ``Domain`` (include/Bathymetry.h:12-49),
``SpaceDisc`` (include/SpaceDisc.h:20-48), ``TimeDisc`` (include/TimeDisc.h:4-23),
``Solvers`` (include/Solvers.h:4-10), ``Fluxes``/``Wavespeeds``
(include/Fluxes.h:4-113).  All numerics run in the HIP library; this module
only moves arrays and maps status codes to the reference's exception types
(include/Exceptions.h:4-10).
"""
import ctypes

import numpy as np

from . import _lib

SWE_OK = 0
SWE_ERR_ARG = 1
SWE_ERR_DEVICE = 2
SWE_ERR_SOLVER = 3
SWE_ERR_DOMAIN = 4

SCHEMES = {"Euler": 0, "SSPRK2": 1, "SSPRK3": 2}
OPT_HEAD_GRADIENT = 1
OPT_HEAD_PARTWET2 = 2
OPT_FRICTION = 16  # linear friction -tau (hu, hv); the reference leaves tau unused
OPT_DRAIN_LENGTH = 32  # draining dt from the outflow F0 * L (the reference omits L)
OPT_EXACT_BALANCE = 64  # flux / RHS groupings that keep equal edge states at rest to the bit
SOLID_WALL = -1  # boundary tags, include/Includes.h:21
FREE_FLOW = -2  # transmissive edge (declared but unimplemented in the reference)
PERIODIC = -3  # paired edges (declared but never paired in the reference): Topology.with_periodic


class DomainError(RuntimeError):
    """include/Exceptions.h:4-6"""


class SolverError(RuntimeError):
    """include/Exceptions.h:8-10"""


class DeviceError(RuntimeError):
    """HIP runtime failure."""


def _check(rc):
    if rc == SWE_OK:
        return
    msg = _lib.load().swe_last_error().decode(errors="replace")
    if rc == SWE_ERR_DOMAIN:
        raise DomainError(msg)
    if rc == SWE_ERR_SOLVER:
        raise SolverError(msg)
    if rc == SWE_ERR_DEVICE:
        raise DeviceError(msg)
    raise ValueError(msg)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a.reshape(shape)


# ------------------------------------------------------------------ fluxes --

class _WaveSpeed:
    def __init__(self, name, code):
        self.name, self.code = name, code

    def __repr__(self):
        return f"Wavespeeds::{self.name}"


class Wavespeeds:
    """Wave-speed estimates (src/Fluxes.cpp:5-26)."""
    Rusanov = _WaveSpeed("Rusanov", 0)
    Davis = _WaveSpeed("Davis", 1)
    Einfeldt = _WaveSpeed("Einfeldt", 2)


class Fluxer:
    """A numerical flux instantiation, e.g. ``Fluxes.HLLC[Wavespeeds.Einfeldt]``."""

    def __init__(self, name, code, ws):
        if not isinstance(ws, _WaveSpeed):
            raise TypeError("flux template argument must be a Wavespeeds member")
        self.name, self.code, self.ws = name, code, ws

    def __repr__(self):
        return f"Fluxes::{self.name}<{self.ws!r}>"


class _FluxTemplate:
    def __init__(self, name, code):
        self.name, self.code = name, code

    def __getitem__(self, ws):
        return Fluxer(self.name, self.code, ws)


class Fluxes:
    """``Fluxes::HLL<W>`` and ``Fluxes::HLLC<W>`` (include/Fluxes.h:14-111)."""
    HLL = _FluxTemplate("HLL", 0)
    HLLC = _FluxTemplate("HLLC", 1)


# ---------------------------------------------------------------- topology --

class Topology:
    """Owning mesh tables (include/TriangMesh.h:14-65).

    Unlike the reference, which keeps references to the caller's arrays, the
    tables are copied into the library.
    """

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def create(numNodes, edgeNodes, edgeElements, elementNodes, elementEdges, elementNeighbours):
        """Topology::create (include/TriangMesh.h:16-25, pybind/Topology.cpp:21-27)."""
        L = _lib.load()
        en = _i64(edgeNodes, (-1, 2))
        ee = _i64(edgeElements, (-1, 2))
        tn = _i64(elementNodes, (-1, 3))
        te = _i64(elementEdges, (-1, 3))
        tt = _i64(elementNeighbours, (-1, 3))
        if en.shape != ee.shape or tn.shape != te.shape or tn.shape != tt.shape:
            raise ValueError("table shapes disagree")
        h = ctypes.c_void_p()
        _check(L.swe_topology_create(int(numNodes), en.shape[0], _ptr(en), _ptr(ee), tn.shape[0],
                                     _ptr(tn), _ptr(te), _ptr(tt), ctypes.byref(h)))
        return Topology(h)

    @staticmethod
    def build(numNodes, triangles, boundaryLines):
        """Tables by the reference mesh reader's conventions from raw triangles."""
        L = _lib.load()
        tri = _i64(triangles, (-1, 3))
        bnd = _i64(boundaryLines, (-1, 2))
        h = ctypes.c_void_p()
        _check(L.swe_topology_build(int(numNodes), tri.shape[0], _ptr(tri), bnd.shape[0],
                                    _ptr(bnd), ctypes.byref(h)))
        return Topology(h)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().swe_topology_destroy(self._h)
            self._h = None

    def NumNodes(self):
        return _lib.load().swe_topology_num_nodes(self._h)

    def NumEdges(self):
        return _lib.load().swe_topology_num_edges(self._h)

    def NumTriangles(self):
        return _lib.load().swe_topology_num_elements(self._h)

    # pybind names (pybind/Topology.cpp:28-30)
    num_nodes = NumNodes
    num_edges = NumEdges
    num_elements = NumTriangles

    def tables(self):
        """dict of EdgePoints, EdgeTriangs, TriangPoints, TriangEdges, TriangTriangs tables."""
        E, T = self.NumEdges(), self.NumTriangles()
        ep = np.empty((E, 2), np.int64)
        et = np.empty((E, 2), np.int64)
        tp = np.empty((T, 3), np.int64)
        te = np.empty((T, 3), np.int64)
        tt = np.empty((T, 3), np.int64)
        _check(_lib.load().swe_topology_copy(self._h, _ptr(ep), _ptr(et), _ptr(tp), _ptr(te), _ptr(tt)))
        out = dict(ep=ep, et=et, tp=tp, te=te, tt=tt)
        if np.any(et[:, 1] == PERIODIC):  # the partner table of periodic pairs
            pe = np.empty(E, np.int64)
            _check(_lib.load().swe_topology_periodic(self._h, _ptr(pe)))
            out["pe"] = pe
        return out

    def IsEdgeBoundary(self, e):
        return bool(self.tables()["et"][e, 1] < 0)

    @staticmethod
    def _pairs(_unused, t):
        """(n, 2) periodic pairs (lower edge first) of tables t."""
        if "pe" not in t:
            return np.zeros((0, 2), np.int64)
        e = np.nonzero(t["pe"] >= 0)[0]
        e = e[e < t["pe"][e]]
        return np.stack([e, t["pe"][e]], 1)

    @staticmethod
    def _with_tables(num_nodes, t, pairs):
        """A new Topology of tables t (periodic tags set by pairs)."""
        et = t["et"].copy()
        et[et[:, 1] == PERIODIC, 1] = SOLID_WALL
        cp = Topology.create(num_nodes, t["ep"], et, t["tp"], t["te"], t["tt"])
        if pairs.shape[0]:
            pairs = np.ascontiguousarray(pairs, dtype=np.int64)
            _check(_lib.load().swe_topology_set_periodic(cp._h, pairs.shape[0], _ptr(pairs)))
        return cp

    def with_periodic(self, pairs):
        """A copy whose boundary edges pairs[k][0] and pairs[k][1] are PERIODIC
        partners (swe_topology_set_periodic): the outside state of each is
        the inside state of the other.  Existing pairs are kept."""
        t = self.tables()
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        return Topology._with_tables(self.NumNodes(), t, np.concatenate([Topology._pairs(None, t), pairs]))

    def with_boundary_tag(self, edges, tag):
        """A copy whose boundary edges `edges` carry `tag` (SOLID_WALL or
        FREE_FLOW) in EdgeTriangs[e][1] (include/Includes.h:21)."""
        t = self.tables()
        edges = np.asarray(edges, dtype=np.int64)
        if np.any(t["et"][edges, 1] >= 0):
            raise DomainError("only boundary edges carry a tag")
        if "pe" in t and np.any(t["pe"][edges] >= 0):
            raise DomainError("periodic edges keep their tag")
        if int(tag) == PERIODIC:
            raise DomainError("PERIODIC edges come in pairs: Topology.with_periodic")
        t["et"][edges, 1] = int(tag)
        return Topology._with_tables(self.NumNodes(), t, Topology._pairs(None, t))


def periodic_pairs(xyz, topo, shift, tol=1e-9):
    """Pairs of boundary edges that map onto each other by the translation
    shift = (dx, dy): edge a (midpoint m) pairs with the boundary edge b
    whose midpoint lies within tol (relative to the domain size) of m +
    shift.  Returns an (n, 2) array for Topology.with_periodic."""
    t = topo.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    mid = 0.5 * (xyz[t["ep"][b, 0], :2] + xyz[t["ep"][b, 1], :2])
    shift = np.asarray(shift, dtype=np.float64)
    eps = tol * max(1.0, float(np.abs(mid).max()))
    q = 1e3 * eps  # hash cells much larger than the tolerance; neighbours checked
    grid = {}
    for k, m in enumerate(mid):
        grid.setdefault(tuple(np.floor(m / q).astype(np.int64)), []).append(k)
    pairs = []
    for k, m in enumerate(mid):
        target = m + shift
        c = np.floor(target / q).astype(np.int64)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for j in grid.get((c[0] + dx, c[1] + dy), ()):
                    if np.all(np.abs(mid[j] - target) <= eps):
                        pairs.append((b[k], b[j]))
    return np.array(pairs, dtype=np.int64).reshape(-1, 2)


def stack_domains(domains):
    """One Domain holding the given domains side by side (a disjoint union:
    node, edge and cell indices of member k offset by the sizes of members
    0..k-1; boundary tags kept).  Every stepping kernel then advances all
    members in one launch, and with a fixed dt each member's cells evolve
    bit for bit as in a run of that member alone.  Returns (Domain, offsets)
    with offsets = dict(nodes=, edges=, cells=) arrays of length n + 1."""
    tabs = [d.topology.tables() for d in domains]
    nv = np.cumsum([0] + [d.geometry.shape[0] for d in domains])
    ne = np.cumsum([0] + [t["ep"].shape[0] for t in tabs])
    nt = np.cumsum([0] + [t["tp"].shape[0] for t in tabs])

    def cells(a, k):  # cell ids >= 0 shift, boundary tags stay
        return np.where(a >= 0, a + nt[k], a)
    ep = np.concatenate([t["ep"] + nv[k] for k, t in enumerate(tabs)])
    et = np.concatenate([cells(t["et"], k) for k, t in enumerate(tabs)])
    tp = np.concatenate([t["tp"] + nv[k] for k, t in enumerate(tabs)])
    te = np.concatenate([t["te"] + ne[k] for k, t in enumerate(tabs)])
    tt = np.concatenate([cells(t["tt"], k) for k, t in enumerate(tabs)])
    pairs = np.concatenate([Topology._pairs(None, t) + ne[k] for k, t in enumerate(tabs)])
    topo = Topology._with_tables(int(nv[-1]), dict(ep=ep, et=et, tp=tp, te=te, tt=tt), pairs)
    geo = np.concatenate([d.geometry for d in domains])
    return Domain(geo, topo), dict(nodes=nv, edges=ne, cells=nt)


def _spread_bits(x):
    """The bits of x (< 2^32) moved to the even bit positions of a 64-bit word."""
    x = x.astype(np.uint64)
    for shift, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF), (4, 0x0F0F0F0F0F0F0F0F),
                        (2, 0x3333333333333333), (1, 0x5555555555555555)):
        x = (x | (x << np.uint64(shift))) & np.uint64(mask)
    return x


def reorder_cells(domain, perm):
    """The domain with its cells in the order perm (new cell k = old cell
    perm[k], vertex order kept), nodes renumbered by first appearance and the
    tables rebuilt by the reference reader's conventions; boundary edges keep
    their order and tags.  Returns (Domain, perm, node_perm) with new node m =
    old node node_perm[m]."""
    t = domain.topology.tables()
    perm = np.asarray(perm, dtype=np.int64)
    V = domain.geometry.shape[0]
    tri = t["tp"][perm]
    flat = tri.ravel()
    _, first = np.unique(flat, return_index=True)
    used = flat[np.sort(first)]
    rest = np.setdiff1d(np.arange(V), used, assume_unique=True)
    node_perm = np.concatenate([used, rest]).astype(np.int64)
    new_id = np.empty(V, np.int64)
    new_id[node_perm] = np.arange(V)
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    topo = Topology.build(V, new_id[tri], new_id[t["ep"][b]])
    tags = t["et"][b, 1]
    for tag in np.unique(tags[(tags != SOLID_WALL) & (tags != PERIODIC)]):
        topo = topo.with_boundary_tag(np.nonzero(tags == tag)[0], int(tag))
    if "pe" in t:  # boundary edge b[k] is edge k of the rebuilt tables
        new_e = np.full(t["ep"].shape[0], -1, np.int64)
        new_e[b] = np.arange(b.shape[0])
        topo = topo.with_periodic(new_e[Topology._pairs(None, t)])
    return Domain(domain.geometry[node_perm], topo), perm, node_perm


def morton_order(domain):
    """Cells along the Morton curve of their centroids (21 bits per axis),
    nodes by first appearance: the layout the kernels are designed for
    (neighbours of a workgroup's cells share its L2 lines and LDS tile).
    Returns reorder_cells' (Domain, perm, node_perm).  The reference's
    results depend on the numbering only through which cell of an edge is
    et0 (the flux orientation), i.e. at round-off for symmetric wave speeds
    and through the Einfeldt estimate's asymmetry otherwise."""
    c = domain.centroids()[:, :2]
    lo = c.min(axis=0)
    span = c.max(axis=0) - lo
    span[span == 0] = 1.0
    q = np.clip((c - lo) / span * 2097151.0, 0, 2097151).astype(np.uint64)
    key = _spread_bits(q[:, 0]) | (_spread_bits(q[:, 1]) << np.uint64(1))
    return reorder_cells(domain, np.argsort(key, kind="stable"))


def _take_mesh(rc, xyz_p, nn, h):
    _check(rc)
    L = _lib.load()
    n = nn.value
    buf = (ctypes.c_double * (3 * n)).from_address(xyz_p.value)
    xyz = np.array(np.frombuffer(buf, dtype=np.float64).reshape(n, 3))
    L.swe_free(xyz_p)
    return xyz, Topology(h)


def read_msh(path):
    """Gmsh 4.1 reader (old TriangMesh(filename)); returns (xyz V x 3, Topology)."""
    L = _lib.load()
    xyz_p, nn, h = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
    rc = L.swe_topology_read_msh(str(path).encode(), ctypes.byref(xyz_p), ctypes.byref(nn), ctypes.byref(h))
    return _take_mesh(rc, xyz_p, nn, h)


def StructTriangMesh(ni, nj, h, x0=0.0, y0=0.0):
    """StructTriangMesh(ni, nj, h) (include/StructTriangMesh.h:4-15): 4 triangles per square."""
    L = _lib.load()
    xyz_p, nn, hh = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
    rc = L.swe_mesh_struct(int(ni), int(nj), float(h), float(x0), float(y0),
                           ctypes.byref(xyz_p), ctypes.byref(nn), ctypes.byref(hh))
    return _take_mesh(rc, xyz_p, nn, hh)


def jittered_mesh(n, jitter=0.25, seed=0, x0=0.0, y0=0.0, L=1.0):
    """Seeded Delaunay mesh of an n x n jittered grid, Morton-ordered cells."""
    lib = _lib.load()
    xyz_p, nn, hh = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
    rc = lib.swe_mesh_jittered(int(n), float(jitter), int(seed) & 0xFFFFFFFFFFFFFFFF, float(x0),
                               float(y0), float(L), ctypes.byref(xyz_p), ctypes.byref(nn), ctypes.byref(hh))
    return _take_mesh(rc, xyz_p, nn, hh)


class Domain:
    """Node geometry (x, y, bathymetry z) on a Topology (include/Bathymetry.h:12-49)."""

    def __init__(self, geometry, topology):
        g = np.ascontiguousarray(geometry, dtype=np.float64)
        if g.ndim != 2 or g.shape[1] != 3:
            raise ValueError("geometry must be V x 3 (x, y, z)")
        if g.shape[0] != topology.NumNodes():
            raise DomainError("geometry and topology disagree on the number of nodes")
        self.geometry = g
        self.topology = topology

    def GetTopology(self):
        return self.topology

    def GetGeometry(self):
        return self.geometry

    def centroids(self):
        """Domain::T for every cell ((p0 c + p1 c) + p2 c, c = 1/3)."""
        tp = self.topology.tables()["tp"]
        c = 1.0 / 3.0
        P = self.geometry
        return (P[tp[:, 0]] * c + P[tp[:, 1]] * c) + P[tp[:, 2]] * c

    def areas(self):
        """Domain::Area for every cell (TriangArea, src/PointOperations.cpp:12-14)."""
        tp = self.topology.tables()["tp"]
        a, b, c = (self.geometry[tp[:, k]] for k in range(3))
        ux, uy, vx, vy = b[:, 0] - a[:, 0], b[:, 1] - a[:, 1], c[:, 0] - a[:, 0], c[:, 1] - a[:, 1]
        return 0.5 * np.abs(ux * vy - vx * uy)


# --------------------------------------------------------------- SpaceDisc --

class SpaceDisc:
    """SpaceDisc(fluxer, domain, v0, cor, tau) (include/SpaceDisc.h:20-48).

    v0 is the primitive VolumeField (w, u, v) per cell, T x 3.  The field and
    all work buffers live on the GPU; getters download them.
    """

    def __init__(self, fluxer, domain, v0, cor=0.0, tau=0.0, options=0, device=0):
        if not isinstance(fluxer, Fluxer):
            raise SolverError(f"unsupported fluxer {fluxer!r}: use Fluxes.HLL[...] or Fluxes.HLLC[...]")
        L = _lib.load()
        self.fluxer = fluxer
        self.domain = domain
        self.cor, self.tau = float(cor), float(tau)
        h = ctypes.c_void_p()
        _check(L.swe_create(domain.topology._h, _ptr(domain.geometry), fluxer.code, fluxer.ws.code,
                            self.cor, self.tau, int(options), int(device), ctypes.byref(h)))
        self._h = h
        self.T = L.swe_num_cells(h)
        self.E = L.swe_num_edges(h)
        self.V = L.swe_num_nodes(h)
        self.SetVolField(v0)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().swe_destroy(self._h)
            self._h = None

    # VolumeField access
    def SetVolField(self, prim):
        p = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.T, 3)
        _check(_lib.load().swe_set_state(self._h, _ptr(p)))

    def GetVolField(self):
        out = np.empty((self.T, 3), np.float64)
        _check(_lib.load().swe_get_state(self._h, _ptr(out)))
        return out

    def GetDomain(self):
        return self.domain

    def ComputeInterfaceValues(self):
        _check(_lib.load().swe_compute_interface_values(self._h))

    def ComputeFluxes(self):
        _check(_lib.load().swe_compute_fluxes(self._h, None))

    def GetMinLenToWavespeed(self):
        r = ctypes.c_double()
        _check(_lib.load().swe_min_len_to_wavespeed(self._h, ctypes.byref(r)))
        return r.value

    def GetEdgField(self):
        out = np.empty((2 * self.E, 3), np.float64)
        _check(_lib.load().swe_get_edge_field(self._h, _ptr(out)))
        return out

    def GetSrcField(self):
        out = np.empty((2 * self.E, 3), np.float64)
        _check(_lib.load().swe_get_src_field(self._h, _ptr(out)))
        return out

    def GetFluxes(self):
        out = np.empty((self.E, 3), np.float64)
        _check(_lib.load().swe_get_fluxes(self._h, _ptr(out)))
        return out

    def GetMaxWp(self):
        out = np.empty(self.V, np.float64)
        _check(_lib.load().swe_get_max_wp(self._h, _ptr(out)))
        return out

    def CellClass(self):
        """0 dry, 1 partially wet, 2 fully wet (src/MUSCLObject.cpp:13-29)."""
        out = np.empty(self.T, np.int8)
        _check(_lib.load().swe_get_cell_class(self._h, _ptr(out)))
        return out

    def GetMuscl(self):
        """Final per-cell reconstruction after ComputeInterfaceValues:
        (w0, gx, gy) arrays, origin at the centroid."""
        w0, gx, gy = (np.empty(self.T, np.float64) for _ in range(3))
        _check(_lib.load().swe_get_muscl(self._h, _ptr(w0), _ptr(gx), _ptr(gy)))
        return w0, gx, gy

    def Integrals(self, per_cell=False):
        """ComputeIntegrals (src/SpaceDisc.cpp:77-103, commented out in the
        reference): (mass, kinetic, potential) of the current state, computed
        on the device; per_cell also returns the (T, 3) cell terms."""
        out = np.zeros(3, np.float64)
        pc = np.zeros((self.T, 3), np.float64) if per_cell else None
        _check(_lib.load().swe_integrals(self._h, _ptr(out), _ptr(pc) if per_cell else None))
        return (out, pc) if per_cell else out

    def set_bathymetry(self, z):
        """Replaces the node bathymetry (V values) on the device and recomputes
        the static geometry; the state (w, u, v) is kept."""
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(self.V)
        _check(_lib.load().swe_set_bathymetry(self._h, _ptr(z)))
        # GetDomain() reflects the device bathymetry (a new Domain: the
        # caller's geometry array is left alone)
        g = self.domain.geometry.copy()
        g[:, 2] = z
        self.domain = Domain(g, self.domain.topology)

    def adjoint_run(self, final_bar, nsteps, scheme="SSPRK2", dt=None, dts=None, checkpoint_every=50,
                    want_z=True):
        """Runs nsteps steps from the current state with fixed step sizes (dt,
        or the sequence dts) and back-propagates final_bar (T x 3 adjoint of
        the final w, u, v): returns (adjoint of the initial state T x 3,
        adjoint of the node bathymetry V or None, [forward ms, reverse ms]).
        The final state is current afterwards, as after run()."""
        fb = np.ascontiguousarray(final_bar, dtype=np.float64).reshape(self.T, 3)
        s0 = np.zeros((self.T, 3), np.float64)
        zb = np.zeros(self.V, np.float64) if want_z else None
        ms = np.zeros(2, np.float64)
        d = None
        if dts is not None:
            d = np.ascontiguousarray(dts, dtype=np.float64).reshape(int(nsteps))
        _check(_lib.load().swe_adjoint_run(self._h, SCHEMES[scheme], float(dt or 0.0),
                                           _ptr(d) if d is not None else None, int(nsteps),
                                           int(checkpoint_every), _ptr(fb), _ptr(s0),
                                           _ptr(zb) if want_z else None, _ptr(ms)))
        return s0, zb, ms

    def adjoint_timing(self):
        """Milliseconds of the last adjoint_run: forward sweep, the reverse
        sweep's recomputation of its segments from their checkpoints, and
        the rest of the reverse sweep (the reverse stages)."""
        out = np.zeros(3, np.float64)
        _check(_lib.load().swe_adjoint_timing(self._h, _ptr(out)))
        return {"forward_sweep_ms": float(out[0]), "recompute_ms": float(out[1]),
                "reverse_stages_ms": float(out[2])}

    def tangent_run(self, state0_dot, nsteps, z_dot=None, scheme="SSPRK2", dt=None, dts=None):
        """Runs nsteps steps from the current state with fixed step sizes,
        carrying the tangent (state0_dot: T x 3, z_dot: V or None) through
        the dual-number build of the stage kernels: returns (tangent of the
        final state T x 3, ms).  The final state is current afterwards and
        equals run()'s bit for bit."""
        sd0 = np.ascontiguousarray(state0_dot, dtype=np.float64).reshape(self.T, 3)
        zd = None if z_dot is None else np.ascontiguousarray(z_dot, dtype=np.float64).reshape(self.V)
        out = np.zeros((self.T, 3), np.float64)
        ms = np.zeros(1, np.float64)
        d = None
        if dts is not None:
            d = np.ascontiguousarray(dts, dtype=np.float64).reshape(int(nsteps))
        _check(_lib.load().swe_tangent_run(self._h, SCHEMES[scheme], float(dt or 0.0),
                                           _ptr(d) if d is not None else None, int(nsteps), _ptr(sd0),
                                           _ptr(zd) if zd is not None else None, _ptr(out), _ptr(ms)))
        return out, float(ms[0])

    def Stats(self):
        """Health of the current state (failure detection): dict with the
        number of cells holding a non-finite value, min / max depth, max
        speed and max |u| + sqrt(h) over wet cells."""
        out = np.zeros(5, np.float64)
        _check(_lib.load().swe_state_stats(self._h, _ptr(out)))
        return {"nonfinite": int(out[0]), "hmin": out[1], "hmax": out[2], "umax": out[3], "cmax": out[4]}

    def GetCor(self):
        return self.cor

    def GetTau(self):
        return self.tau

    # whole-run driver (no host synchronisation between steps)
    def run(self, scheme="SSPRK2", nsteps=1, dt=1e-3, cfl=0.15, sync=True):
        """nsteps steps; dt > 0 fixed, dt <= 0 CFL mode (dt = cfl * r)."""
        _check(_lib.load().swe_run(self._h, SCHEMES[scheme], float(dt), float(cfl), int(nsteps)))
        if sync:
            self.synchronize()

    def dts(self, nsteps):
        out = np.empty(nsteps, np.float64)
        _check(_lib.load().swe_get_dts(self._h, _ptr(out), int(nsteps)))
        return out

    def synchronize(self):
        _check(_lib.load().swe_synchronize(self._h))

    def time_run(self, scheme="SSPRK2", nsteps=1, dt=1e-3, cfl=0.15, per_kernel=False):
        """Milliseconds of nsteps steps (HIP events on the solver stream).

        per_kernel: also return the average per-launch milliseconds of the
        recon1, recon2, flux, drain, update, fused flux-draining-update and
        fused draining-update (after a CFL step's flux pass) kernels (0 for
        a kernel that did not run)."""
        ms = ctypes.c_double()
        if per_kernel:
            k = np.zeros(7, np.float64)
            _check(_lib.load().swe_time_run(self._h, SCHEMES[scheme], float(dt), float(cfl),
                                            int(nsteps), ctypes.byref(ms), _ptr(k)))
            return ms.value, k
        _check(_lib.load().swe_time_run(self._h, SCHEMES[scheme], float(dt), float(cfl), int(nsteps),
                                        ctypes.byref(ms), None))
        return ms.value

    def set_fused(self, on=True):
        """Uses (default) or not the fused tile kernel: flux-draining-update
        for the stages whose step size is known, draining-update after the
        flux pass of a CFL step's first stage; returns whether it is used.
        Results are identical either way."""
        act = ctypes.c_int()
        _check(_lib.load().swe_set_fused(self._h, 1 if on else 0, ctypes.byref(act)))
        return bool(act.value)

    def set_tuning(self, flux_items=2, update_items=1, flux_block=256, update_block=256):
        """Edges per thread of the flux kernel (1, 2, 4), cells per thread of
        the update kernel (1, 2) and threads per workgroup of either (256,
        512); results are identical for every choice."""
        _set_tuning(self._h, flux_items, update_items, flux_block, update_block)

    def tuning(self):
        return _tuning(self._h)

    def autotune(self, reps=10):
        """Times every flux / update instantiation on the current state and
        keeps the fastest; returns the choice and the per-launch times."""
        return _autotune(self._h, reps)

    def set_layout(self, block=0):
        """Device layout of the cell state: 0 = SoA (three arrays), a power
        of two B <= 64 = blocks of B cells (w, then u, then v; B = 1 is AoS).
        Results are identical for every layout."""
        _check(_lib.load().swe_set_layout(self._h, int(block)))

    def layout(self):
        return _layout(self._h)

    def autotune_layout(self, reps=10):
        """The autotuner with the layout dimension (swe_autotune_layout):
        returns the chosen layout, the tuning kept and the cost per layout."""
        return _autotune_layout(self._h, reps)


LAYOUTS = (0, 64, 8, 1)  # the candidates of swe_autotune_layout, in its order


def _layout(h):
    b = ctypes.c_int()
    _check(_lib.load().swe_get_layout(h, ctypes.byref(b)))
    return b.value


def _autotune_layout(h, reps):
    b = ctypes.c_int()
    c = np.zeros(len(LAYOUTS), np.float64)
    _check(_lib.load().swe_autotune_layout(h, int(reps), ctypes.byref(b), _ptr(c)))
    fi, ui, fb, ub = _tuning(h)
    return {"layout": b.value, "flux_items": fi, "update_items": ui, "flux_block": fb, "update_block": ub,
            "stage_ms": {("soa" if L == 0 else f"aosoa{L}"): float(x) for L, x in zip(LAYOUTS, c)}}


def _set_tuning(h, flux_items, update_items, flux_block, update_block):
    _check(_lib.load().swe_set_tuning(h, int(flux_items), int(update_items)))
    _check(_lib.load().swe_set_tuning_blocks(h, int(flux_block), int(update_block)))


def _tuning(h):
    """(flux items, update items, flux block, update block)"""
    fi, ui, fb, ub = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_lib.load().swe_get_tuning(h, ctypes.byref(fi), ctypes.byref(ui)))
    _check(_lib.load().swe_get_tuning_blocks(h, ctypes.byref(fb), ctypes.byref(ub)))
    return fi.value, ui.value, fb.value, ub.value


def _autotune(h, reps):
    fi, ui = ctypes.c_int(), ctypes.c_int()
    t = np.zeros(5, np.float64)
    _check(_lib.load().swe_autotune(h, int(reps), ctypes.byref(fi), ctypes.byref(ui), _ptr(t)))
    _, _, fb, ub = _tuning(h)
    return {"flux_items": fi.value, "update_items": ui.value, "flux_block": fb, "update_block": ub,
            "flux_ms": {1: t[0], 2: t[1], 4: t[2]}, "update_ms": {1: t[3], 2: t[4]}}


class TimeDisc:
    """TimeDisc(SpaceDisc*) (include/TimeDisc.h:4-23)."""

    m_constCFL = 0.15

    def __init__(self, sd=None):
        self.m_sd = sd

    def GetSpaceDisc(self):
        return self.m_sd

    def SetSpaceDisc(self, sd):
        self.m_sd = sd

    def CFLdt(self):
        return self.m_constCFL * self.m_sd.GetMinLenToWavespeed()

    def RHS(self, i, dt):
        out = np.empty(3, np.float64)
        _check(_lib.load().swe_rhs(self.m_sd._h, int(i), float(dt), _ptr(out)))
        return out


class Solvers:
    """Solvers::Euler / SSPRK2 / SSPRK3 (include/Solvers.h:6-8)."""

    @staticmethod
    def Euler(td, dt):
        td.GetSpaceDisc().run("Euler", 1, dt)

    @staticmethod
    def SSPRK2(td, dt):
        td.GetSpaceDisc().run("SSPRK2", 1, dt)

    @staticmethod
    def SSPRK3(td, dt):
        td.GetSpaceDisc().run("SSPRK3", 1, dt)
