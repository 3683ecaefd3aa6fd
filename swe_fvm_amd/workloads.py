"""Seeded synthetic workloads: mesh, bathymetry and initial state.

Equations are in the reference's non-dimensional form with g = 1
(include/SpaceDisc.h:14, src/Fluxes.cpp:6-23).  All generators are
deterministic in their seed.

* thacker      — classic Thacker paraboloid on StructTriangMesh
                 (examples/Tests.h:237-280, examples/Main.cpp:197-223)
* lake_at_rest — flat surface over Gaussian bumps, two emerging
* dam_break    — circular wet/dry dam break over a random Fourier bathymetry
"""
import math

import numpy as np

from .analytic import ClassicThacker
from .core import Domain, StructTriangMesh, jittered_mesh


def triang_average(p0, p1, p2, f, n):
    """TriangAverage<k, n> (include/PointOperations.h:20-44), vectorised over cells.

    p0, p1, p2: (T, 2) vertex arrays; f maps (x, y) arrays to a (k, ...) array.
    """
    h = 1.0 / n
    di = h * (p1 - p0)
    dj = h * (p2 - p0)
    dt = (1.0 / 3.0) * (di + dj)
    total = None
    pi = p0.copy()
    for i in range(n):
        pt = pi + dt
        for _ in range(n - i - 1):
            a = h * np.asarray(f(pt[:, 0], pt[:, 1]))
            q = pt + dt
            b = h * np.asarray(f(q[:, 0], q[:, 1]))
            total = (a if total is None else total + a) + b
            pt = pt + dj
        a = h * np.asarray(f(pt[:, 0], pt[:, 1]))
        total = a if total is None else total + a
        pi = pi + di
    return h * total


def _cell_vertices(xyz, topo):
    tp = topo.tables()["tp"]
    return xyz[tp[:, 0], :2], xyz[tp[:, 1], :2], xyz[tp[:, 2], :2]


def thacker(n=100, L=4.0, quad=100, t=0.0):
    """StructTriangMesh n x n on [-L/2, L/2]^2 (4 triangles per square), bowl
    bathymetry, Thacker state averaged with TriangAverage<3, quad> and
    w = <h> + b_cell (examples/Main.cpp:213-222)."""
    th = ClassicThacker()
    hh = L / n
    xyz, topo = StructTriangMesh(n, n, hh, -0.5 * L, -0.5 * L)
    xyz[:, 2] = th.bath(xyz[:, 0], xyz[:, 1])
    dom = Domain(xyz, topo)
    p0, p1, p2 = _cell_vertices(xyz, topo)
    avg = triang_average(p0, p1, p2, lambda x, y: np.stack([th.h(x, y, t), th.u(x, y, t), th.v(x, y, t)]), quad)
    bc = dom.centroids()[:, 2]
    prim = np.stack([avg[0] + bc, avg[1], avg[2]], axis=1)
    return dom, prim, th


def _tri_wet_volume(z, level):
    """Exact mean of max(0, level - b) over a triangle with linear b through
    the vertex values z (T, 3)."""
    zs = np.sort(z, axis=1)
    a, b, c = zs[:, 0], zs[:, 1], zs[:, 2]
    d1, d2, d3 = level - a, level - b, level - c
    out = np.zeros(z.shape[0])
    full = d3 >= 0
    out[full] = (d1[full] + d2[full] + d3[full]) / 3.0
    # one vertex wet: the wet region is a corner triangle
    m1 = (d1 > 0) & (d2 <= 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        out[m1] = d1[m1] ** 3 / (3.0 * (c[m1] - a[m1]) * (b[m1] - a[m1]))
        # two vertices wet: full mean minus the dry corner at c
        m2 = (d2 > 0) & (d3 < 0)
        mean = (d1[m2] + d2[m2] + d3[m2]) / 3.0
        out[m2] = mean + (-d3[m2]) ** 3 / (3.0 * (c[m2] - a[m2]) * (c[m2] - b[m2]))
    return out


def wall_taper(t, margin=0.05):
    """0 within `margin` of the walls of [0, 1], 1 beyond 2 * margin, C1 smoothstep between."""
    u = np.clip((np.minimum(t, 1.0 - t) - margin) / margin, 0.0, 1.0)
    return u * u * (3.0 - 2.0 * u)


def lake_at_rest(n=707, seed=1, level=1.0, jitter=0.25):
    """Flat surface w = level over 8 Gaussian bumps A ~ U(0.2, 1.2),
    s ~ U(0.02, 0.1); exactly two bumps emerge (A > level).  Partially wet
    cells hold the exact wet volume of the linear bathymetry.

    The bathymetry is tapered to exactly 0 near the walls: the reference's
    wall flux uses the cell-mean depth (src/SpaceDisc.cpp:67-69) while the
    pressure correction uses the edge depth (src/TimeDisc.cpp:35), so a lake
    at rest is steady only where the bottom is flat along the walls."""
    xyz, topo = jittered_mesh(n, jitter, seed)
    rng = np.random.default_rng(seed)
    A = np.concatenate([rng.uniform(1.05, 1.2, 2), rng.uniform(0.2, 0.95, 6)])
    s = rng.uniform(0.02, 0.1, 8)
    c = rng.uniform(0.15, 0.85, (8, 2))
    x, y = xyz[:, 0], xyz[:, 1]
    z = np.zeros_like(x)
    for k in range(8):
        z += A[k] * np.exp(-((x - c[k, 0]) ** 2 + (y - c[k, 1]) ** 2) / (2 * s[k] ** 2))
    z = z * wall_taper(x) * wall_taper(y)
    xyz[:, 2] = z
    dom = Domain(xyz, topo)
    tp = topo.tables()["tp"]
    zc = z[tp]
    bc = dom.centroids()[:, 2]
    zmax = zc.max(axis=1)
    prim = np.zeros((tp.shape[0], 3))
    wet = zmax < level
    prim[wet, 0] = level
    part = ~wet
    hbar = _tri_wet_volume(zc[part], level)
    prim[part, 0] = bc[part] + hbar
    return dom, prim


def dam_break(n=2828, seed=3, modes=64, amp=0.1, dry_frac=0.05, jitter=0.25):
    """Circular dam break (h = 1 for r < 0.2 around (0.5, 0.5), h = 0.1
    elsewhere) over z = amp * S / max|S|, S a sum of `modes` random Fourier
    modes with wave numbers in {-8..8}; circular dry pockets (h = 0) are added
    until dry_frac of the cells are dry; u = v = 0."""
    xyz, topo = jittered_mesh(n, jitter, seed)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, modes)
    m = rng.integers(-8, 9, modes)
    k = rng.integers(-8, 9, modes)
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    x, y = xyz[:, 0], xyz[:, 1]
    S = np.zeros_like(x)
    for q in range(modes):
        S += a[q] * np.cos(2.0 * np.pi * (m[q] * x + k[q] * y) + ph[q])
    xyz[:, 2] = amp * S / np.abs(S).max()
    dom = Domain(xyz, topo)
    cen = dom.centroids()
    bc = cen[:, 2]
    r = np.hypot(cen[:, 0] - 0.5, cen[:, 1] - 0.5)
    h = np.where(r < 0.2, 1.0, 0.1)
    T = h.shape[0]
    dry = np.zeros(T, bool)
    target = int(dry_frac * T)
    while dry.sum() < target:
        pc = rng.uniform(0.05, 0.95, 2)
        rad = rng.uniform(0.01, 0.04)
        if math.hypot(pc[0] - 0.5, pc[1] - 0.5) < 0.25 + rad:
            continue
        dry |= np.hypot(cen[:, 0] - pc[0], cen[:, 1] - pc[1]) < rad
    h[dry] = 0.0
    prim = np.stack([bc + h, np.zeros(T), np.zeros(T)], axis=1)
    return dom, prim


def ensemble_sizes(n_inst=256, total=20_000_000, tmin=10_000, tmax=500_000, seed=11):
    """Member sizes (triangles): log-uniform on [tmin, tmax], then rescaled
    (and clipped to the same range) so they sum to about `total`; each is
    rounded to a jittered grid 2 n^2."""
    rng = np.random.default_rng(seed)
    t = np.exp(rng.uniform(np.log(tmin), np.log(tmax), n_inst))
    for _ in range(50):
        t = np.clip(t * (total / t.sum()), tmin, tmax)
    n = np.maximum(2, np.round(np.sqrt(t / 2.0))).astype(int)
    return n


def ensemble_member(n, seed):
    """One ensemble member on [0, 1]^2: 1-3 Gaussian bumps of the bottom
    (amplitude U(0.05, 0.25), width U(0.05, 0.2)), a Gaussian surface hump
    of amplitude U(0.01, 0.2) and width U(0.05, 0.2) on w = 1, u = v = 0."""
    xyz, topo = jittered_mesh(int(n), 0.25, seed)
    rng = np.random.default_rng(seed + 1_000_003)
    x, y = xyz[:, 0], xyz[:, 1]
    z = np.zeros_like(x)
    for _ in range(int(rng.integers(1, 4))):
        cx, cy = rng.uniform(0.1, 0.9, 2)
        A, s = rng.uniform(0.05, 0.25), rng.uniform(0.05, 0.2)
        z += A * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    xyz[:, 2] = z
    dom = Domain(xyz, topo)
    cen = dom.centroids()
    A, s = rng.uniform(0.01, 0.2), rng.uniform(0.05, 0.2)
    hx, hy = rng.uniform(0.3, 0.7, 2)
    w = 1.0 + A * np.exp(-((cen[:, 0] - hx) ** 2 + (cen[:, 1] - hy) ** 2) / (2 * s * s))
    prim = np.stack([w, np.zeros_like(w), np.zeros_like(w)], axis=1)
    return dom, prim


def ensemble(n_inst=256, total=20_000_000, seed=11, sizes=None):
    """Seeded ensemble: list of (Domain, prim) of log-uniform sizes, smallest
    first (small members reach t_end first, so the running ones stay a
    contiguous block of cells)."""
    ns = np.sort(ensemble_sizes(n_inst, total, seed=seed)) if sizes is None else sizes
    return [ensemble_member(n, seed * 7919 + k) for k, n in enumerate(ns)]


def smooth_field(x, y, rng, sigma, modes=24, kmax=4):
    """Smooth random field: a sum of `modes` Fourier modes with wave numbers
    up to kmax, scaled to standard deviation sigma over the given points."""
    f = np.zeros_like(x)
    for _ in range(modes):
        m, n = rng.integers(-kmax, kmax + 1, 2)
        f += rng.standard_normal() * np.cos(2 * np.pi * (m * x + n * y) + rng.uniform(0, 2 * np.pi))
    sd = f.std()
    return f * (sigma / sd) if sd > 0 else f


def adjoint_inversion(n=354, seed=5, n_gauges=64, noise=0.02, depth=1.0):
    """Bathymetry inversion: a seeded Delaunay mesh (T = 2 n^2; n = 354 gives
    250,632 cells), deep water over Gaussian bumps (every cell wet above its
    highest vertex), a surface hump released at rest, `n_gauges` seeded gauge
    cells, and a "true" bathymetry z + N(0, noise)-smoothed noise whose run
    gives the gauge targets.  Returns (domain, initial state, true node
    bathymetry, gauge cell ids)."""
    xyz, topo = jittered_mesh(n, 0.25, seed)
    rng = np.random.default_rng(seed)
    x, y = xyz[:, 0], xyz[:, 1]
    z = np.zeros_like(x)
    for _ in range(6):
        cx, cy = rng.uniform(0.15, 0.85, 2)
        z += rng.uniform(0.1, 0.4) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * rng.uniform(0.05, 0.15) ** 2))
    z = z * wall_taper(x) * wall_taper(y)
    xyz[:, 2] = z
    z_true = z + smooth_field(x, y, rng, noise) * wall_taper(x) * wall_taper(y)
    dom = Domain(xyz, topo)
    c = dom.centroids()
    # the hump over a smooth seeded swell (amplitude 2e-3): no region of the
    # surface is exactly flat, where the limiter's decisions would make the
    # step non-differentiable
    w = depth + 0.1 * np.exp(-((c[:, 0] - 0.35) ** 2 + (c[:, 1] - 0.6) ** 2) / (2 * 0.08 ** 2))
    w = w + smooth_field(c[:, 0], c[:, 1], rng, 2e-3)
    prim = np.stack([w, np.zeros_like(w), np.zeros_like(w)], axis=1)
    gauges = np.sort(rng.choice(prim.shape[0], n_gauges, replace=False))
    return dom, prim, z_true, gauges


def adjoint_inversion_batch(members=32, n=354, seed=5, n_gauges=64, noise=0.02, depth=1.0):
    """A batch of bathymetry inversions on the mesh of adjoint_inversion:
    member k has its own "true" bathymetry (z + N(0, noise)-smoothed noise,
    seeded by (seed, k)) and its own seeded gauge cells; the initial state is
    shared.  Returns dict(member=(domain, initial state), z=node bathymetry,
    z_true=[V] * members, gauges=[local cell ids] * members)."""
    dom, prim, z_true0, gauges0 = adjoint_inversion(n=n, seed=seed, n_gauges=n_gauges, noise=noise, depth=depth)
    x, y = dom.geometry[:, 0], dom.geometry[:, 1]
    taper = wall_taper(x) * wall_taper(y)
    z = dom.geometry[:, 2].copy()
    zt, gs = [z_true0], [gauges0]
    for k in range(1, members):
        rng = np.random.default_rng([seed, k])
        zt.append(z + smooth_field(x, y, rng, noise) * taper)
        gs.append(np.sort(rng.choice(prim.shape[0], n_gauges, replace=False)))
    return dict(member=(dom, prim), z=z, z_true=zt, gauges=gs)
