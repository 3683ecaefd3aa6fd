"""Shared test helpers: golden fixtures, workload construction, comparisons."""
import gzip
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def read_gz_text(name):
    with gzip.open(os.path.join(GOLDEN, name), "rt") as f:
        return f.read()


def parse_msh_text(text):
    """Minimal Gmsh 4.1 / 4.2 ASCII parser (test-side, independent of the library)."""
    lines = text.split("\n")
    i = 0
    nodes, bnd, tris = {}, [], []
    while i < len(lines):
        s = lines[i].strip()
        if s == "$Nodes":
            nb = int(lines[i + 1].split()[0])
            i += 2
            for _ in range(nb):
                ed, et, par, cnt = map(int, lines[i].split())
                i += 1
                tags = [int(lines[i + k]) for k in range(cnt)]
                i += cnt
                for k in range(cnt):
                    nodes[tags[k]] = list(map(float, lines[i + k].split()[:3]))
                i += cnt
            continue
        if s == "$Elements":
            nb = int(lines[i + 1].split()[0])
            i += 2
            for _ in range(nb):
                ed, et, typ, cnt = map(int, lines[i].split())
                i += 1
                for k in range(cnt):
                    v = list(map(int, lines[i + k].split()))
                    if typ == 1:
                        bnd.append(v[1:3])
                    elif typ == 2:
                        tris.append(v[1:4])
                i += cnt
            continue
        i += 1
    keys = sorted(nodes)
    rank = {k: r for r, k in enumerate(keys)}
    xyz = np.array([nodes[k] for k in keys])
    tri = np.array([[rank[a] for a in t] for t in tris], np.int64)
    b = np.array([[rank[a] for a in l] for l in bnd], np.int64)
    return xyz, tri, b


class Golden:
    """The reference's data fixtures (see tests/golden/make_golden.py)."""

    def __init__(self):
        self._cache = {}

    def bowl(self):
        if "bowl" not in self._cache:
            self._cache["bowl"] = parse_msh_text(read_gz_text("bowl.msh.gz"))
        return self._cache["bowl"]

    def bowl_path(self):
        """bowl.msh decompressed into a temporary file (for the library reader)."""
        return self.msh_path("bowl.msh.gz")

    def msh_path(self, name):
        """A gzipped mesh fixture decompressed into a temporary file."""
        key = "path:" + name
        if key not in self._cache:
            fd, path = tempfile.mkstemp(suffix=".msh")
            with os.fdopen(fd, "w") as f:
                f.write(read_gz_text(name))
            self._cache[key] = path
        return self._cache[key]

    def basic(self):
        """notebooks/basic.msh: (xyz, triangles, boundary lines) by the test-side parser."""
        if "basic" not in self._cache:
            self._cache["basic"] = parse_msh_text(read_gz_text("basic.msh.gz"))
        return self._cache["basic"]

    def topology(self):
        if "topo" not in self._cache:
            lines = read_gz_text("topology.dat.gz").split("\n")
            E = int(lines[0])
            ed = np.array([list(map(int, lines[1 + k].split())) for k in range(E)], np.int64)
            T = int(lines[1 + E])
            td = np.array([list(map(int, lines[2 + E + k].split())) for k in range(T)], np.int64)
            self._cache["topo"] = dict(ep=ed[:, :2], et=ed[:, 2:], tp=td[:, :3], te=td[:, 3:6], tt=td[:, 6:])
        return self._cache["topo"]

    def table(self, name):
        if name not in self._cache:
            self._cache[name] = np.loadtxt(gzip.open(os.path.join(GOLDEN, name), "rt"))
        return self._cache[name]


def gauss_wave_state(xyz, tp):
    """Initial state of testGaussWave (examples/Main.cpp:182-186):
    w = 1 + exp(-5 |T - (4, 4)|^2), u = v = 0 at the centroids T."""
    c = 1.0 / 3.0
    cx = (xyz[tp[:, 0], 0] * c + xyz[tp[:, 1], 0] * c) + xyz[tp[:, 2], 0] * c
    cy = (xyz[tp[:, 0], 1] * c + xyz[tp[:, 1], 1] * c) + xyz[tp[:, 2], 1] * c
    r = (cx - 4.0) ** 2 + (cy - 4.0) ** 2
    return np.stack([1.0 + np.exp(-5.0 * r), np.zeros_like(r), np.zeros_like(r)], axis=1)


def cell_area(xyz, tp):
    a, b, c = xyz[tp[:, 0]], xyz[tp[:, 1]], xyz[tp[:, 2]]
    return 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))


def cell_bath(xyz, tp):
    c = 1.0 / 3.0
    return (xyz[tp[:, 0], 2] * c + xyz[tp[:, 1], 2] * c) + xyz[tp[:, 2], 2] * c


def mass(xyz, tp, prim):
    return float(np.sum(cell_area(xyz, tp) * (prim[:, 0] - cell_bath(xyz, tp))))


def max_rel_diff(a, b, floor=1e-300):
    """max |a - b| / max(|b|, 1) per column."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), 1.0)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def wet_dry_mesh(n, seed, dry_frac=0.08, jitter=0.25):
    """Small jittered mesh with bumps (some emerging), a dam and dry pockets:
    exercises dry, partially wet (all PartWet2 regimes) and full-wet cells."""
    import swe_fvm_amd as swe
    xyz, topo = swe.jittered_mesh(n, jitter, seed)
    rng = np.random.default_rng(seed)
    x, y = xyz[:, 0], xyz[:, 1]
    z = 0.05 * np.sin(3 * x + 1.0) * np.cos(2 * y)
    for _ in range(5):
        cx, cy = rng.uniform(0.15, 0.85, 2)
        A, s = rng.uniform(0.3, 1.3), rng.uniform(0.05, 0.15)
        z += A * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    xyz[:, 2] = z
    tp = topo.tables()["tp"]
    bc = cell_bath(xyz, tp)
    c = 1.0 / 3.0
    cx = (xyz[tp[:, 0], 0] * c + xyz[tp[:, 1], 0] * c) + xyz[tp[:, 2], 0] * c
    cy = (xyz[tp[:, 0], 1] * c + xyz[tp[:, 1], 1] * c) + xyz[tp[:, 2], 1] * c
    level = np.where(np.hypot(cx - 0.3, cy - 0.6) < 0.2, 1.1, 0.8)
    h = np.maximum(0.0, level - bc)
    T = tp.shape[0]
    pockets = rng.random(T) < dry_frac
    h[pockets] = 0.0
    u = rng.normal(0, 0.1, T) * (h > 0)
    v = rng.normal(0, 0.1, T) * (h > 0)
    prim = np.stack([bc + h, u, v], axis=1)
    return xyz, topo, prim


def swe_fvm_module():
    """The pybind module SWE_FVM under the name the reference's scripts
    import it by (one name per process: pybind registers its types once)."""
    import importlib
    import sys
    d = os.path.join(os.path.dirname(HERE), "swe_fvm_amd")
    if d not in sys.path:
        sys.path.insert(0, d)
    return importlib.import_module("SWE_FVM")
