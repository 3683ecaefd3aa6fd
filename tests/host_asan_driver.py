"""Drives the host-only mesh, topology and tile-table code built with
AddressSanitizer and UBSan (tests/test_host_sanitized.py builds the library
and runs this script in a process of its own with the sanitizer runtime
preloaded).  Only numpy and ctypes are imported: the HIP library is not
loaded here.

usage: host_asan_driver.py LIB BOWL_MSH BASIC_MSH TMPDIR
Prints one line per case and "ok" at the end; any sanitizer report aborts
the process with a non-zero status."""
import ctypes
import os
import sys

import numpy as np

lib_path, bowl, basic, tmp = sys.argv[1:5]
L = ctypes.CDLL(lib_path)
P = ctypes.c_void_p
I64 = ctypes.c_int64
L.swe_last_error.restype = ctypes.c_char_p
L.swe_mesh_jittered.argtypes = [I64, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, P, P, P]
L.swe_mesh_struct.argtypes = [I64, I64, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P, P]
L.swe_topology_read_msh.argtypes = [ctypes.c_char_p, P, P, P]
L.swe_topology_build.argtypes = [I64, I64, P, I64, P, P]
L.swe_topology_create.argtypes = [I64, I64, P, P, I64, P, P, P, P]
L.swe_topology_copy.argtypes = [P, P, P, P, P, P]
L.swe_topology_set_periodic.argtypes = [P, I64, P]
L.swe_topology_num_nodes.argtypes = [P]
L.swe_topology_num_nodes.restype = I64
L.swe_topology_num_edges.argtypes = [P]
L.swe_topology_num_edges.restype = I64
L.swe_topology_num_elements.argtypes = [P]
L.swe_topology_num_elements.restype = I64
L.swe_topology_destroy.argtypes = [P]
L.swe_free.argtypes = [P]
L.tiles_check.restype = ctypes.c_int
L.tiles_check.argtypes = [I64] + [P] * 7 + [I64, P, ctypes.c_char_p, ctypes.c_int]


def ptr(a):
    return a.ctypes.data_as(P)


def tables(topo):
    V, E, T = L.swe_topology_num_nodes(topo), L.swe_topology_num_edges(topo), L.swe_topology_num_elements(topo)
    ep, et = np.zeros((E, 2), np.int64), np.zeros((E, 2), np.int64)
    tp, te, tt = (np.zeros((T, 3), np.int64) for _ in range(3))
    assert L.swe_topology_copy(topo, ptr(ep), ptr(et), ptr(tp), ptr(te), ptr(tt)) == 0
    return V, ep, et, tp, te, tt


def check_tiles(name, topo, refused_ok=False):
    V, ep, et, tp, te, tt = tables(topo)
    T = tp.shape[0]
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    et0, et1 = i32(et[:, 0]), i32(et[:, 1])
    fl = np.zeros(T, np.uint32)
    for k in range(3):
        fl |= (et0[te[:, k]] == np.arange(T)).astype(np.uint32) << k
    stats = np.zeros(5, np.int64)
    msg = ctypes.create_string_buffer(256)
    rc = L.tiles_check(T, ptr(i32(tt.T.reshape(-1))), ptr(i32(te.T.reshape(-1))), ptr(fl), ptr(et0), ptr(et1),
                       ptr(i32(ep[:, 0])), ptr(i32(ep[:, 1])), et0.shape[0], ptr(stats), msg, 256)
    text = msg.value.decode()
    assert rc == 0 or (refused_ok and "refused" in text), (name, text)
    print(f"{name}: T={T} tiles rc={rc} {text}", flush=True)
    return tp, ep, et


def mesh(fn, *args):
    xyz, nv, topo = P(), I64(), P()
    rc = fn(*args, ctypes.byref(xyz), ctypes.byref(nv), ctypes.byref(topo))
    assert rc == 0, L.swe_last_error()
    L.swe_free(xyz)
    return topo


# generated, structured and file meshes through the tile builder and checker
for n, seed in ((1, 1), (7, 2), (60, 3)):
    t = mesh(L.swe_mesh_jittered, n, 0.25, seed, 0.0, 0.0, 1.0)
    check_tiles(f"jittered n={n}", t)
    L.swe_topology_destroy(t)
t = mesh(L.swe_mesh_struct, 10, 7, 0.1, 0.0, 0.0)
tp, ep, et = check_tiles("struct 10x7", t)
L.swe_topology_destroy(t)
for path in (bowl, basic):
    xyz, nv, t = P(), I64(), P()
    assert L.swe_topology_read_msh(path.encode(), ctypes.byref(xyz), ctypes.byref(nv), ctypes.byref(t)) == 0
    L.swe_free(xyz)
    check_tiles(os.path.basename(path), t, refused_ok=True)
    L.swe_topology_destroy(t)

# a random cell order: the builder refuses its tiles
t = mesh(L.swe_mesh_jittered, 40, 0.25, 5, 0.0, 0.0, 1.0)
V, ep, et, tp, te, tt = tables(t)
L.swe_topology_destroy(t)
perm = np.random.default_rng(0).permutation(tp.shape[0])
tri = np.ascontiguousarray(tp[perm])
bnd = np.ascontiguousarray(ep[et[:, 1] < 0])
t2 = P()
assert L.swe_topology_build(V, tri.shape[0], ptr(tri), bnd.shape[0], ptr(bnd), ctypes.byref(t2)) == 0
check_tiles("random order", t2, refused_ok=True)

# error paths: rejected inputs return an error code and leave no bad access
bad = P()
ep_bad = ep.copy()
ep_bad[0, 0] = V + 5  # node out of range
rc = L.swe_topology_create(V, ep.shape[0], ptr(ep_bad), ptr(np.ascontiguousarray(et)), tp.shape[0], ptr(tp),
                           ptr(np.ascontiguousarray(te)), ptr(np.ascontiguousarray(tt)), ctypes.byref(bad))
assert rc != 0
print("create with a node out of range:", L.swe_last_error().decode(), flush=True)
interior = int(np.nonzero(et[:, 1] >= 0)[0][0])
boundary = np.nonzero(et[:, 1] < 0)[0]
for pairs in ([[interior, int(boundary[0])]], [[int(boundary[0]), ep.shape[0] + 3]],
              [[int(boundary[0]), int(boundary[0])]]):
    a = np.ascontiguousarray(pairs, dtype=np.int64)
    assert L.swe_topology_set_periodic(t2, a.shape[0], ptr(a)) != 0
    print("set_periodic", pairs, ":", L.swe_last_error().decode(), flush=True)
L.swe_topology_destroy(t2)
missing = os.path.join(tmp, "missing.msh")
assert L.swe_topology_read_msh(missing.encode(), ctypes.byref(P()), ctypes.byref(I64()), ctypes.byref(P())) != 0
text = open(bowl).read()
bad_files = {f"truncated to {cut} chars": text[:cut]
             for cut in (len(text) // 3, (2 * len(text)) // 3, len(text) - 40, text.index("$EndNodes") + 4)}
lines = text.split("\n")
h = lines.index("$Nodes") + 1  # numEntityBlocks numNodes minNodeTag maxNodeTag
nb, nn, mn, mx = lines[h].split()
for name, header in (("negative node count", f"{nb} -5 {mn} {mx}"),
                     ("node count above the blocks'", f"{nb} {int(nn) + 1} {mn} {mx}"),
                     ("node count below the blocks'", f"{nb} {int(nn) - 1} {mn} {mx}")):
    bad_files[name] = "\n".join(lines[:h] + [header] + lines[h + 1:])
for name, body in bad_files.items():
    path = os.path.join(tmp, "bad.msh")
    with open(path, "w") as f:
        f.write(body)
    xyz, nv, t = P(), I64(), P()
    rc = L.swe_topology_read_msh(path.encode(), ctypes.byref(xyz), ctypes.byref(nv), ctypes.byref(t))
    print(f"{name}: rc={rc}", L.swe_last_error().decode(), flush=True)
    assert rc != 0, name

# seeded mutations of the small mesh file (a line deleted or duplicated, a
# token replaced by another integer, a cut): each is read without a
# sanitizer report; the accepted ones also go through the tile checker
small = open(basic).read().split("\n")
rng = np.random.default_rng(7)
accepted = 0
for it in range(300):
    lines = list(small)
    kind = it % 4
    i = int(rng.integers(0, len(lines)))
    if kind == 0:
        del lines[i]
    elif kind == 1:
        lines.insert(i, lines[i])
    elif kind == 2:
        toks = lines[i].split()
        if toks:
            toks[int(rng.integers(0, len(toks)))] = str(int(rng.choice([-1, 0, 1, 2, 3, 7, 99, 2 ** 31, -2 ** 40])))
            lines[i] = " ".join(toks)
    body = "\n".join(lines)
    if kind == 3:
        body = body[:int(rng.integers(0, len(body)))]
    path = os.path.join(tmp, "fuzz.msh")
    with open(path, "w") as f:
        f.write(body)
    xyz, nv, t = P(), I64(), P()
    rc = L.swe_topology_read_msh(path.encode(), ctypes.byref(xyz), ctypes.byref(nv), ctypes.byref(t))
    if rc == 0:
        accepted += 1
        L.swe_free(xyz)
        check_tiles(f"mutation {it}", t, refused_ok=True) if L.swe_topology_num_elements(t) > 0 else None
        L.swe_topology_destroy(t)
print(f"mutations: {accepted} of 300 accepted", flush=True)
print("ok", flush=True)
