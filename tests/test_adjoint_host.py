"""Host-side pieces of the adjoint interface: the per-cell bathymetry map
(node heights as the mean of the incident cells' values) and its transpose."""
import numpy as np

import swe_fvm_amd as swe
from swe_fvm_amd import adjoint as AD


def test_cell_bathymetry_map_and_transpose():
    xyz, topo = swe.jittered_mesh(9, 0.2, 4)
    tp = topo.tables()["tp"]
    V, T = xyz.shape[0], tp.shape[0]
    M = AD.CellBathymetry(tp, V)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(T)
    z = M.to_nodes(b)
    # by definition, node by node
    for n in rng.choice(V, 20, replace=False):
        inc = np.nonzero((tp == n).any(1))[0]
        assert abs(z[n] - b[inc].mean()) <= 1e-14
    # constants are kept
    assert np.allclose(M.to_nodes(np.full(T, 2.5)), 2.5, rtol=0, atol=1e-15)
    # gradient() is the transpose: <M b, y> = <b, M^T y>
    y = rng.standard_normal(V)
    assert abs(np.dot(z, y) - np.dot(b, M.gradient(y))) <= 1e-12 * np.sum(np.abs(z * y))
    # dense check on a small mesh
    Md = np.stack([M.to_nodes(np.eye(T)[k]) for k in range(T)], 1)
    assert np.allclose(Md.T @ y, M.gradient(y), rtol=1e-14, atol=1e-15)
