"""Ensembles of independent meshes advanced together on one GPU.

The members are concatenated into one disjoint mesh; every kernel of the
step runs once per stage over all of them, each member with its own CFL step
size, and a member stops when it reaches t_end.  One step is captured into a
HIP graph and replayed.  Each member's result is bit-identical to advancing
it alone with SpaceDisc.run in CFL mode.
"""
import ctypes

import numpy as np

from . import _lib
from .core import (SCHEMES, Fluxer, SolverError, _autotune, _autotune_layout, _check, _layout, _ptr, _set_tuning,
                   _tuning)


class Ensemble:
    def __init__(self, fluxer, domains, states, cor=0.0, tau=0.0, options=0, device=0):
        if not isinstance(fluxer, Fluxer):
            raise SolverError(f"unsupported fluxer {fluxer!r}")
        L = _lib.load()
        n = len(domains)
        if n == 0 or len(states) != n:
            raise ValueError("need one state per domain")
        self.domains = list(domains)
        topo_arr = (ctypes.c_void_p * n)(*[d.topology._h.value for d in domains])
        self._geo = [np.ascontiguousarray(d.geometry, dtype=np.float64) for d in domains]
        xyz_arr = (ctypes.c_void_p * n)(*[g.ctypes.data for g in self._geo])
        h = ctypes.c_void_p()
        _check(L.swe_batch_create(n, ctypes.cast(topo_arr, ctypes.c_void_p), ctypes.cast(xyz_arr, ctypes.c_void_p),
                                  fluxer.code, fluxer.ws.code, float(cor), float(tau), int(options),
                                  int(device), ctypes.byref(h)))
        self._h = h
        self.n = n
        self.cell_off = np.empty(n + 1, np.int64)
        self.edge_off = np.empty(n + 1, np.int64)
        self.node_off = np.empty(n + 1, np.int64)
        _check(L.swe_batch_info(h, None, _ptr(self.cell_off), _ptr(self.edge_off), _ptr(self.node_off)))
        self.T = int(self.cell_off[-1])
        self.set_states(states)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().swe_destroy(self._h)
            self._h = None

    def set_states(self, states):
        prim = np.concatenate([np.asarray(s, dtype=np.float64).reshape(-1, 3) for s in states])
        if prim.shape[0] != self.T:
            raise ValueError("state sizes disagree with the meshes")
        prim = np.ascontiguousarray(prim)
        _check(_lib.load().swe_set_state(self._h, _ptr(prim)))

    def states(self):
        out = np.empty((self.T, 3), np.float64)
        _check(_lib.load().swe_get_state(self._h, _ptr(out)))
        return [out[self.cell_off[k]:self.cell_off[k + 1]] for k in range(self.n)]

    def reset(self, t0=0.0):
        _check(_lib.load().swe_batch_reset(self._h, float(t0)))

    def run(self, scheme="SSPRK2", cfl=0.15, t_end=1.0, max_steps=1 << 40, graph=True):
        """Steps until every member reached t_end (or max_steps); returns
        (steps taken, members still active)."""
        done, act = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.load().swe_batch_run(self._h, SCHEMES[scheme], float(cfl), float(t_end), int(max_steps),
                                         1 if graph else 0, ctypes.byref(done), ctypes.byref(act)))
        return done.value, act.value

    def times(self):
        t = np.empty(self.n, np.float64)
        st = np.empty(self.n, np.int64)
        _check(_lib.load().swe_batch_time(self._h, _ptr(t), _ptr(st)))
        return t, st

    def set_tuning(self, flux_items=2, update_items=1, flux_block=256, update_block=256):
        _set_tuning(self._h, flux_items, update_items, flux_block, update_block)

    def tuning(self):
        return _tuning(self._h)

    def autotune(self, reps=10):
        """Picks the fastest work items per thread on the current states
        (as SpaceDisc.autotune); a captured step graph is re-captured."""
        return _autotune(self._h, reps)

    def set_layout(self, block=0):
        """State layout of the batch (as SpaceDisc.set_layout)."""
        _check(_lib.load().swe_set_layout(self._h, int(block)))

    def layout(self):
        return _layout(self._h)

    def autotune_layout(self, reps=10):
        """Work items, workgroup size and state layout (SoA or AoSoA) picked
        on the current states (as SpaceDisc.autotune_layout)."""
        return _autotune_layout(self._h, reps)

    def synchronize(self):
        _check(_lib.load().swe_synchronize(self._h))
