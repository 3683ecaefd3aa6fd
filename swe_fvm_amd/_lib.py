"""ctypes binding of libswe_fvm_amd.so (the C ABI declared in include/swe_fvm.h).

The library is built in-tree by swe_fvm_amd.build (hipcc, gfx950).  There is
no fallback: if the shared object is missing, loading raises ImportError.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SWE_FVM_AMD_LIB may name another build of the same sources (tuning variants)
LIB_PATH = os.environ.get("SWE_FVM_AMD_LIB", os.path.join(HERE, "libswe_fvm_amd.so"))

_d = ctypes.c_double
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_p = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# (name, restype, argtypes) for every symbol of include/swe_fvm.h
SIGNATURES = [
    ("swe_last_error", ctypes.c_char_p, []),
    ("swe_topology_create", _int, [_i64, _i64, _p, _p, _i64, _p, _p, _p, _pp]),
    ("swe_topology_build", _int, [_i64, _i64, _p, _i64, _p, _pp]),
    ("swe_topology_set_periodic", _int, [_p, _i64, _p]),
    ("swe_topology_periodic", _int, [_p, _p]),
    ("swe_topology_read_msh", _int, [ctypes.c_char_p, _pp, _pi64, _pp]),
    ("swe_topology_num_nodes", _i64, [_p]),
    ("swe_topology_num_edges", _i64, [_p]),
    ("swe_topology_num_elements", _i64, [_p]),
    ("swe_topology_copy", _int, [_p, _p, _p, _p, _p, _p]),
    ("swe_topology_destroy", None, [_p]),
    ("swe_free", None, [_p]),
    ("swe_mesh_struct", _int, [_i64, _i64, _d, _d, _d, _pp, _pi64, _pp]),
    ("swe_mesh_jittered", _int, [_i64, _d, _u64, _d, _d, _d, _pp, _pi64, _pp]),
    ("swe_create", _int, [_p, _p, _int, _int, _d, _d, _int, _int, _pp]),
    ("swe_destroy", None, [_p]),
    ("swe_set_state", _int, [_p, _p]),
    ("swe_get_state", _int, [_p, _p]),
    ("swe_compute_interface_values", _int, [_p]),
    ("swe_compute_fluxes", _int, [_p, _p]),
    ("swe_get_edge_field", _int, [_p, _p]),
    ("swe_get_src_field", _int, [_p, _p]),
    ("swe_get_fluxes", _int, [_p, _p]),
    ("swe_get_max_wp", _int, [_p, _p]),
    ("swe_get_cell_class", _int, [_p, _p]),
    ("swe_rhs", _int, [_p, _i64, _d, _p]),
    ("swe_run", _int, [_p, _int, _d, _d, _i64]),
    ("swe_get_dts", _int, [_p, _p, _i64]),
    ("swe_synchronize", _int, [_p]),
    ("swe_cfl_dt", _int, [_p, _p]),
    ("swe_min_len_to_wavespeed", _int, [_p, _p]),
    ("swe_num_cells", _i64, [_p]),
    ("swe_num_edges", _i64, [_p]),
    ("swe_num_nodes", _i64, [_p]),
    ("swe_device_state", _int, [_p, _pp, _pp, _pp]),
    ("swe_time_run", _int, [_p, _int, _d, _d, _i64, _p, _p]),
    ("swe_batch_create", _int, [_i64, _p, _p, _int, _int, _d, _d, _int, _int, _pp]),
    ("swe_batch_reset", _int, [_p, _d]),
    ("swe_batch_run", _int, [_p, _int, _d, _d, _i64, _int, _pi64, _pi64]),
    ("swe_batch_info", _int, [_p, _pi64, _p, _p, _p]),
    ("swe_batch_time", _int, [_p, _p, _p]),
    ("swe_integrals", _int, [_p, _p, _p]),
    ("swe_integrals_noalloc", _int, [_p, _p, _p]),
    ("swe_set_fused", _int, [_p, _int, _p]),
    ("swe_reconstruct", _int, [_p, _i64, _int, _p]),
    ("swe_edge_flux", _int, [_p, _int, _int, _i64, _i64, _i64, _p, _p]),
    ("swe_get_muscl", _int, [_p, _p, _p, _p]),
    ("swe_state_stats", _int, [_p, _p]),
    ("swe_adjoint_run", _int, [_p, _int, _d, _p, _i64, _i64, _p, _p, _p, _p]),
    ("swe_adjoint_timing", _int, [_p, _p]),
    ("swe_tangent_run", _int, [_p, _int, _d, _p, _i64, _p, _p, _p, _p]),
    ("swe_set_bathymetry", _int, [_p, _p]),
    ("swe_set_tuning", _int, [_p, _int, _int]),
    ("swe_set_tuning_blocks", _int, [_p, _int, _int]),
    ("swe_get_tuning_blocks", _int, [_p, _p, _p]),
    ("swe_get_tuning", _int, [_p, _p, _p]),
    ("swe_autotune", _int, [_p, _int, _p, _p, _p]),
    ("swe_set_layout", _int, [_p, _int]),
    ("swe_get_layout", _int, [_p, _p]),
    ("swe_autotune_layout", _int, [_p, _int, _p, _p]),
]

_lib = None


def load():
    """Loads the in-tree shared library; raises ImportError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m swe_fvm_amd.build` "
            "(hipcc, gfx950); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
