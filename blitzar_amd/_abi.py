"""ctypes binding of the octree_hip C ABI (include/bh_hip.h, include/bh_ic.h).

The shared library is built in-tree (blitzar_amd/libbh_hip.so) by
`__graft_entry__.build()` / `make -C blitzar_amd/csrc`. There is no fallback:
if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libbh_hip.so")

OK, INVALID_ARGUMENT, NOT_READY, TIMEOUT, INTERNAL_ERROR = range(5)
RESULT_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "NOT_READY", 3: "TIMEOUT", 4: "INTERNAL_ERROR"}


class Params(C.Structure):
    _fields_ = [
        ("theta", C.c_float),
        ("softening", C.c_float),
        ("min_softening", C.c_float),
        ("min_distance2", C.c_float),
        ("min_theta", C.c_float),
        ("max_acceleration", C.c_float),
        ("max_speed", C.c_float),
        ("criterion", C.c_int32),
        ("leaf_capacity", C.c_int32),
        ("bucket_capacity", C.c_int32),
        ("multipole_order", C.c_int32),
        ("key_mode", C.c_int32),
        ("integrator", C.c_int32),
        ("treepm_split_scale", C.c_float),
        ("treepm_cutoff", C.c_float),
        ("waves_per_cu", C.c_int32),
        ("adaptive_max_level", C.c_int32),
        ("adaptive_eta", C.c_float),
        ("adaptive_flags", C.c_int32),
        ("periodic_box", C.c_float),
        ("pm_mesh", C.c_int32),
        ("far_alpha", C.c_float),
        ("relative_alpha", C.c_float),
        ("relative_mono", C.c_float),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("build_ms", C.c_double),
        ("sort_ms", C.c_double),
        ("walk_ms", C.c_double),
        ("integrate_ms", C.c_double),
        ("pp_interactions", C.c_uint64),
        ("pc_interactions", C.c_uint64),
        ("node_count", C.c_uint64),
        ("leaf_depth", C.c_int32),
        ("max_level", C.c_int32),
        ("walks", C.c_int32),
        ("groups", C.c_int32),
        ("particle_steps", C.c_uint64),
        ("far_interactions", C.c_uint64),
        ("far_evaluations", C.c_uint64),
        ("mono_interactions", C.c_uint64),
        ("keys_ms", C.c_double),
        ("permute_ms", C.c_double),
        ("tree_ms", C.c_double),
        ("moments_ms", C.c_double),
        ("lff_pairs", C.c_uint64),
        ("lff_tiles", C.c_uint64),
        ("lff_ms", C.c_double),
    ]


class LffConfig(C.Structure):
    _fields_ = [
        ("theta", C.c_float),
        ("band", C.c_float),
        ("list_cap", C.c_int32),
        ("enabled", C.c_int32),
        ("min_count", C.c_int32),
    ]


LFF_INPUTS = 32
LFF_HIDDEN = 64
LFF_OUTPUTS = 3
LFF_PARAMS = 6467


# Every exported symbol of include/bh_hip.h and include/bh_ic.h.
EXPORTS = [
    "bh_hip_default_params", "bh_hip_version", "bh_hip_create", "bh_hip_destroy",
    "bh_hip_set_params", "bh_hip_upload", "bh_hip_download", "bh_hip_compute_accelerations",
    "bh_hip_step", "bh_hip_build_tree", "bh_hip_tree_info", "bh_hip_dump_tree",
    "bh_hip_resolve_params", "bh_hip_get_params",
    "bh_hip_direct_forces_fp64", "bh_hip_diagnostics", "bh_hip_set_timing",
    "bh_hip_get_timings", "bh_hip_get_last_error", "bh_hip_sort_pairs",
    "bh_hip_sort_pairs_noalloc",
    "bh_hip_lff_configure", "bh_hip_lff_collect", "bh_hip_lff_train", "bh_hip_lff_get_params",
    "bh_hip_lff_set_params", "bh_hip_lff_gradient", "bh_hip_lff_forward",
    "bh_hip_lff_get_tuples", "bh_hip_adaptive_levels", "bh_hip_pm_accelerations",
    "bh_ic_plummer", "bh_ic_hernquist", "bh_ic_hernquist_merger", "bh_ic_uniform_cube",
]

_lib = None


def _p(dtype):
    return np.ctypeslib.ndpointer(dtype, flags="C_CONTIGUOUS")


def _opt(dtype):
    """Pointer argument that also accepts None (NULL)."""
    base = _p(dtype)

    class _Opt(base):  # type: ignore[misc, valid-type]
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return base.from_param(obj)

    return _Opt


def lib() -> C.CDLL:
    """Load libbh_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"octree_hip native library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    R = C.c_int
    L.bh_hip_default_params.argtypes = [C.POINTER(Params)]
    L.bh_hip_default_params.restype = None
    L.bh_hip_version.restype = C.c_char_p
    L.bh_hip_create.argtypes = [C.POINTER(Params), C.c_uint64, C.POINTER(H), C.c_char_p, C.c_size_t]
    L.bh_hip_create.restype = R
    L.bh_hip_destroy.argtypes = [H]
    L.bh_hip_destroy.restype = None
    L.bh_hip_resolve_params.argtypes = [C.POINTER(Params), C.POINTER(Params)]
    L.bh_hip_resolve_params.restype = None
    L.bh_hip_get_params.argtypes = [H, C.POINTER(Params)]
    L.bh_hip_get_params.restype = R
    L.bh_hip_set_params.argtypes = [H, C.POINTER(Params)]
    L.bh_hip_set_params.restype = R
    L.bh_hip_upload.argtypes = [H, _p(np.float64), _opt(np.float64), _p(np.float32)]
    L.bh_hip_upload.restype = R
    L.bh_hip_download.argtypes = [H, _opt(np.float64), _opt(np.float64)]
    L.bh_hip_download.restype = R
    L.bh_hip_compute_accelerations.argtypes = [H, _opt(np.float32)]
    L.bh_hip_compute_accelerations.restype = R
    L.bh_hip_step.argtypes = [H, C.c_double, C.c_int32]
    L.bh_hip_step.restype = R
    L.bh_hip_build_tree.argtypes = [H]
    L.bh_hip_build_tree.restype = R
    L.bh_hip_tree_info.argtypes = [H, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]
    L.bh_hip_tree_info.restype = R
    L.bh_hip_dump_tree.argtypes = [H, _opt(np.uint64), _opt(np.uint32), _opt(np.uint64),
                                   _opt(np.int32), _opt(np.uint32), _opt(np.uint32),
                                   _opt(np.int32), _opt(np.uint32), _opt(np.float32),
                                   _opt(np.float64), C.c_uint64]
    L.bh_hip_dump_tree.restype = R
    L.bh_hip_direct_forces_fp64.argtypes = [H, _p(np.uint32), C.c_uint32, _p(np.float64)]
    L.bh_hip_direct_forces_fp64.restype = R
    L.bh_hip_diagnostics.argtypes = [H, C.c_uint32, _p(np.float64)]
    L.bh_hip_diagnostics.restype = R
    L.bh_hip_set_timing.argtypes = [H, C.c_int32]
    L.bh_hip_set_timing.restype = R
    L.bh_hip_get_timings.argtypes = [H, C.POINTER(Timings)]
    L.bh_hip_get_timings.restype = R
    L.bh_hip_get_last_error.argtypes = [H, C.c_char_p, C.c_size_t]
    L.bh_hip_get_last_error.restype = R
    L.bh_hip_sort_pairs.argtypes = [_p(np.uint64), _p(np.uint32), C.c_uint64, C.c_int32,
                                    _p(np.uint64), _p(np.uint32), C.c_char_p, C.c_size_t]
    L.bh_hip_sort_pairs.restype = R
    # device pointers and a stream handle, passed as plain addresses
    L.bh_hip_sort_pairs_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p,
                                            C.c_size_t]
    L.bh_hip_sort_pairs_noalloc.restype = R
    L.bh_hip_lff_configure.argtypes = [H, C.POINTER(LffConfig)]
    L.bh_hip_lff_collect.argtypes = [H, C.c_int32, C.c_uint64, C.POINTER(C.c_uint64)]
    L.bh_hip_lff_train.argtypes = [H, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                   _opt(np.float32)]
    L.bh_hip_lff_get_params.argtypes = [H, _p(np.float32)]
    L.bh_hip_lff_set_params.argtypes = [H, _p(np.float32)]
    L.bh_hip_lff_gradient.argtypes = [H, _p(np.float32), _p(np.float32), C.c_uint32,
                                      _p(np.float32), _p(np.float32), C.POINTER(C.c_float)]
    L.bh_hip_lff_forward.argtypes = [H, _p(np.float32), C.c_uint32, _p(np.float32)]
    L.bh_hip_lff_get_tuples.argtypes = [H, _opt(np.float32), _opt(np.float32), C.c_uint64,
                                        C.POINTER(C.c_uint64)]
    L.bh_hip_adaptive_levels.argtypes = [H, _p(np.int32)]
    L.bh_hip_adaptive_levels.restype = R
    L.bh_hip_pm_accelerations.argtypes = [H, _p(np.float32)]
    L.bh_hip_pm_accelerations.restype = R
    for name in ("bh_hip_lff_configure", "bh_hip_lff_collect", "bh_hip_lff_train",
                 "bh_hip_lff_get_params", "bh_hip_lff_set_params", "bh_hip_lff_gradient",
                 "bh_hip_lff_forward", "bh_hip_lff_get_tuples"):
        getattr(L, name).restype = R
    L.bh_ic_plummer.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, _p(np.float64),
                                _p(np.float64), _p(np.float32)]
    L.bh_ic_plummer.restype = None
    L.bh_ic_hernquist.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                  C.c_double, _p(np.float64), _p(np.float64), _p(np.float32)]
    L.bh_ic_hernquist.restype = None
    L.bh_ic_hernquist_merger.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                         _p(np.float64), _p(np.float64), _p(np.float32)]
    L.bh_ic_hernquist_merger.restype = None
    L.bh_ic_uniform_cube.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                     _p(np.float64), _p(np.float64), _p(np.float32)]
    L.bh_ic_uniform_cube.restype = None
    _lib = L
    return L


def default_params() -> Params:
    p = Params()
    lib().bh_hip_default_params(C.byref(p))
    return p


def resolve_params(p: Params) -> Params:
    """The configuration a handle runs with for `p` (host only)."""
    out = Params()
    lib().bh_hip_resolve_params(C.byref(p), C.byref(out))
    return out
