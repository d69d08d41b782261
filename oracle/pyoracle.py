"""ctypes bindings for the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker (or the timed CPU baseline).

`liboracle.so` is the CPU restatement of the reference path (see the headers
of octree_cpu.cpp, linear_octree.cpp, direct.cpp for the reference file:line
each function follows). `_ref/libref.so`, when present, is the reference's own
CPU octree compiled from its sources by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class OcParams(C.Structure):
    _fields_ = [
        ("theta", C.c_float),
        ("softening", C.c_float),
        ("min_softening", C.c_float),
        ("min_distance2", C.c_float),
        ("min_theta", C.c_float),
        ("max_acceleration", C.c_float),
        ("cutoff2", C.c_float),
        ("short_range_scale", C.c_float),
        ("criterion", C.c_int),
        ("deterministic", C.c_int),
    ]


def make_params(theta=0.5, softening=0.01, min_softening=1e-4, min_distance2=1e-12,
                min_theta=0.05, max_acceleration=0.0, cutoff2=0.0, short_range_scale=0.0,
                criterion=0, deterministic=0) -> OcParams:
    return OcParams(theta, softening, min_softening, min_distance2, min_theta,
                    max_acceleration, cutoff2, short_range_scale, criterion, deterministic)


def build() -> None:
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oc_octree_forces.argtypes = [C.c_int, _f32p, C.POINTER(OcParams), _f32p, C.c_int]
        L.oc_octree_forces.restype = C.c_int
        L.oc_octree_forces_subset.argtypes = [C.c_int, _f32p, C.POINTER(OcParams), C.c_int,
                                              _i32p, _f32p, C.c_int]
        L.oc_octree_forces_subset.restype = C.c_int
        L.oc_octree_export.argtypes = [C.c_int, _f32p, C.c_int, _f32p, _i32p, _i32p, _i32p]
        L.oc_octree_export.restype = C.c_int
        L.oc_euler_update.argtypes = [C.c_int, _f32p, _f32p, C.c_float, C.c_float,
                                      C.POINTER(OcParams), _f32p, C.c_int]
        L.oc_kdk_step.argtypes = L.oc_euler_update.argtypes
        L.oc_pair_accel.argtypes = [_f32p, _f32p, C.c_float, C.POINTER(OcParams), _f32p]
        L.oc_resolve_policy.argtypes = [C.POINTER(OcParams), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.lo_encode.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        L.lo_encode.restype = C.c_uint64
        L.lo_aabb.argtypes = [C.c_int, _f64p, _f32p, _f32p]
        L.lo_keys.argtypes = [C.c_int, _f64p, C.c_int, _u64p, C.c_void_p]
        L.lo_sort.argtypes = [C.c_int, _u64p, _u64p, _u32p]
        L.lo_leaf_depth.argtypes = [C.c_longlong, C.c_int]
        L.lo_leaf_depth.restype = C.c_int
        L.lo_reference_tree.argtypes = [C.c_int, _u64p, _u32p, _f64p, _f32p, C.c_int, C.c_int,
                                        _i32p, _u64p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                        _f64p, _f64p]
        L.lo_reference_tree.restype = C.c_int
        L.lo_hip_tree.argtypes = [C.c_int, _u64p, C.c_int, C.c_int, C.c_int, _i32p, _u64p,
                                  _i32p, _i32p, _i32p, _i32p]
        L.lo_hip_tree.restype = C.c_int
        L.dx_direct_fp64.argtypes = [C.c_int, _f64p, _f64p, C.c_int, _i32p, C.c_double,
                                     C.c_double, _f64p, C.c_int]
        L.dx_force_error.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.dx_direct_tpm_fp64.argtypes = [C.c_int, _f64p, _f64p, C.c_int, _i32p, C.c_double,
                                         C.c_double, C.c_double, _f64p, C.c_int]
        L.dx_direct_tpm_fp64_box.argtypes = [C.c_int, _f64p, _f64p, C.c_int, _i32p, C.c_double,
                                             C.c_double, C.c_double, C.c_double, _f64p, C.c_int]
        L.dx_energy.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_double, _f64p,
                                C.c_int]
        L.dx_moments.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        _i32p_ = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.ad_steps.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _i32p_, _u64p,
                               C.POINTER(C.c_uint64), C.c_double, C.c_int, C.c_int, C.c_double,
                               C.c_double, C.c_int, C.c_double, C.c_double, C.c_int]
        L.ad_steps.restype = None
        L.ad_choose_level.argtypes = [_f64p, _f64p, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int]
        L.ad_choose_level.restype = C.c_int
        L.dx_multipole_accel.argtypes = [_f64p, _f64p, C.c_double, C.c_int, _f64p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own CPU octree (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_octree_forces.argtypes = [C.c_int, _f32p, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_int, C.c_float, _f32p]
        R.ref_octree_forces.restype = C.c_int
        R.ref_octree_forces_subset.argtypes = [C.c_int, _f32p, C.c_float, C.c_float, C.c_float,
                                               C.c_float, C.c_float, C.c_int, C.c_int, _i32p,
                                               _f32p, C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]
        R.ref_octree_forces_subset.restype = C.c_int
        _ref = R
    return _ref


# ----------------------------------------------------------------------------
# octree_cpu restatement
# ----------------------------------------------------------------------------
def hot(pos, mass) -> np.ndarray:
    """(n,4) float32 hot records x, y, z, m (PhyParticleHotData.cpp:15-27)."""
    pos = np.asarray(pos)
    out = np.empty((pos.shape[0], 4), np.float32)
    out[:, :3] = pos.astype(np.float32)
    out[:, 3] = np.asarray(mass, np.float32)
    return np.ascontiguousarray(out)


def octree_forces(xyzm: np.ndarray, params: OcParams | None = None, nthreads: int = 0):
    params = params or make_params()
    n = xyzm.shape[0]
    acc = np.zeros((n, 3), np.float32)
    nodes = lib().oc_octree_forces(n, np.ascontiguousarray(xyzm, np.float32).reshape(-1),
                                   C.byref(params), acc.reshape(-1), nthreads)
    return acc, nodes


def octree_forces_subset(xyzm, targets, params=None, nthreads=0):
    params = params or make_params()
    t = np.ascontiguousarray(targets, np.int32)
    acc = np.zeros((t.size, 3), np.float32)
    lib().oc_octree_forces_subset(xyzm.shape[0], np.ascontiguousarray(xyzm, np.float32).reshape(-1),
                                  C.byref(params), t.size, t, acc.reshape(-1), nthreads)
    return acc


def ref_octree_forces(xyzm, params=None):
    """Reference octree forces (unclamped) from oracle/_ref, or None."""
    R = ref_lib()
    if R is None:
        return None
    p = params or make_params()
    n = xyzm.shape[0]
    acc = np.zeros((n, 3), np.float32)
    R.ref_octree_forces(n, np.ascontiguousarray(xyzm, np.float32).reshape(-1), p.theta,
                        p.softening, p.min_softening, p.min_distance2, p.min_theta, p.criterion,
                        p.cutoff2, acc.reshape(-1))
    return acc


def euler_update(xyzm, vel, dt, max_speed, params=None, nthreads=0):
    params = params or make_params()
    acc = np.zeros((xyzm.shape[0], 3), np.float32)
    lib().oc_euler_update(xyzm.shape[0], xyzm.reshape(-1), vel.reshape(-1), dt, max_speed,
                          C.byref(params), acc.reshape(-1), nthreads)
    return acc


def kdk_step(xyzm, vel, acc, dt, max_speed, params=None, nthreads=0):
    params = params or make_params()
    lib().oc_kdk_step(xyzm.shape[0], xyzm.reshape(-1), vel.reshape(-1), dt, max_speed,
                      C.byref(params), acc.reshape(-1), nthreads)


def pair_accel(self3, src3, mass, params=None):
    params = params or make_params()
    out = np.zeros(3, np.float32)
    lib().oc_pair_accel(np.asarray(self3, np.float32), np.asarray(src3, np.float32), mass,
                        C.byref(params), out)
    return out


def resolve_policy(params):
    t, s, d = C.c_float(), C.c_float(), C.c_float()
    lib().oc_resolve_policy(C.byref(params), C.byref(t), C.byref(s), C.byref(d))
    return t.value, s.value, d.value


def octree_export(xyzm):
    n = xyzm.shape[0]
    cap = 2 * n + 16
    node8 = np.zeros((cap, 8), np.float32)
    child8 = np.zeros((cap, 8), np.int32)
    mfc = np.zeros((cap, 3), np.int32)
    leaf = np.zeros(max(n, 1), np.int32)
    nn = lib().oc_octree_export(n, np.ascontiguousarray(xyzm, np.float32).reshape(-1), cap,
                                node8.reshape(-1), child8.reshape(-1), mfc.reshape(-1), leaf)
    assert nn >= 0
    return dict(node=node8[:nn], child=child8[:nn], mask=mfc[:nn, 0], first=mfc[:nn, 1],
                count=mfc[:nn, 2], leaf_idx=leaf)


# ----------------------------------------------------------------------------
# linear octree (topology oracle)
# ----------------------------------------------------------------------------
def encode(qx, qy, qz, key_mode=0) -> int:
    return int(lib().lo_encode(qx, qy, qz, key_mode))


def aabb(pos):
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().lo_aabb(p.size // 3, p, lo, hi)
    return lo, hi


def keys(pos, key_mode=0, with_q=False):
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    n = p.size // 3
    k = np.zeros(n, np.uint64)
    q = np.zeros((n, 3), np.uint32) if with_q else None
    lib().lo_keys(n, p, key_mode, k, q.ctypes.data if with_q else None)
    return (k, q) if with_q else k


def stable_sort(k):
    k = np.ascontiguousarray(k, np.uint64)
    sk = np.zeros_like(k)
    order = np.zeros(k.size, np.uint32)
    lib().lo_sort(k.size, k, sk, order)
    return sk, order


def leaf_depth(n, leaf_capacity=256) -> int:
    return int(lib().lo_leaf_depth(n, leaf_capacity))


def reference_tree(sorted_keys, order, pos, mass, depth):
    n = sorted_keys.size
    cap = 2 * n + 64 + 22 * 8
    arrs = dict(level=np.zeros(cap, np.int32), prefix=np.zeros(cap, np.uint64),
                start=np.zeros(cap, np.int32), count=np.zeros(cap, np.int32),
                child=np.zeros((cap, 8), np.int32), parent=np.zeros(cap, np.int32),
                next=np.zeros(cap, np.int32), msum=np.zeros(cap, np.float64),
                mxyz=np.zeros((cap, 3), np.float64))
    nn = lib().lo_reference_tree(
        n, np.ascontiguousarray(sorted_keys, np.uint64), np.ascontiguousarray(order, np.uint32),
        np.ascontiguousarray(pos, np.float64).reshape(-1), np.ascontiguousarray(mass, np.float32),
        depth, cap, arrs["level"], arrs["prefix"], arrs["start"], arrs["count"],
        arrs["child"].reshape(-1), arrs["parent"], arrs["next"], arrs["msum"],
        arrs["mxyz"].reshape(-1))
    assert nn >= 0, nn
    return {k: v[:nn] for k, v in arrs.items()}


def hip_tree(sorted_keys, depth, bucket):
    n = sorted_keys.size
    cap = 4 * n + 64
    while True:
        arrs = dict(level=np.zeros(cap, np.int32), prefix=np.zeros(cap, np.uint64),
                    start=np.zeros(cap, np.int32), count=np.zeros(cap, np.int32),
                    first_child=np.zeros(cap, np.int32), nchild=np.zeros(cap, np.int32))
        nn = lib().lo_hip_tree(n, np.ascontiguousarray(sorted_keys, np.uint64), depth, bucket,
                               cap, arrs["level"], arrs["prefix"], arrs["start"], arrs["count"],
                               arrs["first_child"], arrs["nchild"])
        if nn >= 0:
            return {k: v[:nn] for k, v in arrs.items()}
        cap = -nn


# ----------------------------------------------------------------------------
# FP64 direct sum and metrics
# ----------------------------------------------------------------------------
def direct_fp64(pos, mass, targets, eps, min_d2=0.0, nthreads=0):
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    m = np.ascontiguousarray(mass, np.float64)
    t = np.ascontiguousarray(targets, np.int32)
    acc = np.zeros((t.size, 3), np.float64)
    lib().dx_direct_fp64(m.size, p, m, t.size, t, eps, min_d2, acc.reshape(-1), nthreads)
    return acc


def direct_tpm_fp64(pos, mass, targets, eps, rs, rc, nthreads=0, box=0.0):
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    m = np.ascontiguousarray(mass, np.float64)
    t = np.ascontiguousarray(targets, np.int32)
    acc = np.zeros((t.size, 3), np.float64)
    lib().dx_direct_tpm_fp64_box(m.size, p, m, t.size, t, eps, rs, rc, box, acc.reshape(-1),
                                 nthreads)
    return acc


def force_error(approx, ref) -> dict:
    a = np.ascontiguousarray(approx, np.float64).reshape(-1)
    r = np.ascontiguousarray(ref, np.float64).reshape(-1)
    out = np.zeros(5, np.float64)
    lib().dx_force_error(a.size // 3, a, r, out)
    return dict(l2=out[0], linf=out[1], p99=out[2], median=out[3], max=out[4])


def energy(pos, vel, mass, eps, min_d2=0.0, nthreads=0):
    out = np.zeros(8, np.float64)
    lib().dx_energy(len(mass), np.ascontiguousarray(pos, np.float64).reshape(-1),
                    np.ascontiguousarray(vel, np.float64).reshape(-1),
                    np.ascontiguousarray(mass, np.float64), eps, min_d2, out, nthreads)
    return dict(ke=out[0], pe=out[1], p=out[2:5].copy(), l=out[5:8].copy())


def moments(pos, mass):
    out = np.zeros(10, np.float64)
    lib().dx_moments(len(mass), np.ascontiguousarray(pos, np.float64).reshape(-1),
                     np.ascontiguousarray(mass, np.float64), out)
    return out


def multipole_accel(m10, x3, eps, order=2):
    out = np.zeros(3, np.float64)
    lib().dx_multipole_accel(np.ascontiguousarray(m10, np.float64),
                             np.ascontiguousarray(x3, np.float64), eps, order, out)
    return out


def ref_octree_forces_subset(xyzm, targets, params=None, nthreads=0):
    """Reference octree (oracle/_ref) on listed targets, OpenMP over targets.
    Returns (acc, build_seconds, walk_seconds) or None when _ref is absent."""
    R = ref_lib()
    if R is None:
        return None
    p = params or make_params()
    t = np.ascontiguousarray(targets, np.int32)
    acc = np.zeros((t.size, 3), np.float32)
    tb, tw = C.c_double(), C.c_double()
    R.ref_octree_forces_subset(xyzm.shape[0], np.ascontiguousarray(xyzm, np.float32).reshape(-1),
                               p.theta, p.softening, p.min_softening, p.min_distance2,
                               p.min_theta, p.criterion, t.size, t, acc.reshape(-1), nthreads,
                               C.byref(tb), C.byref(tw))
    return acc, tb.value, tw.value


class AdaptiveState:
    """Dyadic block time-step schedule of oracle/adaptive.cpp (FP64 state,
    FP64 direct-sum forces)."""

    def __init__(self, pos, vel, mass, dt, max_level, eta, soft, eps, velocity_criterion=True,
                 max_speed=0.0):
        self.pos = np.ascontiguousarray(pos, np.float64).copy()
        self.vel = np.ascontiguousarray(vel, np.float64).copy()
        self.mass = np.ascontiguousarray(mass, np.float64)
        n = len(self.mass)
        self.acc = np.zeros((n, 3), np.float64)
        self.level = np.zeros(n, np.int32)
        self.last = np.zeros(n, np.uint64)
        self.tick = C.c_uint64(0)
        self.args = (dt, max_level, eta, soft, 1 if velocity_criterion else 0, max_speed, eps)
        self.primed = False

    def steps(self, nsteps):
        dt, L, eta, soft, vc, vmax, eps = self.args
        lib().ad_steps(len(self.mass), self.pos.reshape(-1), self.vel.reshape(-1), self.mass,
                       self.acc.reshape(-1), self.level, self.last, C.byref(self.tick), dt,
                       nsteps, L, eta, soft, vc, vmax, eps, 0 if self.primed else 1)
        self.primed = True


def choose_level(a3, v3, eta, soft, dt, max_level, velocity_criterion=True):
    return int(lib().ad_choose_level(np.ascontiguousarray(a3, np.float64),
                                     np.ascontiguousarray(v3, np.float64), eta, soft, dt,
                                     max_level, 1 if velocity_criterion else 0))
