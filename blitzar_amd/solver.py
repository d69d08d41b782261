"""Python host mirror of the octree_hip solver, over the C ABI.

`OctreeHip` follows the reference's ParticleSystem solver surface
(engine/physics/core/particle/PhyParticleSystem.hpp:109-718): the setters that
configureSystem calls (engine/batch/runner/BatRunner.cpp:48-97), `update(dt)`
returning False on failure (the run is then marked faulted,
BatRunner.cpp:189-193), and the force pass. It holds no physics itself: every
call goes to libbh_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import Params, Timings, lib

SOLVER_NAMES = ("octree_hip", "octree-hip", "octree_gpu_hip")
CRITERIA = {"com": 0, "center_of_mass": 0, "bounds": 1, "extent": 2, "relative": 3}
INTEGRATORS = {"leapfrog": 0, "kdk": 0, "euler": 1}


class BhHipError(RuntimeError):
    """Raised when the C ABI reports an error."""


def _check(rc: int, h=None, what: str = "") -> None:
    if rc == _abi.OK:
        return
    msg = ""
    if h is not None:
        buf = C.create_string_buffer(1024)
        lib().bh_hip_get_last_error(h, buf, 1024)
        msg = buf.value.decode(errors="replace")
    raise BhHipError(f"{what}: {_abi.RESULT_NAMES.get(rc, rc)} {msg}".strip())


def normalize_solver(name: str) -> str:
    """octree_hip aliases (CfgNormalize.cpp:46-67 style: lower-case, '-' -> '_')."""
    n = name.strip().lower().replace("-", "_")
    if n in ("octree_hip", "octree_gpu_hip"):
        return "octree_hip"
    return n


class OctreeHip:
    """One octree_hip solver instance (owns its device buffers and stream)."""

    def __init__(self, n: int, params: Params | None = None, **overrides):
        self.n = int(n)
        p = params if params is not None else _abi.default_params()
        for k, v in overrides.items():
            if k == "criterion" and isinstance(v, str):
                v = CRITERIA[v]
            if k == "integrator" and isinstance(v, str):
                v = INTEGRATORS[v]
            setattr(p, k, v)
        self.params = p
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().bh_hip_create(C.byref(p), self.n, C.byref(h), err, 1024)
        if rc != _abi.OK:
            raise BhHipError(f"bh_hip_create: {_abi.RESULT_NAMES.get(rc, rc)} "
                             f"{err.value.decode(errors='replace')}")
        self._h = h

    # -- lifecycle -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().bh_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration (ParticleSystem setters) --------------------------------
    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            if k == "criterion" and isinstance(v, str):
                v = CRITERIA[v]
            if k == "integrator" and isinstance(v, str):
                v = INTEGRATORS[v]
            setattr(self.params, k, v)
        _check(lib().bh_hip_set_params(self._h, C.byref(self.params)), self._h, "set_params")

    def resolved_params(self) -> Params:
        """The configuration the handle runs with (after resolveForceLawPolicy)."""
        out = Params()
        _check(lib().bh_hip_get_params(self._h, C.byref(out)), self._h, "get_params")
        return out

    def set_octree_theta(self, theta: float) -> None:
        self.set_params(theta=theta)

    def set_octree_softening(self, eps: float) -> None:
        self.set_params(softening=eps)

    def set_octree_opening_criterion(self, criterion: str) -> None:
        self.set_params(criterion=criterion)

    def set_integrator_mode(self, integrator: str) -> None:
        self.set_params(integrator=integrator)

    def set_physics_stability_constants(self, max_acceleration, min_softening, min_distance2,
                                        min_theta) -> None:
        self.set_params(max_acceleration=max_acceleration, min_softening=min_softening,
                        min_distance2=min_distance2, min_theta=min_theta)

    # -- state -------------------------------------------------------------------
    def upload(self, pos, vel=None, mass=None) -> None:
        pos = np.ascontiguousarray(pos, np.float64).reshape(self.n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(self.n, 3)
        if mass is None:
            mass = np.full(self.n, 1.0 / self.n, np.float32)
        mass = np.ascontiguousarray(mass, np.float32).reshape(self.n)
        _check(lib().bh_hip_upload(self._h, pos, vel, mass), self._h, "upload")

    def download(self):
        pos = np.zeros((self.n, 3), np.float64)
        vel = np.zeros((self.n, 3), np.float64)
        _check(lib().bh_hip_download(self._h, pos, vel), self._h, "download")
        return pos, vel

    # -- force pass and stepping -----------------------------------------------
    def compute_accelerations(self, to_host: bool = True):
        acc = np.zeros((self.n, 3), np.float32) if to_host else None
        _check(lib().bh_hip_compute_accelerations(self._h, acc), self._h,
               "compute_accelerations")
        return acc

    def step(self, dt: float, nsteps: int = 1) -> None:
        _check(lib().bh_hip_step(self._h, float(dt), int(nsteps)), self._h, "step")

    def update(self, dt: float) -> bool:
        """ParticleSystem::update(dt): one step; False on failure."""
        if dt <= 0.0:
            return False
        return lib().bh_hip_step(self._h, float(dt), 1) == _abi.OK

    def last_error(self) -> str:
        buf = C.create_string_buffer(1024)
        lib().bh_hip_get_last_error(self._h, buf, 1024)
        return buf.value.decode(errors="replace")

    # -- tree and diagnostics ---------------------------------------------------
    def build_tree(self) -> None:
        _check(lib().bh_hip_build_tree(self._h), self._h, "build_tree")

    def tree_info(self):
        nn = C.c_uint64()
        d = C.c_int32()
        m = C.c_int32()
        _check(lib().bh_hip_tree_info(self._h, C.byref(nn), C.byref(d), C.byref(m)), self._h,
               "tree_info")
        return int(nn.value), int(d.value), int(m.value)

    def dump_tree(self) -> dict:
        nn, depth, maxl = self.tree_info()
        out = dict(
            sorted_keys=np.zeros(self.n, np.uint64), sorted_ids=np.zeros(self.n, np.uint32),
            prefix=np.zeros(nn, np.uint64), level=np.zeros(nn, np.int32),
            start=np.zeros(nn, np.uint32), count=np.zeros(nn, np.uint32),
            first_child=np.zeros(nn, np.int32), nchild=np.zeros(nn, np.uint32),
            moments=np.zeros((nn, 16), np.float32), geom=np.zeros(12, np.float64))
        _check(lib().bh_hip_dump_tree(self._h, out["sorted_keys"], out["sorted_ids"],
                                      out["prefix"], out["level"], out["start"], out["count"],
                                      out["first_child"], out["nchild"],
                                      out["moments"].reshape(-1), out["geom"], nn),
               self._h, "dump_tree")
        out["leaf_depth"] = depth
        out["max_level"] = maxl
        out["lattice"] = out["moments"].view(np.int32)[:, 12:15].copy()
        return out

    def direct_forces_fp64(self, targets) -> np.ndarray:
        t = np.ascontiguousarray(targets, np.uint32)
        out = np.zeros((t.size, 3), np.float64)
        _check(lib().bh_hip_direct_forces_fp64(self._h, t, t.size, out.reshape(-1)), self._h,
               "direct_forces_fp64")
        return out

    def diagnostics(self, sample_limit: int = 0) -> dict:
        out = np.zeros(8, np.float64)
        _check(lib().bh_hip_diagnostics(self._h, sample_limit, out), self._h, "diagnostics")
        return dict(ke=out[0], pe=out[1], p=out[2:5].copy(), l=out[5:8].copy())

    def set_timing(self, enabled: bool = True) -> None:
        _check(lib().bh_hip_set_timing(self._h, 1 if enabled else 0), self._h, "set_timing")

    def timings(self) -> dict:
        t = Timings()
        _check(lib().bh_hip_get_timings(self._h, C.byref(t)), self._h, "get_timings")
        return {k: getattr(t, k) for k, _ in Timings._fields_}


    def pm_accelerations(self) -> np.ndarray:
        """Long-range (mesh) part of the periodic TreePM force (caller order)."""
        out = np.zeros((self.n, 3), np.float32)
        _check(lib().bh_hip_pm_accelerations(self._h, out.reshape(-1)), self._h,
               "pm_accelerations")
        return out

    def adaptive_levels(self) -> np.ndarray:
        """Block time-step level of every particle (caller order)."""
        out = np.zeros(self.n, np.int32)
        _check(lib().bh_hip_adaptive_levels(self._h, out), self._h, "adaptive_levels")
        return out

    # -- learned far-field correction -----------------------------------------
    def lff_configure(self, theta: float = 0.9, band: float = 0.5, list_cap: int = 192,
                      enabled: bool = False, min_count: int = 32) -> None:
        """Set the corrected walk's opening angle, near band and the smallest
        corrected node; enable it."""
        cfg = _abi.LffConfig(theta, band, list_cap, 1 if enabled else 0, min_count)
        _check(lib().bh_hip_lff_configure(self._h, C.byref(cfg)), self._h, "lff_configure")
        self._lff_cfg = cfg

    def lff_enable(self, enabled: bool = True) -> None:
        cfg = getattr(self, "_lff_cfg", None) or _abi.LffConfig(0.9, 0.5, 192, 0, 32)
        cfg.enabled = 1 if enabled else 0
        _check(lib().bh_hip_lff_configure(self._h, C.byref(cfg)), self._h, "lff_configure")
        self._lff_cfg = cfg

    def lff_collect(self, group_stride: int = 1, max_tuples: int = 1 << 22) -> int:
        """Walk at the correction's theta and append training tuples; returns the count."""
        total = C.c_uint64()
        _check(lib().bh_hip_lff_collect(self._h, group_stride, max_tuples, C.byref(total)),
               self._h, "lff_collect")
        return int(total.value)

    def lff_train(self, epochs: int = 20, batch: int = 4096, lr: float = 2e-3,
                  seed: int = 1) -> np.ndarray:
        """Train on the device tuples; returns the loss per epoch."""
        loss = np.zeros(epochs, np.float32)
        _check(lib().bh_hip_lff_train(self._h, epochs, batch, lr, seed, loss), self._h,
               "lff_train")
        return loss

    def lff_params(self) -> np.ndarray:
        p = np.zeros(_abi.LFF_PARAMS, np.float32)
        _check(lib().bh_hip_lff_get_params(self._h, p), self._h, "lff_get_params")
        return p

    def lff_set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != _abi.LFF_PARAMS:
            raise ValueError(f"expected {_abi.LFF_PARAMS} parameters, got {p.size}")
        _check(lib().bh_hip_lff_set_params(self._h, p), self._h, "lff_set_params")

    def lff_gradient(self, X, Y, params):
        """(gradient, loss) of mean |MLP(X) - Y|^2 through the FP32 training kernels."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros(_abi.LFF_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib().bh_hip_lff_gradient(self._h, X, Y, X.shape[0], p, g, C.byref(loss)),
               self._h, "lff_gradient")
        return g, float(loss.value)

    def lff_forward(self, X) -> np.ndarray:
        """MLP outputs for feature rows through the BF16 MFMA inference kernel."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], _abi.LFF_OUTPUTS), np.float32)
        _check(lib().bh_hip_lff_forward(self._h, X, X.shape[0], Y), self._h, "lff_forward")
        return Y

    def lff_tuples(self, cap: int | None = None):
        """Copy the training tuples (features, normalised residuals) to the host."""
        n = C.c_uint64()
        _check(lib().bh_hip_lff_get_tuples(self._h, None, None, 0, C.byref(n)), self._h,
               "lff_get_tuples")
        m = int(n.value) if cap is None else min(int(n.value), cap)
        X = np.zeros((m, _abi.LFF_INPUTS), np.float32)
        Y = np.zeros((m, _abi.LFF_OUTPUTS), np.float32)
        if m:
            _check(lib().bh_hip_lff_get_tuples(self._h, X, Y, m, C.byref(n)), self._h,
                   "lff_get_tuples")
        return X, Y


def sort_pairs(keys, vals, key_bits: int = 64):
    """Device stable radix sort of (uint64 key, uint32 value) pairs."""
    k = np.ascontiguousarray(keys, np.uint64)
    v = np.ascontiguousarray(vals, np.uint32)
    ko = np.zeros_like(k)
    vo = np.zeros_like(v)
    err = C.create_string_buffer(512)
    rc = lib().bh_hip_sort_pairs(k, v, k.size, key_bits, ko, vo, err, 512)
    if rc != _abi.OK:
        raise BhHipError(f"bh_hip_sort_pairs: {_abi.RESULT_NAMES.get(rc, rc)} "
                         f"{err.value.decode(errors='replace')}")
    return ko, vo


# -- initial conditions (host, deterministic) -----------------------------------
def plummer(n: int, seed: int = 42, total_mass: float = 1.0, scale_radius: float = 1.0):
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    m = np.zeros(n, np.float32)
    lib().bh_ic_plummer(n, seed, total_mass, scale_radius, pos.reshape(-1), vel.reshape(-1), m)
    return pos, vel, m


def hernquist(n: int, seed: int = 11, stream: int = 0, total_mass: float = 1.0,
              scale_radius: float = 1.0, r_max: float = 200.0):
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    m = np.zeros(n, np.float32)
    lib().bh_ic_hernquist(n, seed, stream, total_mass, scale_radius, r_max, pos.reshape(-1),
                          vel.reshape(-1), m)
    return pos, vel, m


def hernquist_merger(n_per_halo: int, seed: int = 11, separation: float = 20.0,
                     pericentre: float = 2.0):
    n = 2 * n_per_halo
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    m = np.zeros(n, np.float32)
    lib().bh_ic_hernquist_merger(n_per_halo, seed, separation, pericentre, pos.reshape(-1),
                                 vel.reshape(-1), m)
    return pos, vel, m


def uniform_cube(n: int, seed: int = 1, half_extent: float = 1.0, total_mass: float = 1.0):
    pos = np.zeros((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    m = np.zeros(n, np.float32)
    lib().bh_ic_uniform_cube(n, seed, half_extent, total_mass, pos.reshape(-1), vel.reshape(-1),
                             m)
    return pos, vel, m


def zeldovich_lattice(ng, seed=3, rms_cells=0.3, box=1.0):
    """ng^3 lattice displaced by a first-order Zel'dovich field: Gaussian
    random density with P(k) ~ k^-2, displacement psi_k = i k delta_k / k^2,
    scaled to an RMS displacement of rms_cells lattice cells per axis."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((ng, ng, ng)).astype(np.float32)
    dk = np.fft.rfftn(white)
    del white
    k = np.fft.fftfreq(ng, d=1.0 / ng).astype(np.float32)
    kz = np.arange(ng // 2 + 1, dtype=np.float32)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    dk *= 1.0 / np.sqrt(k2)  # amplitude sqrt(P) ~ 1/k
    dk[0, 0, 0] = 0.0
    g = (np.arange(ng, dtype=np.float64) + 0.5) / ng * box
    pos = np.empty((ng ** 3, 3), np.float64)
    disp = []
    for a, K in enumerate((k[:, None, None], k[None, :, None], kz[None, None, :])):
        comp = np.fft.irfftn(1j * K * dk / k2, s=(ng, ng, ng), axes=(0, 1, 2))
        disp.append(comp.astype(np.float32).reshape(-1))
        del comp
    scale = rms_cells * (box / ng) / np.sqrt(np.mean([np.mean(d * d) for d in disp]))
    grid = np.meshgrid(g, g, g, indexing="ij")
    for a in range(3):
        pos[:, a] = np.mod(grid[a].reshape(-1) + scale * disp[a], box)
    return pos
