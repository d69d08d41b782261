"""Benchmark of the octree_hip hot path on one GPU.

Metric: particle-steps/s at N = 8,388,608 with the median relative force error
<= 1e-3 against an FP64 direct sum on 4,096 sampled particles.

Workload (default): two Hernquist halos of 4,194,304 particles each (mass 1,
scale radius 1, isotropic distribution function, seed 11), centres 20 apart
on a parabolic approach (pericentre 2); eps = 0.01, FP32 walk with the
second-order softened expansion (monopole + traceless quadrupole + eps^2
trace term), FP64 leapfrog kick-drift-kick with a global step dt = 2^-10.
One step = drift/kick + tree rebuild (AABB, keys, radix sort, tree, moments)
+ walk + kick, all on the device with the state resident in HBM. Synthetic
initial conditions (generated on the host, uploaded once).

The opening test is the accuracy knob of the metric. The headline runs the
relative criterion: a node is accepted when its estimated truncation error,
M bmax^3 / (d^2 + eps^2)^(5/2), is below alpha = 0.0125 times the group's
smallest |a| of the previous force evaluation, with the opening angle bounded
by theta = 2; accepted nodes whose quadrupole term, M bmax^2 / (d^2 +
eps^2)^2, is below 0.2 alpha |a| enter as point masses. On this state that
holds the median error near 6.6e-4, a third below the bar. The same run also
times the plain COM walk at theta = 0.5 under "theta_0.5", the opening angle
the workload is usually quoted at. Both force errors are measured in the run.
`--criterion com --theta 0.75 --far-alpha 0.75` runs the geometric criterion
with the group far-field expansion instead. The run also times the block
time-step, periodic TreePM and learned-correction workloads (theta 0.9,
trained on a second realisation, seed 12); `--block-steps 0`, `--treepm 0`
and `--no-lff` skip them.

Prints one JSON line (see the keys below). `--gpus N` with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 8M, median |dF|/F ≤1e-3 vs FP64 direct sum"
PEAK_FP32_TFLOPS = 157.3
FLOP_PP = 20.0   # particle-particle interaction
FLOP_PC = 50.0   # particle-cell interaction, monopole + quadrupole
FLOP_PC_MONO = 20.0
FLOP_FAR_NODE = 180.0  # node -> group local expansion (monopole to 2nd, quadrupole to 1st order)
FLOP_FAR_EVAL = 48.0   # evaluation of the group expansion at one target


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_state(workload: str, n: int):
    import blitzar_amd as B
    if workload == "merger":
        return B.hernquist_merger(n // 2, seed=11, separation=20.0, pericentre=2.0)
    if workload == "plummer":
        return B.plummer(n, seed=7)
    raise SystemExit(f"unknown workload {workload}")


def cpu_baseline(pos, m, theta, eps, n_targets, seed=123):
    """octree_cpu on the host cores: the reference's own CPU octree compiled
    from its sources (oracle/_ref) when present, else the oracle restatement.
    Bounded sample: the full serial tree build over all particles plus the
    OpenMP walk on `n_targets` random targets, extrapolated to one full force
    pass over all particles."""
    from oracle import pyoracle as O
    n = pos.shape[0]
    x = O.hot(pos, m)
    rng = np.random.default_rng(seed)
    t = np.sort(rng.choice(n, n_targets, replace=False)).astype(np.int32)
    params = O.make_params(theta=theta, softening=eps)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = O.ref_octree_forces_subset(x, t, params, nthreads=threads)
    single = None
    if ref is not None:
        _, t_build, t_walk = ref
        kind = "reference"
        # one thread: the same build and the walk on the first 1,024 targets
        one = O.ref_octree_forces_subset(x, t[:1024], params, nthreads=1)
        single = n / (one[1] + one[2] * (n / 1024))
    else:
        t0 = time.perf_counter()
        O.octree_forces_subset(x, t[:1], params, nthreads=threads)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.octree_forces_subset(x, t, params, nthreads=threads)
        t_walk = time.perf_counter() - t0 - t_build
        kind = "port"
    t_full = t_build + t_walk * (n / n_targets)
    return {
        "value": n / t_full,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": kind,
        "single_thread_value": single,
        "cpu_model": cpu_model(),
        "sample": (f"octree_cpu (theta {theta}, eps {eps}, monopole, FP32) on the same state: "
                   f"serial tree build over all {n} particles ({t_build:.2f} s) + walk on "
                   f"{n_targets} random targets ({t_walk:.2f} s, {threads} threads), "
                   f"extrapolated to one full force pass = one Euler step; single thread: the "
                   f"same build and 1,024 of the targets, extrapolated the same way"),
    }


PMC_SUMMARY = os.path.join(ROOT, "profiles", "walk_build_merger8M_relative0.0125_mono0.2_pmc.md")


def walk_traffic(args, n):
    """Bytes per walk launch between L2 and the memory fabric, from the
    committed rocprofv3 summary of this workload (raw read / write request
    counters; 128 B per read request and 64 B per 64-byte write request, as a
    1 GiB copy in the same run measures them), or None for other workloads."""
    if not (args.workload == "merger" and n == 8388608 and args.order == 2 and args.bucket == 16
            and args.criterion == "relative" and args.theta == 2.0
            and args.relative_alpha == 0.0125 and args.relative_mono == 0.2
            and args.far_alpha == 0.0):
        return None
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("walk_hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def hbm_fractions(tm, n, walks, args):
    """Algorithmic bytes over device time, against 8 TB/s, for the
    bandwidth-bound phases of one step (HIP events on the library's stream
    around each phase, averaged over the timed force evaluations):
      keys       36 B/particle (FP64 positions in, 63-bit key + index out);
      sort      192 B/particle (8 passes x 24 B, the fixed 8-bit-digit
                convention, whatever the pass count);
      permute    60 B/particle (positions and masses: the rest of the store
                 follows on the second stream beside the moments, outside
                 this timer);
      tree       40 B/particle + 128 B/node (levels, moments, leaf packing
                and walk groups together);
      integrate 108 B/particle per kick + drift (the closing half kick of a
                call, 60 B, is in the time but not in the bytes)."""
    nodes = float(tm["node_count"])
    ph = {
        "keys": (36.0 * n, tm["keys_ms"] / walks),
        "sort": (192.0 * n, tm["sort_ms"] / walks),
        "permute": (60.0 * n, tm["permute_ms"] / walks),
        "tree": (40.0 * n + 128.0 * nodes, (tm["tree_ms"] + tm["moments_ms"]) / walks),
        "integrate": (108.0 * n, tm["integrate_ms"] / max(1, args.steps)),
    }
    out = {}
    for k, (b, ms) in ph.items():
        rate = b / (ms * 1e-3) if ms > 0 else 0.0
        out[k] = {"bytes": b, "ms": ms, "tb_per_s": rate / 1e12, "frac": rate / HBM_PEAK}
    return out


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB, three times: each dispatch
    reads and writes exactly mib MiB, which fixes the scale of the profiler's
    byte counters for the other kernels of the same run."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    log(f"[bench] copy calibration: 3 copies of {mib} MiB")


def workload_name(args, n):
    crit = (f"relative{args.relative_alpha}_mono{args.relative_mono}" if args.criterion == "relative"
            else args.criterion)
    return (f"{'hernquist_merger' if args.workload == 'merger' else 'plummer'}_{n}_{crit}"
            f"_theta{args.theta}_{'quadrupole' if args.order >= 2 else 'monopole'}"
            f"{f'_far{args.far_alpha}' if args.far_alpha > 0 else ''}")


def force_error(acc, ref, samples):
    acc = np.asarray(acc, np.float64)
    r = np.linalg.norm(acc - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    return {"median": float(np.median(r)), "p99": float(np.percentile(r, 99)),
            "max": float(r.max()),
            "l2": float(np.linalg.norm(acc - ref) / max(np.linalg.norm(ref), 1e-300)),
            "samples": int(samples), "sample_seed": 2024}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def periodic_treepm(steps, ng=256):
    """Periodic TreePM box: ng^3 Zel'dovich lattice (seed 3, rms 0.3 cells),
    L = 1, rocFFT mesh ng^3, r_s = 1.25 cells, short-range cut at 4.5 r_s,
    theta 0.6, eps 0.05 cells, KDK dt 1e-3, zero initial velocities."""
    import blitzar_amd as B
    t0 = time.perf_counter()
    pos = B.zeldovich_lattice(ng)
    n = pos.shape[0]
    t_ic = time.perf_counter() - t0
    h = 1.0 / ng
    with B.OctreeHip(n, theta=0.6, softening=0.05 * h, treepm_split_scale=1.25 * h,
                     treepm_cutoff=4.5 * 1.25 * h, periodic_box=1.0, pm_mesh=ng,
                     integrator="leapfrog") as s:
        s.set_timing(True)
        s.upload(pos, None, np.full(n, 1.0 / n, np.float32))
        s.step(1e-3, 1)
        t0 = time.perf_counter()
        s.step(1e-3, steps)
        el = time.perf_counter() - t0
        tm = s.timings()
    walks = max(1, tm["walks"])
    return {"workload": f"zeldovich_{ng}^3_box1_pm{ng}_rs1.25cells_theta0.6", "n": n,
            "steps": steps, "value": n * steps / el, "unit": "particle-steps/s",
            "ms_per_step": el * 1e3 / steps, "ic_seconds": t_ic,
            "build_ms": tm["build_ms"] / walks, "walk_and_mesh_ms": tm["walk_ms"] / walks}


def small_configs(with_cpu):
    """The two smaller reference configurations, each on its own handle:
    Plummer 65,536 (seed 42, eps 0.01, theta 0.5, KDK dt 1e-3, 100 steps, FP64
    energy drift) and Plummer 1,048,576 (seed 7, eps 0.005, theta 0.5, dt 5e-4,
    20 steps, forces at steps 0 and 20 against the FP64 direct sum and
    octree_cpu)."""
    import blitzar_amd as B
    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    pos, vel, m = B.plummer(65536, seed=42)
    with B.OctreeHip(65536, theta=0.5, softening=0.01, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        d0 = s.diagnostics(sample_limit=65536)
        s.step(1e-3, 1)
        s.upload(pos, vel, m)
        t0 = time.perf_counter()
        s.step(1e-3, 100)
        el = time.perf_counter() - t0
        d1 = s.diagnostics(sample_limit=65536)
    e0, e1 = d0["ke"] + d0["pe"], d1["ke"] + d1["pe"]
    c1 = {"workload": "plummer_65536_seed42_eps0.01_theta0.5_kdk_dt1e-3",
          "steps": 100, "value": 65536 * 100 / el, "unit": "particle-steps/s",
          "ms_per_step": el * 10.0, "energy_drift": abs(e1 - e0) / abs(e0)}
    if with_cpu:
        from oracle import pyoracle as O
        x = O.hot(pos, m)
        params = O.make_params(theta=0.5, softening=0.01)
        ref = O.ref_octree_forces_subset(x, np.arange(65536), params, nthreads=threads)
        if ref is not None:
            _, tb, tw = ref
            sub = np.arange(0, 65536, 16)
            _, tb1, tw1 = O.ref_octree_forces_subset(x, sub, params, nthreads=1)
            c1["cpu_baseline"] = {
                "value": 65536 / (tb + tw), "unit": "particle-steps/s", "cores": threads,
                "kind": "reference", "cpu_model": cpu_model(),
                "single_thread_value": 65536 / (tb1 + tw1 * 16),
                "sample": "octree_cpu full force pass (one per KDK step) over all 65,536; "
                          "single thread: 4,096 targets extrapolated"}
    out["plummer_65536"] = c1

    n2 = 1 << 20
    pos, vel, m = B.plummer(n2, seed=7)
    rng = np.random.default_rng(2024)
    t = np.sort(rng.choice(n2, 4096, replace=False)).astype(np.uint32)
    with B.OctreeHip(n2, theta=0.5, softening=0.005, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        a0 = s.compute_accelerations()
        r0 = s.direct_forces_fp64(t)
        p0 = pos
        t0 = time.perf_counter()
        s.step(5e-4, 20)
        el = time.perf_counter() - t0
        a20 = s.compute_accelerations()
        r20 = s.direct_forces_fp64(t)
        p20, _ = s.download()
    c2 = {"workload": "plummer_1048576_seed7_eps0.005_theta0.5_kdk_dt5e-4",
          "steps": 20, "value": n2 * 20 / el, "unit": "particle-steps/s",
          "ms_per_step": el * 1e3 / 20, "force_error_step0": force_error(a0[t], r0, 4096),
          "force_error_step20": force_error(a20[t], r20, 4096)}
    if with_cpu:
        from oracle import pyoracle as O
        params = O.make_params(theta=0.5, softening=0.005)
        for tag, pp, aa in (("step0", p0, a0), ("step20", p20, a20)):
            ref = O.ref_octree_forces_subset(O.hot(pp, m), t.astype(np.int32), params,
                                             nthreads=threads)
            if ref is not None:
                rc = ref[0].astype(np.float64)
                d = np.linalg.norm(aa[t] - rc, axis=1) / np.linalg.norm(rc, axis=1)
                c2[f"vs_octree_cpu_{tag}"] = {"median": float(np.median(d)),
                                             "p99": float(np.percentile(d, 99))}
    out["plummer_1048576"] = c2
    return out


def block_time_steps(s, args, n):
    """Dyadic block steps on the same state: theta 0.5, dt_i = 0.02 sqrt(eps /
    |a_i|) clipped to [2^-14, 2^-6] (base step 2^-6, 8 levels, no velocity
    bound). value = particle steps (one per particle per own step) per second."""
    s.set_params(theta=args.ref_theta, adaptive_max_level=8, adaptive_eta=0.02,
                 adaptive_flags=0)
    s.step(2.0 ** -6, 1)  # initialises the schedule, one warm-up base step
    t0 = time.perf_counter()
    s.step(2.0 ** -6, args.block_steps)
    el = time.perf_counter() - t0
    tm = s.timings()
    levels = np.bincount(s.adaptive_levels(), minlength=9)
    slices = max(1, tm["walks"])
    out = {"base_dt": 2.0 ** -6, "levels": 8, "eta": 0.02, "theta": args.ref_theta,
           "base_steps": args.block_steps, "slices": tm["walks"],
           "particle_steps": tm["particle_steps"], "value": tm["particle_steps"] / el,
           "unit": "particle-steps/s", "ms_per_base_step": el * 1e3 / args.block_steps,
           "build_ms_per_slice": tm["build_ms"] / slices, "walk_ms_per_slice": tm["walk_ms"] / slices,
           "level_histogram": levels.tolist()}
    s.set_params(adaptive_max_level=0)
    return out


def learned_far_field(s, args, n, m_now, t, with_error, t05_ms):
    """theta 0.9 walk with the learned correction (32 -> 64 -> 64 -> 3 SiLU
    MLP over the node's moments up to third order and the target geometry,
    BF16 MFMA), trained on a second realisation (seed 12) of the workload,
    applied to the timed state. Reported beside the same walk with a zero
    network (its near nodes of < min_count particles opened, the rest
    uncorrected) and with the ratio to the theta 0.5 walk."""
    import blitzar_amd as B
    band, min_count = args.lff_band, 64
    p_now, v_now = s.download()
    if args.workload == "merger":
        p_tr, v_tr, m_tr = B.hernquist_merger(n // 2, seed=12, separation=20.0, pericentre=2.0)
    else:
        p_tr, v_tr, m_tr = B.plummer(n, seed=12)
    s.set_params(theta=0.9)
    s.lff_configure(theta=0.9, band=band, min_count=min_count)
    s.upload(p_tr, v_tr, m_tr)
    t0 = time.perf_counter()
    tuples = s.lff_collect(group_stride=1, max_tuples=1 << 24)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = s.lff_train(epochs=6, batch=16384, lr=2e-3, seed=3)
    t_train = time.perf_counter() - t0
    params = s.lff_params()
    out = {"theta": 0.9, "band": band, "min_count": min_count, "tuples": tuples,
           "collect_s": t_collect, "train_s": t_train, "final_loss": float(loss[-1])}
    for tag, p in (("zero_network", np.zeros_like(params)), ("learned", params)):
        s.lff_set_params(p)
        s.lff_enable(True)
        s.upload(p_now, v_now, m_now)
        err = None
        if with_error:
            err = force_error(s.compute_accelerations()[t], s.direct_forces_fp64(t), len(t))
        s.step(args.dt, 1)
        t0 = time.perf_counter()
        s.step(args.dt, args.ref_steps)
        el = time.perf_counter() - t0
        tm = s.timings()
        w = max(1, tm["walks"])
        line = {"value": n * args.ref_steps / el, "ms_per_step": el * 1e3 / args.ref_steps,
                "walk_and_mlp_ms": tm["walk_ms"] / w, "mlp_ms": tm["lff_ms"] / w,
                "pairs_per_walk": tm["lff_pairs"] / w, "tiles_per_walk": tm["lff_tiles"] / w,
                "ratio_to_theta_0.5": t05_ms / (el * 1e3 / args.ref_steps),
                "force_error": err}
        if tag == "learned":
            # MLP FLOP: 2 (32 x 64 + 64 x 64 + 64 x 16) per pair slot of a
            # 16-pair tile, on the matrix cores
            flop = 2.0 * (32 * 64 + 64 * 64 + 64 * 16) * 16 * tm["lff_tiles"] / w
            line["mfma_tflops"] = flop / max(tm["lff_ms"] / w * 1e-3, 1e-12) / 1e12
            out.update(line)
        else:
            out[tag] = line
    s.lff_enable(False)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8388608)
    ap.add_argument("--workload", default="merger", choices=["merger", "plummer"])
    ap.add_argument("--criterion", default="relative", choices=["relative", "com", "bounds", "extent"])
    ap.add_argument("--theta", type=float, default=2.0,
                    help="opening angle (under the relative criterion: its upper bound)")
    ap.add_argument("--relative-alpha", type=float, default=0.0125)
    ap.add_argument("--relative-mono", type=float, default=0.2,
                    help="relative criterion: point-mass tolerance as a fraction of alpha (0 = off)")
    ap.add_argument("--far-alpha", type=float, default=0.0,
                    help="group far-field expansion (0 = off)")
    ap.add_argument("--ref-theta", type=float, default=0.5)
    ap.add_argument("--ref-steps", type=int, default=5)
    ap.add_argument("--no-lff", action="store_true",
                    help="skip the learned far-field correction (theta 0.9 + MLP)")
    ap.add_argument("--lff-band", type=float, default=0.75,
                    help="learned correction: near nodes with size / distance >= band")
    ap.add_argument("--block-steps", type=int, default=64,
                    help="base steps of the dyadic block time-step workload (0 = skip)")
    ap.add_argument("--treepm", type=int, default=10,
                    help="steps of the periodic TreePM box, 256^3 Zel'dovich lattice (0 = skip)")
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--bucket", type=int, default=16)
    ap.add_argument("--dt", type=float, default=2.0 ** -10)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-targets", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 65,536 and 1,048,576 Plummer configurations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of MIB MiB three times "
                         "(reads and writes MIB MiB each): the known-size kernel that "
                         "calibrates rocprofv3's read / write request counters")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "octree_hip runs on one GPU: --gpus above 1 is not "
                                   "supported (the walk would shard by Morton ranges of "
                                   "targets, which is not implemented)"}))
        return 2

    import blitzar_amd as B

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)

    t0 = time.perf_counter()
    pos, vel, m = make_state(args.workload, args.n)
    n = pos.shape[0]
    log(f"[bench] state {args.workload} n={n} generated in {time.perf_counter() - t0:.2f} s")

    s = B.OctreeHip(n, theta=args.theta, softening=args.eps, multipole_order=args.order,
                    bucket_capacity=args.bucket, integrator="leapfrog", far_alpha=args.far_alpha,
                    criterion=args.criterion, relative_alpha=args.relative_alpha,
                    relative_mono=args.relative_mono)
    s.upload(pos, vel, m)
    s.set_timing(True)
    s.compute_accelerations(to_host=False)
    if args.warmup > 0:
        s.step(args.dt, args.warmup)
    log("[bench] warmup done")

    t0 = time.perf_counter()
    s.step(args.dt, args.steps)  # returns after the device finished (stream sync)
    elapsed = time.perf_counter() - t0
    tm = s.timings()
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; {tm}")

    walks = max(1, tm["walks"])
    walk_ms = tm["walk_ms"] / walks
    flop_pc = FLOP_PC if args.order >= 2 else FLOP_PC_MONO
    flop = (FLOP_PP * tm["pp_interactions"] + flop_pc * tm["pc_interactions"] +
            FLOP_FAR_NODE * tm["far_interactions"] + FLOP_FAR_EVAL * tm["far_evaluations"]) / walks
    achieved = flop / (walk_ms * 1e-3) / 1e12

    err = None
    ref = t = None
    if not args.no_accuracy:
        acc = s.compute_accelerations()
        p_now, _ = s.download()
        rng = np.random.default_rng(2024)
        t = np.sort(rng.choice(n, args.samples, replace=False)).astype(np.uint32)
        ref = s.direct_forces_fp64(t)
        err = force_error(acc[t], ref, args.samples)
        log(f"[bench] force error vs FP64 direct sum: {err}")
        pos = p_now

    # the same state at the reference opening angle (timed the same way)
    s.set_params(theta=args.ref_theta, far_alpha=0.0, criterion="com")
    s.step(args.dt, 1)
    t0 = time.perf_counter()
    s.step(args.dt, args.ref_steps)
    el_ref = time.perf_counter() - t0
    tm_ref = s.timings()
    ref_line = {"theta": args.ref_theta, "value": n * args.ref_steps / el_ref,
                "ms_per_step": el_ref * 1e3 / args.ref_steps, "steps": args.ref_steps,
                "walk_ms": tm_ref["walk_ms"] / max(1, tm_ref["walks"])}
    if ref is not None:
        ref_line["force_error"] = force_error(s.compute_accelerations()[t],
                                              s.direct_forces_fp64(t), args.samples)
    log(f"[bench] theta {args.ref_theta}: {ref_line}")

    block_line = None
    if args.block_steps > 0:
        block_line = block_time_steps(s, args, n)
        log(f"[bench] block time steps: {block_line}")

    lff_line = None
    if not args.no_lff:
        lff_line = learned_far_field(s, args, n, m, t, ref is not None,
                                     ref_line["ms_per_step"])
        log(f"[bench] learned far field: {lff_line}")

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(pos, m, args.ref_theta, args.eps, args.cpu_targets)
        log(f"[bench] cpu baseline {cpu} ({time.perf_counter() - t0:.1f} s)")
    s.close()

    treepm_line = None
    if args.treepm > 0:
        t0 = time.perf_counter()
        treepm_line = periodic_treepm(args.treepm)
        log(f"[bench] periodic TreePM ({time.perf_counter() - t0:.1f} s): {treepm_line}")

    configs = None
    if not args.no_configs:
        t0 = time.perf_counter()
        configs = small_configs(not args.no_cpu)
        log(f"[bench] smaller configurations ({time.perf_counter() - t0:.1f} s): {configs}")

    value = n * args.steps / elapsed
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": workload_name(args, n),
            "n": n, "criterion": args.criterion, "theta": args.theta,
            "relative_alpha": args.relative_alpha if args.criterion == "relative" else None,
            "relative_mono": args.relative_mono if args.criterion == "relative" else None,
            "far_alpha": args.far_alpha, "softening": args.eps,
            "dt": args.dt,
            "integrator": "leapfrog_kdk_fp64", "bucket": args.bucket,
            "parallelism": "none (1 GPU)",
        },
        "force_error": err,
        "phases_ms_per_step": {
            "build": tm["build_ms"] / walks, "sort": tm["sort_ms"] / walks,
            "walk": walk_ms, "integrate": tm["integrate_ms"] / max(1, args.steps),
        },
        "hbm_fraction": hbm_fractions(tm, n, walks, args),
        "roofline": {
            "kernel": "walk_kernel", "bound": "valu", "pipe": "fp32 VALU",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": walk_traffic(args, n),
            "traffic_source": f"profiles/{os.path.basename(PMC_SUMMARY)} (rocprofv3 PMC)",
            "interactions_per_launch": {"pp": tm["pp_interactions"] / walks,
                                        "pp_point_mass_cells": tm["mono_interactions"] / walks,
                                        "pc": tm["pc_interactions"] / walks,
                                        "far_node_group": tm["far_interactions"] / walks,
                                        "far_evaluations": tm["far_evaluations"] / walks},
        },
        "cpu_baseline": cpu,
        f"theta_{args.ref_theta}": ref_line,
    }
    if lff_line is not None:
        line["learned_far_field"] = lff_line
    if block_line is not None:
        line["block_time_steps"] = block_line
    if configs is not None:
        line["configs"] = configs
    if treepm_line is not None:
        line["periodic_treepm"] = treepm_line
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
