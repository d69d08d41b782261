"""The reference's own result checks, run on octree_hip through the C ABI.

The reference carries no golden force vectors; what it checks are physics
envelopes and replays (engine/physics/octree/tests/*.cpp,
docs/quality/manifest/numerical_campaign.json). Each test below restates one
of them for octree_hip, with the reference's scenario parameters and
thresholds:

  * TST_UNT_PHYS_009 solver parity (solver.cpp:91-152, campaign
    "disk_solver_parity"): a 96-body disk orbit, 36 Euler steps, against the
    FP64 pairwise sum driven by the same update;
  * TST_UNT_PHYS_001-005 two-body cases (orbit.cpp:18-112) on
    tests/golden/two_body_rest.xyz;
  * TST_UNT_PHYS_012 Plummer calibration (multibody.cpp:104-133);
  * the 100-particle multi-step stability regression
    (octree_gpu_stability.cpp:24-69);
  * TST_UNT_PHYS_014/015 bitwise replay (extended_solvers.cpp:15-56), here
    also at 200,000 particles with the relative criterion, for block time
    steps and for the periodic short-range walk;
  * TST_UNT_PHYS_019 (octree_criteria.cpp:91-112) on the committed golden
    vectors of the reference build (tests/golden/octree_cpu_ref.npz);
  * the resolved theta / softening of a created handle
    (PhyForceLawPolicy.cpp:22-32).

The scenario generators are restated from the reference's
(SrvDisk.cpp:16-80, SrvPlummer.cpp:17-104) with numpy's generator, so the
states have the reference's distributions, not its exact draws. The
reference runs some of these with RK4, which octree_hip does not have: those
run the leapfrog (kick-drift-kick) instead, and the thresholds are the
reference's.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_xyz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def disk_orbit(n, seed, disk_mass, r_min=1.5, r_max=11.5, central_mass=1.0,
               max_acc=64.0):
    """SrvDisk.cpp:16-80: a central body at rest at the origin plus n - 1
    particles on circular orbits, radius uniform in [r_min, r_max], in the
    plane; speed sqrt(min(M(<r) / r^2, max_acc) r) with M(<r) the central mass
    plus the disk mass inside r (by r^2)."""
    rng = np.random.default_rng(seed)
    k = n - 1
    r = rng.uniform(r_min, r_max, k)
    ang = rng.uniform(0.0, 2.0 * np.pi, k)
    frac = np.clip((r * r - r_min * r_min) / (r_max * r_max - r_min * r_min), 0.0, 1.0)
    enc = np.maximum(1e-6, central_mass + disk_mass * frac)
    v = np.sqrt(np.minimum(enc / (r * r), max_acc) * r)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    pos[1:, 0], pos[1:, 1] = r * np.cos(ang), r * np.sin(ang)
    vel[1:, 0], vel[1:, 1] = -np.sin(ang) * v, np.cos(ang) * v
    m = np.full(n, disk_mass / k)
    m[0] = central_mass
    return pos, vel, m.astype(np.float32)


def plummer_preset(n, seed, scale, particle_mass, velocity_scale):
    """SrvPlummer.cpp:17-104: r = a / sqrt(u^(-2/3) - 1), u in [1e-4, 0.9999],
    isotropic directions, Gaussian velocities of sigma sqrt(M / 6a) times the
    velocity scale, recentred on the mean position and velocity."""
    rng = np.random.default_rng(seed)
    total = particle_mass * n
    sigma = np.sqrt(total / (6.0 * scale)) * velocity_scale
    u = rng.uniform(1e-4, 0.9999, n)
    ct = rng.uniform(-1.0, 1.0, n)
    az = rng.uniform(0.0, 2.0 * np.pi, n)
    rad = scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    pos = np.stack([rad * st * np.cos(az), rad * st * np.sin(az), rad * ct], 1)
    vel = rng.normal(0.0, sigma, (n, 3))
    pos -= pos.mean(0)
    vel -= vel.mean(0)
    return pos, vel, np.full(n, total / n, np.float32)


def pairwise_euler(oracle, pos, vel, m, eps, dt, steps):
    """The FP64 pairwise baseline with the reference CPU update
    (OctParticleSystemHost.cpp:61-69): v += a dt, then x += v dt."""
    pos, vel = pos.copy(), vel.copy()
    t = np.arange(len(m), dtype=np.int32)
    for _ in range(steps):
        a = oracle.direct_fp64(pos, m.astype(np.float64), t, eps=eps)
        vel += a * dt
        pos += vel * dt
    return pos, vel


def com(pos, m):
    return (pos * m[:, None]).sum(0) / m.sum()


def avg_radius(pos):
    return np.linalg.norm(pos, axis=1).mean()


def test_solver_parity_disk_orbit_96(oracle):
    import blitzar_amd as B
    n, dt, steps, theta, eps = 96, 0.004, 36, 0.35, 0.08
    say(f"case solver parity disk orbit n={n} euler {steps} steps")
    pos, vel, m = disk_orbit(n, 12345, disk_mass=0.75)
    m64 = m.astype(np.float64)
    eps32 = float(np.float32(eps))
    p_ref, v_ref = pairwise_euler(oracle, pos, vel, m, eps32, dt, steps)
    with B.OctreeHip(n, theta=theta, softening=eps, integrator="euler") as s:
        s.upload(pos, vel, m)
        s.step(dt, steps)
        p_hip, v_hip = s.download()
    # the reference's own CPU solver in the same envelope (oracle restatement)
    x = oracle.hot(pos, m)
    v32 = vel.astype(np.float32)
    params = oracle.make_params(theta=theta, softening=eps)
    for _ in range(steps):
        oracle.euler_update(x, v32, dt, 0.0, params)
    e_ref = oracle.energy(p_ref, v_ref, m64, eps32)
    for name, p, v in (("octree_hip", p_hip, v_hip),
                       ("octree_cpu", x[:, :3].astype(np.float64), v32.astype(np.float64))):
        e = oracle.energy(p, v, m64, eps32)
        de = abs((e["ke"] + e["pe"]) - (e_ref["ke"] + e_ref["pe"])) / abs(e_ref["ke"] + e_ref["pe"])
        dcom = np.linalg.norm(com(p, m64) - com(p_ref, m64))
        drad = abs(avg_radius(p) - avg_radius(p_ref))
        dmax = np.linalg.norm(p - p_ref, axis=1).max()
        say(f"  {name}: com {dcom:.2e} radius {drad:.2e} energy {100 * de:.3e}% "
            f"max particle {dmax:.2e}")
        assert dcom <= dt              # EXPECT_LE(comDelta, kDefaultSimulationDt)
        assert drad <= 0.05
        assert 100.0 * de <= 0.25
        assert dmax <= 0.05


def _two_body(steps, dt, s_kw=None):
    import blitzar_amd as B
    pos, mass = read_xyz(os.path.join(GOLDEN, "two_body_rest.xyz"))
    kw = dict(theta=1.2, softening=2.5, integrator="leapfrog")
    kw.update(s_kw or {})
    energies = []
    with B.OctreeHip(2, **kw) as s:
        s.upload(pos, None, mass.astype(np.float32))
        d = s.diagnostics()
        energies.append(d["ke"] + d["pe"])
        for _ in range(steps):
            s.step(dt, 1)
            d = s.diagnostics()
            energies.append(d["ke"] + d["pe"])
        p, v = s.download()
    return pos, mass, p, v, np.array(energies)


def test_two_body_attraction_com_energy_and_long_run():
    say("case two body orbit.cpp cases")
    # TST_UNT_PHYS_001 attraction and TST_UNT_PHYS_003 centre of mass: 400 x 0.002
    p0, m, p, _, _ = _two_body(400, 0.002)
    ratio = np.linalg.norm(p[0] - p[1]) / np.linalg.norm(p0[0] - p0[1])
    assert 0.05 <= ratio <= 0.999
    assert np.linalg.norm(com(p, m) - com(p0, m)) <= 1e-2
    # TST_UNT_PHYS_002 energy: 800 x 0.002, max |drift| <= 5 %
    _, _, _, _, e = _two_body(800, 0.002)
    assert np.max(np.abs(e - e[0]) / abs(e[0])) * 100.0 <= 5.0
    # TST_UNT_PHYS_004 time-step convergence: 100 x 0.002 against 200 x 0.001
    _, _, pc, _, _ = _two_body(100, 0.002)
    _, _, pf, _, _ = _two_body(200, 0.001)
    assert np.linalg.norm(pc - pf, axis=1).max() <= 0.05
    # TST_UNT_PHYS_005 long run: 2000 x 0.001, finite and bounded
    _, _, p, _, _ = _two_body(2000, 0.001)
    assert np.all(np.isfinite(p)) and np.linalg.norm(p, axis=1).max() <= 50.0


def test_plummer_calibration_96():
    import blitzar_amd as B
    n, dt, steps = 96, 0.001, 80
    say(f"case plummer calibration n={n} euler {steps} steps")
    pos, vel, m = plummer_preset(n, 42, scale=4.0, particle_mass=0.01, velocity_scale=0.75)
    drift = 0.0
    with B.OctreeHip(n, theta=1.2, softening=2.5, integrator="euler") as s:
        s.upload(pos, vel, m)
        d = s.diagnostics()
        e0 = d["ke"] + d["pe"]
        for _ in range(steps):
            s.step(dt, 1)
            d = s.diagnostics()
            drift = max(drift, abs(d["ke"] + d["pe"] - e0) / abs(e0))
        p, _ = s.download()
    m64 = m.astype(np.float64)
    rmax = np.linalg.norm(p, axis=1).max()
    say(f"  energy drift {100 * drift:.3e}% avg radius {avg_radius(p):.3f} max radius {rmax:.2f}")
    assert 100.0 * drift <= 0.01
    assert 4.0 <= avg_radius(p) <= 12.0
    assert rmax <= 45.0
    assert np.linalg.norm(com(p, m64)) <= 0.02


def test_small_batch_multistep_stability_100():
    import time
    import blitzar_amd as B
    n = 100
    say(f"case small batch stability n={n} 10 steps")
    rng = np.random.default_rng(42)
    pos = rng.uniform(-1.8, 1.8, (n, 3))
    m = np.ones(n, np.float32)
    t0 = time.perf_counter()
    with B.OctreeHip(n, theta=0.5, softening=0.01, integrator="leapfrog") as s:
        s.upload(pos, None, m)
        s.step(0.001, 10)
        p, v = s.download()
    assert time.perf_counter() - t0 < 10.0
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))


def _replay(n, make, steps, dt, **kw):
    import blitzar_amd as B
    out = []
    for _ in range(2):
        pos, vel, m = make()
        with B.OctreeHip(n, **kw) as s:
            s.upload(pos, vel, m)
            s.step(dt, steps)
            out.append(s.download())
    (pa, va), (pb, vb) = out
    return np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_bitwise_replay():
    import blitzar_amd as B
    say("case replay disk orbit n=128 euler 25 steps")
    assert _replay(128, lambda: disk_orbit(128, 33333, 0.75, r_max=12.0), 25, 0.004,
                   theta=0.35, softening=0.08, integrator="euler")
    say("case replay plummer n=200000 relative criterion 5 kdk steps")
    assert _replay(200000, lambda: B.plummer(200000, seed=3), 5, 1e-3, theta=2.0,
                   softening=0.01, criterion="relative", relative_alpha=0.0125,
                   relative_mono=0.2, integrator="leapfrog")
    say("case replay hernquist n=100000 block time steps")
    assert _replay(100000, lambda: B.hernquist(100000, seed=4), 2, 2.0 ** -6, theta=0.5,
                   softening=0.01, adaptive_max_level=6, adaptive_eta=0.02, adaptive_flags=0)
    say("case replay periodic short range n=60000")
    rng = np.random.default_rng(8)
    box = rng.random((60000, 3))
    h = 1.0 / 64
    assert _replay(60000, lambda: (box, None, np.full(60000, 1.0 / 60000, np.float32)), 3,
                   1e-3, theta=0.6, softening=0.1 * h, treepm_split_scale=1.25 * h,
                   treepm_cutoff=4.5 * 1.25 * h, periodic_box=1.0)


def test_criterion_fixtures_of_the_reference_build(oracle):
    """TST_UNT_PHYS_019 and the reference build's golden vectors through the
    GPU walk. The 9-body fixture fits one leaf of 16 in both trees, so both
    criteria sum it directly: equal to the reference's forces to its rsqrt
    rounding. On the 2,000-particle cloud (theta 0.5 and 0.7, monopole like
    the reference), the GPU group walk opens more than the reference's
    per-target walk: its error against the FP64 sum stays within the
    reference's own, and the bounds criterion is again the more conservative."""
    import blitzar_amd as B
    g = np.load(os.path.join(GOLDEN, "octree_cpu_ref.npz"))
    x9 = g["criterion9_xyzm"]
    say("case criterion fixture 9 bodies")
    for crit, name in ((0, "com"), (1, "bounds")):
        with B.OctreeHip(9, theta=0.9, softening=0.01, criterion=name, multipole_order=1) as s:
            s.upload(x9[:, :3].astype(np.float64), None, x9[:, 3].astype(np.float32))
            a = s.compute_accelerations()
        ref = g[f"criterion9_acc_c{crit}"].astype(np.float64)
        assert np.linalg.norm(a - ref, axis=1).max() <= 1e-5 * np.linalg.norm(ref, axis=1).max()
    xc = g["cloud2k_xyzm"]
    n = xc.shape[0]
    pos, m = xc[:, :3].astype(np.float64), xc[:, 3].astype(np.float32)
    t = np.arange(n, dtype=np.int32)
    exact = oracle.direct_fp64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64),
                               t, eps=float(np.float32(0.01)))
    for th in (0.5, 0.7):
        errs = {}
        for crit, name in ((0, "com"), (1, "bounds")):
            with B.OctreeHip(n, theta=th, softening=0.01, criterion=name,
                             multipole_order=1) as s:
                s.upload(pos, None, m)
                a = s.compute_accelerations()
            mine = np.median(np.linalg.norm(a - exact, axis=1) / np.linalg.norm(exact, axis=1))
            ref = g[f"cloud2k_acc_t{int(th * 10)}_c{crit}"].astype(np.float64)
            theirs = np.median(np.linalg.norm(ref - exact, axis=1) / np.linalg.norm(exact, axis=1))
            say(f"  cloud2k theta={th} {name}: octree_hip {mine:.3e} reference {theirs:.3e}")
            assert mine <= theirs + 1e-6
            errs[name] = mine
        assert errs["bounds"] <= errs["com"] + 1e-6


def test_created_handle_resolves_theta_and_softening(oracle):
    import blitzar_amd as B
    say("case resolved parameters")
    cases = [dict(theta=0.01, softening=0.0, min_softening=1e-4, min_theta=0.05),
             dict(theta=9.0, softening=0.5, min_softening=1e-4, min_theta=0.05),
             dict(theta=0.3, softening=1e-6, min_softening=1e-3, min_theta=0.4),
             dict(theta=0.7, softening=0.02, min_softening=-1.0, min_theta=7.0),
             dict(theta=0.5, softening=0.01, min_softening=1e-4, min_theta=0.0)]
    for kw in cases:
        with B.OctreeHip(16, **kw) as s:
            got = s.resolved_params()
        th, so, _ = oracle.resolve_policy(oracle.make_params(
            theta=kw["theta"], softening=kw["softening"], min_softening=kw["min_softening"],
            min_theta=kw["min_theta"]))
        assert got.theta == pytest.approx(th, rel=1e-7), kw
        assert got.softening == pytest.approx(so, rel=1e-7), kw
