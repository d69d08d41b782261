"""The headless `--run --solver octree_hip` path end to end on the GPU: a
simulation.ini case runs, exports NBSIMBIN v1, and the exported state equals
the same run driven through the C ABI from Python (same initial state,
parameters, substeps)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "blitzar_amd", "bh-headless")


def read_nbsimbin(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"NBSIMBIN"
    version, count = np.frombuffer(raw[8:16], np.uint32)
    assert version == 1
    rec = np.frombuffer(raw[16:], np.float32).reshape(count, 8)
    return rec


def test_headless_run_matches_c_abi(tmp_path):
    import blitzar_amd as B
    n = 3000
    ini = tmp_path / "case.ini"
    ini.write_text(
        f"simulation(particle_count={n}, dt=0.001, solver=octree_hip, integrator=leapfrog)\n"
        "octree(theta=0.5, softening=0.01, criterion=com, leaf_capacity=256)\n"
        "physics(max_acceleration=64, min_softening=0.0001, min_distance2=1e-12, min_theta=0.05)\n"
        "generation(seed=42)\n"
        "scene(style=preset, preset=plummer)\n"
        "export(directory=out, format=bin)\n")
    print("case headless run", flush=True)
    r = subprocess.run([BIN, "--config", str(ini), "--run", "--target-steps", "3",
                        "--export-path", str(tmp_path / "final.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "completed steps=3" in r.stdout
    rec = read_nbsimbin(tmp_path / "final.bin")
    assert rec.shape == (n, 8)

    pos, vel, m = B.plummer(n, 42)
    with B.OctreeHip(n, theta=0.5, softening=0.01, max_acceleration=64.0, max_speed=120.0,
                     integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        dt = float(np.float32(0.001))  # the case file's dt is an FP32 value
        for _ in range(3):
            s.step(dt, 1)
        p, v = s.download()
    assert np.array_equal(rec[:, :3], p.astype(np.float32))
    assert np.array_equal(rec[:, 3:6], v.astype(np.float32))
    assert np.array_equal(rec[:, 6], m)


def test_headless_reads_nbsimbin_input(tmp_path):
    n = 500
    rng = np.random.default_rng(3)
    rec = np.zeros((n, 8), np.float32)
    rec[:, :3] = rng.standard_normal((n, 3))
    rec[:, 6] = 1.0 / n
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(b"NBSIMBIN" + np.array([1, n], np.uint32).tobytes() + rec.tobytes())
    ini = tmp_path / "case.ini"
    ini.write_text("simulation(particle_count=10, dt=0.001, solver=octree_hip, integrator=euler)\n"
                   "octree(theta=0.5, softening=0.05)\n"
                   "scene(style=file, file=in.bin, format=bin)\n")
    r = subprocess.run([BIN, "--config", str(ini), "--run", "--target-steps", "1",
                        "--export-path", str(tmp_path / "o.xyz")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(tmp_path / "o.xyz").read().splitlines()
    assert int(lines[0]) == n and len(lines) == n + 2
