"""Host logic of the headless runner (bh-headless): directive INI parsing of
the reference's own simulation.ini (tests/golden/simulation_reference.ini,
copied from the reference as data), solver normalisation, the theta clamp and
the exit codes of apps/headless/src/main.cpp. No GPU is touched."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "blitzar_amd", "bh-headless")


def run(*args):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=60)


def effective(out):
    lines = out.split("[headless] effective configuration begin\n")[1]
    return lines.split("[headless] effective configuration end")[0]


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(BIN):
        pytest.skip("bh-headless not built")


def test_reference_simulation_ini_loads_unchanged():
    r = run("--config", os.path.join(GOLDEN, "simulation_reference.ini"), "--inspect")
    assert r.returncode == 0, r.stderr
    eff = effective(r.stdout)
    assert "particle_count=3073" in eff and "solver=octree_gpu" in eff
    assert "theta=1.2" in eff and "softening=2.5" in eff and "leaf_capacity=256" in eff
    assert "max_acceleration=64" in eff and "min_distance2=1e-12" in eff
    assert "malformed" not in r.stderr


def test_validate_rejects_other_solvers_and_accepts_aliases(tmp_path):
    ini = tmp_path / "c.ini"
    ini.write_text("simulation(particle_count=100, dt=0.001, solver=octree_cpu)\n")
    assert run("--config", str(ini), "--validate").returncode == 4
    for alias in ("octree_hip", "octree-hip", "OCTREE_GPU_HIP"):
        assert run("--config", str(ini), "--validate", "--solver", alias).returncode == 0


def test_exit_codes(tmp_path):
    assert run("--config", str(tmp_path / "missing.ini"), "--run").returncode == 3
    ini = tmp_path / "c.ini"
    ini.write_text("simulation(solver=octree_hip)\n")
    assert run("--config", str(ini)).returncode == 2          # no command
    assert run("--config", str(ini), "--bogus").returncode == 2
    assert run("--config", str(ini), "--run", "--target-steps", "-1").returncode == 2
    assert run("--help").returncode == 0


def test_directive_grammar(tmp_path):
    ini = tmp_path / "c.ini"
    ini.write_text(
        "# comment\n"
        "simulation(particle_count = 77 , dt=0.002, solver=\"octree_hip\", integrator=leapfrog)\n"
        "octree(theta=0.3, theta_auto_min=0.4, theta_auto_max=1.2, criterion=bounds)\n"
        "scene(style=preset, preset=plummer, file=\"a, b.bin\")\n"
        "broken(line\n"
        "octree_hip(bucket=8, order=1)\n")
    r = run("--config", str(ini), "--inspect")
    assert r.returncode == 0
    eff = effective(r.stdout)
    assert "particle_count=77" in eff and "dt=0.002" in eff and "integrator=leapfrog" in eff
    assert "criterion=bounds" in eff and "bucket=8, order=1" in eff
    assert 'file="' in eff and "a, b.bin" in eff
    assert "malformed directive" in r.stderr


def resolved(out):
    line = [ln for ln in out.splitlines() if ln.startswith("[headless] resolved ")][0]
    return dict(kv.split("=") for kv in line.split()[2:])


def test_theta_is_clamped_into_the_auto_bounds(tmp_path):
    """resolveOctreeTheta (SrvFormatAndTheta.cpp:54-69) clamps theta into
    [theta_auto_min, theta_auto_max] even with auto-tune off, then
    resolveForceLawPolicy clamps it into [min_theta, 4]."""
    ini = tmp_path / "c.ini"
    cases = [("theta=0.3, theta_auto_min=0.4, theta_auto_max=1.2", 0.4),
             ("theta=1.5, theta_auto_min=0.4, theta_auto_max=1.2", 1.2),
             ("theta=0.7, theta_auto_min=0.4, theta_auto_max=1.2", 0.7),
             ("theta=0.3, theta_auto_min=0.01, theta_auto_max=1.2", 0.3),
             ("theta=0.01, theta_auto_min=0.01, theta_auto_max=0.02", 0.05),
             ("theta=0.3, theta_auto_min=0.9, theta_auto_max=0.5", 0.9)]
    for octree, want in cases:
        ini.write_text(f"simulation(solver=octree_hip)\noctree({octree}, softening=0.02)\n"
                       "physics(min_softening=0.05)\n")
        r = run("--config", str(ini), "--inspect")
        assert r.returncode == 0, r.stderr
        got = resolved(r.stdout)
        assert float(got["theta"]) == pytest.approx(want, rel=1e-6), octree
        assert float(got["softening"]) == pytest.approx(0.05, rel=1e-6)  # max(eps, min_softening)


@pytest.mark.parametrize("payload", ["truncated", "oversized", "too_few"])
def test_snapshot_header_counts_are_checked_before_allocating(tmp_path, payload):
    """A corrupt NBSIMBIN header must fail the parse (and fall back to the
    generated scene, as the reference does) instead of asking for gigabytes."""
    import struct
    path = tmp_path / "bad.bin"
    count = {"truncated": 1000, "oversized": 0xFFFFFFF0, "too_few": 1}[payload]
    body = b"\0" * 32 * (10 if payload != "too_few" else 1)
    path.write_bytes(b"NBSIMBIN" + struct.pack("<II", 1, count) + body)
    ini = tmp_path / "c.ini"
    ini.write_text("simulation(particle_count=64, dt=0.001, solver=octree_hip)\n"
                   f"scene(style=file, file={path}, format=bin)\n")
    r = run("--config", str(ini), "--run", "--target-steps", "1")
    assert "failed to parse input state" in r.stderr
    assert "outside [2, 2^31 - 1] or beyond the file size" in r.stderr
    assert r.returncode in (0, 5)  # 5 where no GPU is present; never an abort
