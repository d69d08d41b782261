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
