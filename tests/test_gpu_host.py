"""GPU checks of the C++ drop-ins (ParticleSolverHIP / ParticleSolverBHutHIP via
the headless CLI), the particle-file path and bit-exact snapshot resume."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "n-body_amd" / "bin" / "nbody_cli"
GOLDEN = Path(__file__).resolve().parent / "golden"


def run_cli(*args):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cli_direct_matches_binding(nb, tmp_path):
    out = tmp_path / "final.txt"
    txt = run_cli("-v", 4, "-i", 8, "-n", 2048, "-t", 0.0035, "-s", 40, "--steps", 10, "--seed", 3,
                  "--out", out, "--energy")
    assert "interactions/s" in txt and "drift" in txt
    p, v, a, m = nb.read_particle_file(out)
    pos, vel, mass = nb.ic_plummer(3, 2048, 50.0, 1e4)
    with nb.Simulation(2048, 40.0, 0.0035) as sim:
        sim.upload(nb.to_vec4(pos.astype(np.float32)), nb.to_vec4(vel.astype(np.float32)), mass)
        sim.step(10)
        gp, gv, ga = sim.download()
    np.testing.assert_array_equal(p, gp)
    np.testing.assert_array_equal(v, gv)
    np.testing.assert_array_equal(a, ga)


@pytest.mark.parametrize("init,gen", [(1, "ic_cube"), (2, "ic_galaxy"), (4, "ic_sphere_surface"),
                                      (5, "ic_ball_world"), (6, "ic_cube_surface")])
def test_cli_reference_initializers(nb, tmp_path, init, gen):
    """-i 1, 2, 4, 5, 6 with the reference's defaults (dt 0.001, eps2 0.0935, world
    edge 5): the direct-sum run matches the binding on the same seeded bodies bit
    for bit, and Barnes-Hut runs the same start."""
    n, seed = 3000, 9
    out = tmp_path / "final.txt"
    run_cli("-v", 4, "-i", init, "-n", n, "--steps", 5, "--seed", seed, "--out", out)
    p, v, a, m = nb.read_particle_file(out)
    pos, vel, mass = getattr(nb, gen)(seed, n)
    with nb.Simulation(n, 0.0935, 0.001) as sim:
        sim.upload(nb.to_vec4(pos.astype(np.float32)), nb.to_vec4(vel.astype(np.float32)), mass)
        sim.step(5)
        gp, gv, ga = sim.download()
    np.testing.assert_array_equal(p, gp)
    np.testing.assert_array_equal(v, gv)
    np.testing.assert_array_equal(m, mass)
    out7 = tmp_path / "bh.txt"
    txt = run_cli("-v", 7, "-i", init, "-n", n, "--steps", 5, "--seed", seed, "--out", out7,
                  "--energy")
    assert "drift" in txt
    p7, _, _, _ = nb.read_particle_file(out7)
    assert np.all(np.isfinite(p7))
    # five steps of dt 0.001 move no body far from the direct-sum result
    assert np.abs(p7 - gp).max() < 1e-2 * 5.0


def test_cli_barnes_hut_from_reference_file(tmp_path):
    out = tmp_path / "o.txt"
    txt = run_cli("-v", 7, "-f", GOLDEN / "readme_particles.txt", "-t", 0.0035, "-s", 40,
                  "--steps", 5, "--out", out)
    assert "3 bodies" in txt
    assert out.read_text().startswith("Particle System with 3 particles:")


def test_snapshot_resume_is_bit_exact(nb, tmp_path):
    for solver in (nb.SOLVER_DIRECT, nb.SOLVER_BH):
        pos, vel, m = nb.ic_plummer(77, 5000, 20.0, 500.0)
        snap = tmp_path / f"s{solver}.bin"
        with nb.Simulation(5000, 4.0, 0.002, solver=solver) as sim:
            sim.upload(pos, vel, m)
            sim.step(6)
            sim.save_snapshot(snap)
            sim.step(6)
            full = sim.download_f64()
        with nb.Simulation(5000, 4.0, 0.002, solver=solver) as sim:
            sim.load_snapshot(snap)
            sim.step(6)
            resumed = sim.download_f64()
        for x, y in zip(full, resumed):
            assert np.array_equal(x, y)


def test_cli_snapshot_resume(tmp_path):
    snap, a, b = tmp_path / "s.bin", tmp_path / "a.txt", tmp_path / "b.txt"
    run_cli("-v", 4, "-i", 8, "-n", 3000, "--steps", 4, "--snapshot", snap)
    run_cli("-v", 4, "-i", 8, "-n", 3000, "--steps", 4, "--resume", snap, "--out", a)
    run_cli("-v", 4, "-i", 8, "-n", 3000, "--steps", 8, "--out", b)
    assert a.read_text() == b.read_text()


def test_cli_frames_and_reorder(nb, tmp_path):
    # frames of the render consumer, with and without a Morton reorder: the same
    # bodies in the caller's order, and the last frame is the final state
    prefix = tmp_path / "fr"
    out = tmp_path / "final.txt"
    run_cli("-v", 7, "-i", 8, "-n", 3000, "-t", 0.0035, "-s", 40, "--steps", 6, "--seed", 5,
            "--frames", 3, "--frame-out", prefix, "--reorder", "--out", out)
    frames = sorted(tmp_path.glob("fr_*.xyzs"))
    assert [f.name for f in frames] == ["fr_00000.xyzs", "fr_00001.xyzs", "fr_00002.xyzs"]
    first = np.fromfile(frames[0], np.float32).reshape(-1, 4)
    last = np.fromfile(frames[-1], np.float32).reshape(-1, 4)
    pos, vel, _ = nb.ic_plummer(5, 3000, 50.0, 1e4)
    np.testing.assert_array_equal(first[:, :3], pos.astype(np.float32))
    p, v, _, _ = nb.read_particle_file(out)
    np.testing.assert_array_equal(last[:, :3], p[:, :3].astype(np.float32))
    np.testing.assert_allclose(last[:, 3], np.linalg.norm(v[:, :3].astype(np.float64), axis=1),
                               rtol=1e-6)
