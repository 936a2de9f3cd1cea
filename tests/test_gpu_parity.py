"""GPU parity: the HIP collision step against the CPU oracle, bitwise.

Every test drives both through the reference's per-step loop (predict,
resolveCollision) on the same seeded scene and compares, after each step,
the per-pass candidate / narrow / hit counts, the set of hit element pairs
with their sub-test masks, the impact-zone membership, and the complete
vertex state (positions, velocities, average velocities, has_collsn), all
bit for bit.
"""
import numpy as np
import pytest

from collision_amd import scenes as S
from tests.parity import run_parity

pytestmark = pytest.mark.gpu


def test_config0_full_300_steps():
    """Config 0: 32x32 cloth onto a static 642-vertex icosphere, 300 steps.
    Contact starts near step 249 and includes CCD passes, an impact-zone
    escalation and strain limiting."""
    st = run_parity(S.config0(), steps=300, verbose=True)
    assert st["hit_pairs"] > 1000
    assert st["zone_steps"] >= 1


def test_config1_reduced_self_collision():
    """Config 1 at 64x64 (same generator, amplitude and velocity law)."""
    sc = S.self_collision_cloth(n=64, steps=12)
    st = run_parity(sc, steps=12, verbose=True)
    assert st["hit_pairs"] > 0


@pytest.mark.parametrize("full", [0, 1], ids=["incremental", "full_passes"])
def test_config4_reduced_cloth_strings_boxes(full):
    """Config 4 geometry at 96x96 cloth, 60 strings, friction 0.3; with the
    incremental later passes (default) and with every pass recomputed."""
    sc = S.combined_scene(n=96, nstrings=60, steps=60)
    st = run_parity(sc, steps=60, check_every=5, verbose=True, gpu_kw=dict(force_full_passes=full))
    assert st["hit_pairs"] > 0
    assert st["zone_steps"] > 0


def test_strain_more_sweeps_than_one_launch():
    """The config-4 scene at 96x96 with the strain loop allowed 23 sweeps
    instead of ten (dcollid.cpp:590): from step 26 on every step runs all 23,
    which the device does as three launches of sweeps in flight (10, 10, 3),
    each starting from the previous launch's velocities."""
    sc = S.combined_scene(n=96, nstrings=60, steps=40)
    st = run_parity(sc, steps=40, check_every=5, verbose=True, max_strain_sweeps=23)
    assert st["strain_sweeps"] > 10 * 23


def test_strain_fan_one_sweep_in_flight():
    """A disc meshed as a fan of 96 triangles around one vertex: its visits
    span so many levels that only one sweep is in flight per launch, and
    velocity changes for visits 64 or more levels ahead are flagged in
    global memory instead of the on-chip ring; 23 sweeps allowed."""
    sc = S.fan_disc_on_box()
    st = run_parity(sc, steps=sc.steps, verbose=True, max_strain_sweeps=23)
    assert st["strain_sweeps"] > 5 * 23
    assert st["hit_pairs"] > 0


@pytest.mark.parametrize("maker", [S.deck_ball_ball, S.deck_string_box, S.deck_string_string,
                                   S.deck_string_on_static_box])
def test_reference_deck_scenarios(maker):
    sc = maker()
    st = run_parity(sc, steps=min(sc.steps, 120), check_every=10)
    assert st["steps"] > 0


def test_cpp_host_api_matches_oracle(tmp_path):
    """The C++ CollisionSolver3d wrapper (include/collision_amd/CollisionSolver3d.hpp)
    driving a cloth onto a static box gives the oracle's positions bit for bit."""
    import os
    import subprocess

    from oracle import oracle as O

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "cloth_on_box"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "cloth_on_box.cpp"), "-o", str(exe),
                    "-L", os.path.join(root, "collision_amd"), "-lcollision_amd",
                    "-Wl,-rpath," + os.path.join(root, "collision_amd")], check=True)
    out = tmp_path / "x.bin"
    steps = 40
    r = subprocess.run([str(exe), str(out), str(steps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x_gpu = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    b = S._Builder()
    x, t = S.cloth_grid(24, 0.5, origin=(-0.25, -0.25), z=0.012)
    v = np.zeros_like(x)
    v[:, 2] = -1.0
    b.add_surface(x, v, t, fixed=False)
    bx, bt = S.box_mesh((0.0, 0.0, -0.05), (0.2, 0.2, 0.1))
    b.add_surface(bx, np.zeros_like(bx), bt, fixed=True)
    sc = b.build("cloth_on_box", thickness=1e-3, dt=1e-3, steps=steps)
    orc = O.OracleSolver(sc)
    hits = 0
    for _ in range(steps):
        hits += orc.step().hit_pairs
    assert hits > 0
    assert np.array_equal(orc.get_state()["x"], x_gpu)
