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


# ---------------------------------------------------------------------------
def test_device_math_is_bitwise_the_oracles():
    """The device's acos / cos / sin / pow(x, 1/3) (cs_math_eval) equal the
    oracle's correctly rounded functions bit for bit on 10^6 arguments each
    (the oracle's are pinned against 70-digit arithmetic in test_oracle_math)."""
    from collision_amd.solver import math_eval
    from oracle import oracle as O
    rng = np.random.default_rng(21)
    n = 1 << 20
    cases = {"acos": np.concatenate([rng.uniform(-1, 1, n), [-1.0, 1.0, 0.0, 0.5, -0.5]]),
             "cos": np.concatenate([rng.uniform(-2.2, 2.2, n), rng.uniform(-40, 40, n // 8)]),
             "sin": np.concatenate([rng.uniform(-2.2, 2.2, n), rng.uniform(-1e-3, 1e-3, n // 8)]),
             "pow_third": np.concatenate([10.0 ** rng.uniform(-300, 300, n), [0.0, 8.0, 27.0]])}
    for name, x in cases.items():
        g, o = math_eval(name, x), O.math_eval(name, x)
        assert np.array_equal(g.view(np.uint64), o.view(np.uint64)), name


@pytest.mark.parametrize("maker", [S.deck_ball_plane, S.deck_box_boundary, S.deck_string_slide])
def test_movable_rigid_body_decks(maker):
    """Movable rigid bodies (is_movableRG): body zones, RG impulses spread
    over the body, rigid averaging in every apply, centre-of-mass update;
    restated from in-ball_plane, in-box_boundary and in-string_slide."""
    sc = maker()
    st = run_parity(sc, steps=min(sc.steps, 150), verbose=True)
    assert st["hit_pairs"] > 0


def test_domain_box_fires_on_gpu():
    """Strings driven into a tight domain box, one starting outside it on one
    axis: the domain update repeats once per element of a point
    (dcollid.cpp:119-156)."""
    st = run_parity(S.domain_wall_strings(), steps=40, verbose=True)
    assert st["steps"] == 40


def test_nonfinite_average_velocity_stops_before_any_pass():
    """computeAverageVelocity calls clean_up(ERROR) on NaN (dcollid.cpp:178-184):
    both stop with the non-finite status before the proximity pass."""
    from collision_amd.solver import CollisionError, solver_for_scene
    from oracle import oracle as O
    sc = S.cloth_over_sphere(n=8, level=1)
    sc.vel[3, 1] = np.nan
    orc = O.OracleSolver(sc)
    ro = orc.step()
    orc.close()
    gpu = solver_for_scene(sc)
    with pytest.raises(CollisionError) as ei:
        gpu.step()
    assert ei.value.status == ro.status == 2
    assert gpu.L.cs_pass_count(gpu.h) == 0
    gpu.close()


def test_zone_cap_warning_matches():
    """With the impact-zone loop capped at one pass both report
    CS_WARN_ZONE_CAP in the same steps and finish those steps identically."""
    sc = S.config0()
    sc.x = sc.x + sc.vel * (sc.dt * 236)
    st = run_parity(sc, steps=64, max_zone_passes=1, verbose=True)
    assert st["zone_steps"] >= 1


@pytest.mark.parametrize("full", [0, 1], ids=["incremental", "full_passes"])
def test_shipped_mode_without_parity_taps(full):
    """The library as shipped (log_pairs = 0: no pair stream carried between
    incremental passes, their candidate counts -1) against the oracle: hit
    counts, zones and the complete state after every step."""
    sc = S.combined_scene(n=96, nstrings=60, steps=60)
    st = run_parity(sc, steps=60, gpu_log_pairs=0, gpu_kw=dict(force_full_passes=full))
    assert st["hit_pairs"] > 0 and st["zone_steps"] > 0


def test_shipped_mode_without_parity_taps_config0():
    sc = S.config0()
    sc.x = sc.x + sc.vel * (sc.dt * 236)
    st = run_parity(sc, steps=64, gpu_log_pairs=0)
    assert st["hit_pairs"] > 0


def _golden_run(name):
    import json
    import os

    from collision_amd.solver import solver_for_scene
    from tests.golden.make_step_golden import PASS_KEYS, REPORT_KEYS, WORKLOADS, log_hash, state_hash
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"steps_{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    gold = json.load(open(path))
    sc = WORKLOADS[name][0]()
    assert (sc.nv, sc.nt, sc.nb) == (gold["nv"], gold["nt"], gold["nb"])
    gpu = solver_for_scene(sc, log_pairs=1)
    try:
        for k, rec in enumerate(gold["records"]):
            r = gpu.step().as_dict()
            print(f"[golden] {name} step {k} passes {r['n_passes']} hits {r['hit_pairs']}", flush=True)
            assert {key: int(r[key]) for key in REPORT_KEYS} == rec["report"], k
            assert [{key: p[key] for key in PASS_KEYS} for p in gpu.passes()] == rec["passes"], k
            assert log_hash(gpu.pairLog()) == rec["log"], k
            assert state_hash(gpu.getState()) == rec["state"], k
    finally:
        gpu.close()
    return gold


def test_config1_full_size_against_golden():
    """Config 1 at full size (256 x 256 cloth, 100 steps): every step's
    counts, hit pairs and complete state bitwise equal the oracle's
    (committed golden records, tests/golden/make_step_golden.py)."""
    g = _golden_run("config1_256")
    assert sum(r["report"]["hit_pairs"] for r in g["records"]) > 10000


def test_config4_256_against_golden():
    g = _golden_run("config4_256")
    assert sum(r["report"]["n_zone_passes"] for r in g["records"]) > 0


def test_config4_full_size_incremental_equals_full_passes():
    """Full-size config 4 (1024 x 1024 cloth, 2000 strings), 15 contact steps
    from 8 steps before the first box contact: status OK, no non-finite value,
    and the incremental passes give bitwise the state of all-full passes."""
    from collision_amd.solver import solver_for_scene
    sc = S.config4()
    sc.x = sc.x + sc.vel * (sc.dt * 2)
    a = solver_for_scene(sc, force_full_passes=0)
    b = solver_for_scene(sc, force_full_passes=1)
    hits = 0
    try:
        for k in range(15):
            ra, rb = a.step(), b.step()
            print(f"[c4full] step {k} passes {ra.n_passes}/{rb.n_passes} hits {ra.hit_pairs}",
                  flush=True)
            assert ra.status == 0 and rb.status == 0
            for key in ("n_passes", "n_ccd_passes", "n_zone_passes", "hit_pairs", "hit_subtests",
                        "strain_sweeps", "n_zones"):
                assert getattr(ra, key) == getattr(rb, key), (k, key)
            hits += ra.hit_pairs
        sa, sb = a.getState(), b.getState()
        for key in ("x", "vel", "avg_vel", "has_collsn"):
            assert np.array_equal(sa[key], sb[key]), key
        assert np.all(np.isfinite(sa["x"])) and np.all(np.isfinite(sa["vel"]))
        assert hits > 0
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("maker,n", [(lambda: S.combined_scene(n=96, nstrings=60), 20),
                                     (S.deck_box_boundary, 40)], ids=["config4_96", "box_boundary"])
def test_checkpoint_restart_is_bitwise(tmp_path, maker, n):
    """N + N steps equal N steps, cs_save, cs_load into a fresh handle, N
    steps: every per-vertex array and the body state, bit for bit."""
    from collision_amd.solver import solver_for_scene
    sc = maker()
    a = solver_for_scene(sc)
    for _ in range(n):
        a.step()
    ck = tmp_path / "state.ckpt"
    a.save(ck)
    for _ in range(n):
        a.step()
    b = solver_for_scene(sc, friction=0.77)     # parameters come from the checkpoint
    b.load(ck)
    assert b.params.friction == sc.friction
    for _ in range(n):
        b.step()
    sa, sb = a.getState(), b.getState()
    for k in ("x", "vel", "avg_vel", "x_old", "has_collsn"):
        assert np.array_equal(sa[k], sb[k]), k
    (ca, va), (cb, vb) = a.getBodies(), b.getBodies()
    assert np.array_equal(ca, cb) and np.array_equal(va, vb)
    a.close()
    b.close()


def test_vtk_dump_of_the_last_pass(tmp_path):
    """cs_write_vtk (vtkplotVectorSurface, vtk.cpp:14-93): the summed impulses
    of each step's last pass with a hit equal the oracle's bit for bit, and
    the .vtp file holds
    the current points, one polygon per element, the impulses and red cells
    exactly where a point had a hit."""
    import xml.etree.ElementTree as ET

    from collision_amd.solver import solver_for_scene
    from oracle import oracle as O
    sc = S.config0()
    sc.x = sc.x + sc.vel * (sc.dt * 246)
    gpu = solver_for_scene(sc, record_impulses=1)
    orc = O.OracleSolver(sc, record_impulses=1)
    for _ in range(6):
        gpu.step()
        orc.step()
        ig, ng = gpu.lastImpulses()
        io, no = orc.last_impulses()
        assert np.array_equal(ig, io) and np.array_equal(ng, no)
    assert ng.sum() > 0
    path = tmp_path / "surf.vtp"
    gpu.writeVtk(path)
    piece = ET.parse(path).getroot().find("PolyData/Piece")
    assert int(piece.get("NumberOfPoints")) == sc.nv and int(piece.get("NumberOfPolys")) == sc.nt
    # 9 significant digits give back every float32 value exactly
    arrays = {d.get("Name"): np.array(d.text.split(), np.float64).astype(
        np.float32 if d.get("type") == "Float32" else np.float64) for d in piece.iter("DataArray")}
    order = []
    for t in sc.tris:
        for v in t:
            if v not in order:
                order.append(v)
    order = np.array(order)
    x = gpu.getState()["x"]
    np.testing.assert_array_equal(arrays["Points"].reshape(-1, 3), x[order].astype(np.float32))
    np.testing.assert_array_equal(arrays["CollsnImpulse"].reshape(-1, 3),
                                  ig[order].astype(np.float32))
    red = (ng[sc.tris] > 0).any(axis=1)
    col = arrays["CollsnRegion"].reshape(-1, 3)
    assert np.array_equal(col[:, 0] == 255, red) and np.array_equal(col[:, 1] == 255, ~red)
    gpu.close()
    orc.close()


# ---------------------------------------------------------------------------
# mixed precision: FP32 proximity screen + certified fast coplanarity screen,
# both exact by construction (every uncertain sub-test runs the FP64 path)
@pytest.mark.parametrize("maker,steps", [
    (lambda: S.cloth_over_sphere(), 64),
    (lambda: S.self_collision_cloth(n=64, steps=12), 12),
    (lambda: S.combined_scene(n=96, nstrings=60, steps=60), 60),
    (S.fan_disc_on_box, 8), (S.deck_string_string, 120), (S.deck_box_boundary, 150),
], ids=["config0_contact", "config1_64", "config4_96", "fan_disc", "string_string", "box_boundary"])
def test_mixed_precision_form_is_bitwise_plain(maker, steps):
    sc = maker()
    if sc.name == "config0_cloth_sphere":
        sc.x = sc.x + sc.vel * (sc.dt * 236)
    st = run_parity(sc, steps=steps, verbose=True, gpu_kw=dict(precision=1, precision_stats=1))
    assert st["hit_pairs"] > 0


def test_mixed_precision_screens_and_reports_errors():
    """The FP32 screen rejects most proximity sub-tests of config 4 (reduced);
    every screened quantity's FP32 error stays within the forward-error bound
    the screen certified it with, and the largest raw errors are reported
    (the line distance of nearly parallel edges can be far off in FP32 --
    with a bound to match, so such a pair is passed on to FP64)."""
    from collision_amd.solver import solver_for_scene
    sc = S.combined_scene(n=96, nstrings=60, steps=20)
    g = solver_for_scene(sc, precision=1, precision_stats=1)
    scr = chk = 0
    ed = ep = er = 0.0
    for _ in range(20):
        r = g.step()
        scr += r.mx_screened
        chk += r.mx_rechecked
        ed, ep = max(ed, r.mx_err_dist), max(ep, r.mx_err_param)
        er = max(er, r.mx_bound_ratio)
    g.close()
    print(f"[mixed] screened {scr} rechecked {chk} err_dist/h {ed:.3e} err_param {ep:.3e} "
          f"error/bound {er:.3e}")
    assert scr > 4 * chk
    assert ed > 0 and ep > 0
    assert 0 < er <= 1.0


def test_config1_full_size_mixed_form_against_golden():
    import json
    import os

    from collision_amd.solver import solver_for_scene
    from tests.golden.make_step_golden import REPORT_KEYS, WORKLOADS, state_hash
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "steps_config1_256.json")
    gold = json.load(open(path))
    sc = WORKLOADS["config1_256"][0]()
    g = solver_for_scene(sc, precision=1)
    try:
        for k, rec in enumerate(gold["records"]):
            r = g.step().as_dict()
            keys = [key for key in REPORT_KEYS if key not in ("candidates", "narrow_pairs")]
            assert {key: int(r[key]) for key in keys} == {key: rec["report"][key] for key in keys}, k
            assert state_hash(g.getState()) == rec["state"], k
    finally:
        g.close()
