"""Batched scenes on the GPU: eight independent cloth-over-sphere scenes in
one mesh, run by one launch per kernel (cs_step_batch), against the CPU
oracle running every scene alone: each scene's state bit for bit, and its
own pass counts, collision flag and hit pairs, every step.  The window holds
steps where one scene's impact-zone loop ends before another's (the zone
updates of a finished scene are gated)."""
import numpy as np
import pytest

from collision_amd import scenes as S
from collision_amd.solver import solver_for_scene
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_batched_scenes_match_separate_oracle_runs():
    sc = S.batched_scenes(count=8, nmin=16, nmax=32, steps=60)
    sc.x = sc.x + sc.vel * (sc.dt * 200)  # start close to contact
    parts = S.scene_parts(sc)
    g = solver_for_scene(sc, log_pairs=0)
    assert g.nscene == 8
    ops = [O.OracleSolver(p, order_mode=O.ORDER_CANONICAL, math_mode=O.MATH_DET) for p in parts]
    gated = zone_steps = 0
    try:
        for k in range(60):
            rep, per = g.stepBatch()
            rs = [o.step() for o in ops]
            print(f"[batch] step {k} passes {rep.n_passes} per scene "
                  f"{[p['n_passes'] for p in per]}", flush=True)
            assert rep.status == 0
            st = g.getState()
            off = 0
            for i, (o, r, p) in enumerate(zip(ops, rs, per)):
                assert (p["n_ccd_passes"], p["n_zone_passes"], p["has_collision"],
                        p["hit_pairs"]) == (r.n_ccd_passes, r.n_zone_passes, r.has_collision,
                                            r.hit_pairs), (k, i, p, r.as_dict())
                sp = o.get_state()
                n = sp["x"].shape[0]
                for key in ("x", "vel", "avg_vel", "x_old", "has_collsn"):
                    assert np.array_equal(st[key][off:off + n], sp[key]), (k, i, key)
                off += n
            nz = [r.n_zone_passes for r in rs if r.n_zone_passes]
            zone_steps += bool(nz)
            gated += bool(nz) and min(nz) != rep.n_zone_passes
            assert rep.n_passes == max(p["n_passes"] for p in per)
    finally:
        g.close()
        for o in ops:
            o.close()
    assert zone_steps > 0 and gated > 0


def test_batched_assembly_rejects_shared_vertices():
    from collision_amd.solver import CollisionError
    sc = S.batched_scenes(count=2, nmin=16, nmax=16, steps=1)
    bad = sc.elem_scene.copy()
    bad[0] = 1  # a cloth triangle of scene 0 claimed by scene 1
    sc.elem_scene = bad
    with pytest.raises(CollisionError):
        solver_for_scene(sc)


@pytest.mark.parametrize("batched", [False, True], ids=["config4_96", "batch6"])
def test_graph_step_equals_eager_step(batched):
    """graph_step = 1: the predictor and the resolve prologue replay one
    captured graph each step; every report and state bit equals the eager
    launches, also across a re-capture (record_impulses switched on midway)
    and with host reads of the state between replays."""
    if batched:
        sc = S.batched_scenes(count=6, nmin=16, nmax=24, steps=40)
        sc.x = sc.x + sc.vel * (sc.dt * 150)
    else:
        sc = S.combined_scene(n=96, nstrings=60, steps=40)
    ga = solver_for_scene(sc, graph_step=1, log_pairs=0)
    gb = solver_for_scene(sc, graph_step=0, log_pairs=0)
    keys = ("status", "n_passes", "n_ccd_passes", "n_zone_passes", "has_collision", "n_zones",
            "strain_sweeps", "hit_pairs", "hit_subtests")
    hits = 0
    try:
        for k in range(40):
            if k == 20:
                for g in (ga, gb):
                    g.params.record_impulses = 1
                    g._push()
            if batched:
                (ra, pa), (rb, pb) = ga.stepBatch(), gb.stepBatch()
                assert pa == pb, k
            else:
                ra, rb = ga.step(), gb.step()
            da, db = ra.as_dict(), rb.as_dict()
            assert {x: da[x] for x in keys} == {x: db[x] for x in keys}, k
            hits += ra.hit_pairs
            sa, sb = ga.getState(), gb.getState()
            for key in ("x", "vel", "avg_vel", "x_old", "has_collsn"):
                assert np.array_equal(sa[key], sb[key]), (k, key)
            if k >= 20:
                ia, ib = ga.lastImpulses(), gb.lastImpulses()
                assert np.array_equal(ia[0], ib[0]) and np.array_equal(ia[1], ib[1]), k
    finally:
        ga.close()
        gb.close()
    assert batched or hits > 0


def test_autotuner_keeps_results_and_caches_its_choice(tmp_path, monkeypatch):
    """autotune = 1: the first 8 steps run candidate launch shapes; the state
    equals an untuned run bit for bit every step, the choice is kept, stored
    in CS_TUNE_CACHE and taken by the next assembly of the same key."""
    from collision_amd.solver import CollisionError
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("CS_TUNE_CACHE", str(cache))
    sc = S.combined_scene(n=96, nstrings=60, steps=30)
    ga = solver_for_scene(sc, autotune=1, log_pairs=0)
    gb = solver_for_scene(sc, log_pairs=0)
    assert ga.tuning()["tuned"] == 0
    try:
        for k in range(30):
            ra, rb = ga.step(), gb.step()
            assert ra.hit_pairs == rb.hit_pairs and ra.n_passes == rb.n_passes, k
            sa, sb = ga.getState(), gb.getState()
            for key in ("x", "vel", "avg_vel", "has_collsn"):
                assert np.array_equal(sa[key], sb[key]), (k, key)
        t = ga.tuning()
        print(f"[tune] {t}", flush=True)
        assert t["tuned"] == 1
        assert cache.exists()
        gc = solver_for_scene(sc, log_pairs=0)
        tc = gc.tuning()
        assert tc["tuned"] == 2 and {k: tc[k] for k in ("traverse_blocks_per_cu", "narrow_threads",
                                                        "query_grid")} == \
            {k: t[k] for k in ("traverse_blocks_per_cu", "narrow_threads", "query_grid")}
        with pytest.raises(CollisionError):
            gc.setTuning(8, 96, 4096)
        gc.close()
    finally:
        ga.close()
        gb.close()
