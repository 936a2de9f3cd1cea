"""Step-by-step comparison of the HIP library against the CPU oracle.

Both run the reference driver loop (predict, resolve) from the same scene;
after every step the per-pass counts, the hit-pair log, and the full state
(Coords, vel, avgVel, x_old, has_collsn) are compared.  In the canonical
order with the deterministic math mode the two are bitwise identical.
"""
from __future__ import annotations

import numpy as np

from collision_amd.solver import solver_for_scene
from oracle import oracle as O

PASS_KEYS = ("kind", "index", "candidates", "narrow_pairs", "hit_pairs", "hit_subtests")
REPORT_KEYS = ("n_passes", "n_ccd_passes", "n_zone_passes", "has_collision", "n_zones",
               "strain_sweeps", "strain_edges", "degenerate_edges", "candidates",
               "narrow_pairs", "hit_pairs", "hit_subtests")


def sorted_log(log):
    if len(log) == 0:
        return log
    order = np.lexsort((log[:, 3], log[:, 2], log[:, 1], log[:, 0]))
    return log[order]


def run_parity(scene, steps, check_every=1, compare_zones=True, verbose=False, gpu_kw=None,
               gpu_log_pairs=1, gpu_setup=None, **kw):
    """gpu_log_pairs=0 runs the library as shipped (no parity taps: the pair
    stream is not carried between incremental passes and their candidate and
    pair counts are -1); the hit counts, zones and the full state are still
    compared every step (check_every only applies with the taps on).
    gpu_setup(solver) runs once after the library's solver is made."""
    gpu = solver_for_scene(scene, log_pairs=gpu_log_pairs, **(gpu_kw or {}), **kw)
    if gpu_setup is not None:
        gpu_setup(gpu)
    orc = O.OracleSolver(scene, log_pairs=1, order_mode=O.ORDER_CANONICAL,
                         math_mode=O.MATH_DET, **kw)
    stats = dict(steps=0, hit_pairs=0, zone_steps=0, strain_sweeps=0)
    try:
        for step in range(steps):
            rg = gpu.step()
            ro = orc.step()
            print(f"[parity] {scene.name} step {step} gpu_status={rg.status} "
                  f"orc_status={ro.status} passes={rg.n_passes} hits={rg.hit_pairs}",
                  flush=True) if verbose else None
            assert rg.status == ro.status, (step, rg.status, ro.status)
            dg, do = rg.as_dict(), ro.as_dict()
            taps = gpu_log_pairs != 0
            for k in REPORT_KEYS:
                if not taps and k in ("candidates", "narrow_pairs") and dg[k] == -1:
                    continue
                assert dg[k] == do[k], (step, k, dg[k], do[k])
            pg, po = gpu.passes(), orc.passes()
            assert len(pg) == len(po), (step, len(pg), len(po))
            for a, b in zip(pg, po):
                for k in PASS_KEYS:
                    if not taps and k in ("candidates", "narrow_pairs") and a[k] == -1:
                        continue
                    assert a[k] == b[k], (step, k, a, b)
            if taps:
                lg, lo = sorted_log(gpu.pairLog()), sorted_log(orc.pair_log())
                assert lg.shape == lo.shape, (step, lg.shape, lo.shape)
                assert np.array_equal(lg, lo), (step, "pair log differs")
            if compare_zones and ro.n_zone_passes:
                assert np.array_equal(gpu.zoneLabels(), orc.zone_labels()), (step, "zones")
                stats["zone_steps"] += 1
            if not taps or (step + 1) % check_every == 0 or step == steps - 1:
                if getattr(scene, "nbody", 0):
                    (cg, vg), (co, vo) = gpu.getBodies(), orc.bodies()
                    assert np.array_equal(cg, co) and np.array_equal(vg, vo), (step, "bodies")
                sg, so = gpu.getState(), orc.get_state()
                for k in ("x", "vel", "avg_vel", "x_old", "has_collsn"):
                    if not np.array_equal(sg[k], so[k]):
                        d = np.abs(sg[k].astype(np.float64) - so[k].astype(np.float64))
                        raise AssertionError(f"step {step}: {k} differs, max abs {d.max():.3e}, "
                                             f"count {(d > 0).sum()}")
            stats["steps"] += 1
            stats["hit_pairs"] += ro.hit_pairs
            stats["strain_sweeps"] += ro.strain_sweeps
    finally:
        gpu.close()
        orc.close()
    return stats
