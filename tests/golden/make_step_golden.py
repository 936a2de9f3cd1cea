"""Writes golden step records of the oracle (canonical order, deterministic
math) for workloads too large to run the oracle beside the GPU in a test:

  config1_256   config 1 at full size (256 x 256 cloth), 100 steps
  config4_256   config-4 geometry at 256 x 256 cloth + 125 strings, 60 steps

Each step record holds the report counts, the per-pass counts, a SHA-256 of
the sorted hit-pair log and a SHA-256 of the complete vertex state (x, vel,
avgVel, x_old as little-endian float64, has_collsn as bytes).  The oracle is
the checker; this script only runs in the development container.

    python tests/golden/make_step_golden.py config1_256 [config4_256 ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from collision_amd import scenes as S  # noqa: E402

WORKLOADS = {
    "config1_256": (lambda: S.self_collision_cloth(n=256), 100),
    "config4_256": (lambda: S.combined_scene(n=256, nstrings=125), 60),
}
REPORT_KEYS = ("status", "n_passes", "n_ccd_passes", "n_zone_passes", "has_collision", "n_zones",
               "strain_sweeps", "strain_edges", "degenerate_edges", "candidates", "narrow_pairs",
               "hit_pairs", "hit_subtests")
PASS_KEYS = ("kind", "index", "candidates", "narrow_pairs", "hit_pairs", "hit_subtests")


def state_hash(st) -> str:
    h = hashlib.sha256()
    for k in ("x", "vel", "avg_vel", "x_old"):
        h.update(np.ascontiguousarray(st[k], "<f8").tobytes())
    h.update(np.ascontiguousarray(st["has_collsn"], np.uint8).tobytes())
    return h.hexdigest()


def log_hash(log) -> str:
    log = np.asarray(log, np.int32).reshape(-1, 4)
    if len(log):
        log = log[np.lexsort((log[:, 3], log[:, 2], log[:, 1], log[:, 0]))]
    return hashlib.sha256(np.ascontiguousarray(log, "<i4").tobytes()).hexdigest()


def main(names):
    from oracle import oracle as O
    for name in names:
        make, steps = WORKLOADS[name]
        sc = make()
        s = O.OracleSolver(sc, log_pairs=1, order_mode=O.ORDER_CANONICAL, math_mode=O.MATH_DET)
        recs = []
        t0 = time.time()
        for k in range(steps):
            r = s.step().as_dict()
            recs.append(dict(report={key: int(r[key]) for key in REPORT_KEYS},
                             passes=[{key: p[key] for key in PASS_KEYS} for p in s.passes()],
                             log=log_hash(s.pair_log()), state=state_hash(s.get_state())))
            print(f"{name} step {k} passes {r['n_passes']} hits {r['hit_pairs']} "
                  f"{time.time() - t0:.0f}s", flush=True)
        s.close()
        out = dict(workload=name, steps=steps, nv=sc.nv, nt=sc.nt, nb=sc.nb, records=recs)
        with open(os.path.join(HERE, f"steps_{name}.json"), "w") as f:
            json.dump(out, f, indent=0)


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
