#!/usr/bin/env python3
"""Benchmark of the collision step on one MI355X.

Workload (the headline configuration): a 1024 x 1024 cloth (side 2.0) falling
onto 4 static boxes, with 2,000 strings of 64 bonds above it; thickness
5e-4, dt 5e-4, friction 0.3; the reference driver loop (predictor, then
resolveCollision) for every step.  State stays resident in HBM for the whole
timed region; one "step" is one predictor + resolveCollision call.

Prints one JSON line with steps/s (value), narrow-phase pairs/s, the
roofline of the dominant kernel (per-launch HIP-event durations from the
library's own stream), and a CPU baseline: the oracle in the reference's
processing order (incremental AABB tree, libm), single core, on a bounded
sample of the same geometry, scaled to the full element count.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
HIT_RECORD_BYTES = 64          # R: compulsory bytes per contact record


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024, help="cloth side vertex count")
    ap.add_argument("--strings", type=int, default=2000)
    ap.add_argument("--cpu-n", type=int, default=256, help="cloth side of the CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three 256 MiB device copies (scale for HBM byte counters)")
    return ap.parse_args()


def copy_probe():
    """Three plain device copies of 256 MiB: the reference for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the collision kernels to bytes."""
    import torch
    n = 256 * 1024 * 1024 // 4
    x = torch.arange(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        y = x.clone()
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copies": 3, "last": float(y[-1])}))


def cpu_baseline(args, ne_full):
    from collision_amd import scenes as S
    from oracle import oracle as O
    ns = max(1, int(round(args.strings * (args.cpu_n / args.n) ** 2)))
    sc = S.combined_scene(n=args.cpu_n, nstrings=ns)
    s = O.OracleSolver(sc, order_mode=O.ORDER_REFERENCE_TREE, math_mode=O.MATH_LIBM)
    s.step()  # first step builds both AABB trees (not timed)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        s.step()
    dt = time.perf_counter() - t0
    s.close()
    ne_s = sc.nt + sc.nb
    sps_sample = args.cpu_steps / dt
    return {
        "value": sps_sample * ne_s / ne_full,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle in the reference's order (incremental AABB tree, glibc libm), one "
                   f"core, {args.cpu_steps} steps after a tree-building step on the same "
                   f"geometry at {args.cpu_n}x{args.cpu_n} cloth + {ns} strings "
                   f"({ne_s} elements, {sps_sample:.4f} steps/s), scaled linearly by element "
                   f"count to {ne_full}; linear scaling favours the CPU (its tree query "
                   f"grows faster than linearly)"),
        "sample_seconds": dt,
        "host_cpu": platform.processor() or platform.machine(),
        "host_nproc": os.cpu_count(),
    }


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the collision step "
              f"of one scene is not sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    from collision_amd import scenes as S
    from collision_amd.solver import solver_for_scene

    sc = S.combined_scene(n=args.n, nstrings=args.strings)
    ne = sc.nt + sc.nb
    # timed steps keep only the kernel timing points (traversal and narrow
    # phase, for the roofline and pairs/s); the phase breakdown is taken from
    # a few extra steps afterwards
    solver = solver_for_scene(sc, pass_timing=1)
    print(f"[bench] scene nv={sc.nv} nt={sc.nt} nb={sc.nb}; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        solver.step()
    solver.synchronize()
    reports, passes = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = solver.step()
        reports.append(r.as_dict())
        passes.extend(solver.passes())
        if (k + 1) % 20 == 0:
            print(f"[bench] step {k + 1}/{args.steps}", file=sys.stderr, flush=True)
    solver.synchronize()
    wall = time.perf_counter() - t0
    solver.setPassTiming(2)
    phase_steps = 3
    phase_reports = [solver.step().as_dict() for _ in range(phase_steps)]
    solver.close()

    steps_per_s = args.steps / wall
    nv, nt, nb = sc.nv, sc.nt, sc.nb
    full = [p for p in passes if not p["incremental"]]
    incr = [p for p in passes if p["incremental"]]
    tested = sum(p["tested_pairs"] for p in passes)
    t_narrow_all = sum(p["ms_narrow_kernel"] for p in passes) * 1e-3
    # algorithmic bytes per launch (compulsory HBM traffic, s = 8 B FP64 state):
    #   traversal (broad) : 6 s Nv + 12 Nt + 8 Nb + 8 C   (C = box-overlap candidates)
    #   narrow            : 8 P + 48 Nv + 12 Nt + 8 Nb + R H
    #                       (P = pairs evaluated, H = hit sub-tests, R = 64 B)
    def avg(lst, key):
        return sum(p[key] for p in lst) / max(len(lst), 1)
    kern = {
        "traverse": dict(t=avg(full, "ms_traverse_kernel") * 1e-3,
                         bytes=48 * nv + 12 * nt + 8 * nb + 8 * avg(full, "candidates"),
                         total=sum(p["ms_traverse_kernel"] for p in full) * 1e-3),
        "narrow": dict(t=avg(full, "ms_narrow_kernel") * 1e-3,
                       bytes=8 * avg(full, "tested_pairs") + 48 * nv + 12 * nt + 8 * nb
                       + HIT_RECORD_BYTES * avg(full, "hit_subtests"),
                       total=sum(p["ms_narrow_kernel"] for p in full) * 1e-3),
    }
    for k in kern.values():
        k["achieved"] = k["bytes"] / k["t"] / 1e9 if k["t"] > 0 else 0.0
    dom_name = max(kern, key=lambda k: kern[k]["total"])
    dom = kern[dom_name]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom_name}.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": "collision steps/s + narrow-phase pairs/s, 1M-vert cloth; learned vs plain",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"cloth {args.n}x{args.n} (side 2.0) over 4 static boxes + "
                         f"{args.strings} strings x 64 bonds; thickness 5e-4, dt 5e-4, "
                         f"friction 0.3; steps {args.warmup + 1}..{args.warmup + args.steps}"),
            "vertices": nv, "triangles": nt, "bonds": nb,
            "form": "plain FP64 (bitwise equal to the oracle)",
        },
        "pairs_per_s": tested / t_narrow_all if t_narrow_all > 0 else None,
        "pairs_per_s_note": "element pairs evaluated by the narrow-phase kernel per second of "
                            "its run time (full and incremental passes)",
        "passes_per_step": len(passes) / args.steps,
        "full_passes_per_step": len(full) / args.steps,
        "candidates_per_full_pass": avg(full, "candidates"),
        "narrow_pairs_per_full_pass": avg(full, "narrow_pairs"),
        "tested_pairs_per_incremental_pass": avg(incr, "tested_pairs"),
        "hit_subtests_per_incremental_pass": avg(incr, "hit_subtests"),
        "strain_sweeps_per_step": sum(r["strain_sweeps"] for r in reports) / args.steps,
        "strain_levels_per_step": sum(r["strain_levels"] for r in reports) / args.steps,
        "strain_slots_per_step": sum(r["strain_slots"] for r in reports) / args.steps,
        "phase_ms_per_step": {
            k: sum(r[k] for r in phase_reports) / phase_steps
            for k in ("ms_total", "ms_broad", "ms_narrow", "ms_response", "ms_zone", "ms_strain")
        },
        "phase_ms_note": f"steps {args.warmup + args.steps}..{args.warmup + args.steps + phase_steps - 1} "
                         "(after the timed steps) with every phase boundary timed",
        "kernel_ms_per_launch_full_pass": {k: 1e3 * v["t"] for k, v in kern.items()},
        "broad_ms_per_incremental_pass": avg(incr, "ms_traverse_kernel"),
        "narrow_ms_per_incremental_pass": avg(incr, "ms_narrow_kernel"),
        "roofline": {
            "kernel": dom_name,
            "bound": "hbm",
            "achieved": dom["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["achieved"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom["bytes"],
            "avg_launch_ms": 1e3 * dom["t"],
        },
        "roofline_other": {k: {"achieved_GBs": v["achieved"], "frac": v["achieved"] / HBM_PEAK_GBS,
                               "bytes_per_launch": v["bytes"], "avg_launch_ms": 1e3 * v["t"]}
                           for k, v in kern.items() if k != dom_name},
    }
    if not args.no_cpu:
        cb = cpu_baseline(args, ne)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = steps_per_s / cb["value"] if cb["value"] > 0 else None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
