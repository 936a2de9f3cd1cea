#!/usr/bin/env python3
"""Benchmark of the collision step on one MI355X.

Workload (the headline configuration): a 1024 x 1024 cloth (side 2.0) falling
onto 4 static boxes, with 2,000 strings of 64 bonds above it; thickness
5e-4, dt 5e-4, friction 0.3; the reference driver loop (predictor, then
resolveCollision) for every step.  State stays resident in HBM for the whole
timed region; one "step" is one predictor + resolveCollision call.

Prints one JSON line:
  value         steps/s over the timed steps (after --warmup untimed steps)
  config_*      the whole 100-step configuration (steps 1..100) run again from
                t = 0 in the same process, when the timed window is not it
  pairs_per_s   element pairs through the narrow-phase kernel per second of it
  roofline      the kernel with the most measured time over the timed steps
                (HIP events on the library's own stream around every launch
                of the traversal, the dirty-element queries, the narrow phase
                and the strain sweep); traffic from the committed rocprofv3
                counter summary of that kernel, if there is one
  cpu_baseline  the oracle in the reference's processing order (incremental
                AABB tree, glibc libm), one core, timed on the same steps of
                the same geometry at 256 x 256 cloth + 125 strings (1/16 of
                the elements), scaled by the element count
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
CONFIG_STEPS = 100             # the configuration's step count
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
# algorithmic FLOPs per sample of the 24-128-128-2 network
MLP_FWD_FLOP = 2 * (24 * 128 + 128 * 128 + 128 * 2)                    # 39,424
MLP_TRAIN_FLOP = MLP_FWD_FLOP + 2 * (2 * 128 * 2 + 2 * 128 * 128 + 128 * 24)  # + backward
ROOFLINE_NOTES = {
    "strain": "one work-group runs about 2,900 dependent level iterations per step (the exact "
              "Gauss-Seidel order; a second work-group on its XCD reads the next levels' records into "
              "L2); bound by that chain's latency, not by HBM: it moves about 4.7e8 B "
              "per launch (profiles/traffic_strain.json; mostly the prefetching work-group streaming the "
              "records once) against the 1.6e8 B of reading every vertex and element once",
    "traverse": "per-lane stackless walks, bound by dependent L2 misses; walks longer than 512 steps "
                "are continued over 64 sub-ranges by k_traverse_split (its time is included)",
    "narrow": "gather of each sub-test's four vertices; FP64 VALU co-bound",
}
FORMS = {"plain": {}, "mixed": {"precision": 1}, "learned": {"learned": 1},
         "learned_mixed": {"learned": 1, "precision": 1}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=CONFIG_STEPS)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=1024, help="cloth side vertex count")
    ap.add_argument("--strings", type=int, default=2000)
    ap.add_argument("--cpu-n", type=int, default=256, help="cloth side of the CPU sample")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of timed CPU work (at least one step)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config-run", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the plain / mixed / learned / learned+mixed comparison")
    ap.add_argument("--form", default="plain", choices=sorted(FORMS),
                    help="the form of the timed steps")
    ap.add_argument("--train-steps", type=int, default=10,
                    help="config-1 (256x256) plain steps whose CCD sub-tests train the network")
    ap.add_argument("--train-stride", type=int, default=16)
    ap.add_argument("--train-samples", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-scenes measurement (config 2)")
    ap.add_argument("--batch-scenes", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=170,
                    help="steps of the batch (scene 284 of the seeded configuration reaches the "
                         "reference's clean_up(ERROR), a degenerate proximity contact, at step "
                         "176 -- the oracle alone stops there too)")
    ap.add_argument("--batch-separate-steps", type=int, default=10,
                    help="steps of the one-handle-per-scene comparison")
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


def cpu_baseline(args, ne_full, gpu_ms_by_step):
    """The oracle on the 1/16 sample, on the steps right after the warm-up
    (the GPU's timed window starts there too); the contact timing of the
    geometry does not depend on the cloth resolution."""
    from collision_amd import scenes as S
    from oracle import oracle as O
    ns = max(1, int(round(args.strings * (args.cpu_n / args.n) ** 2)))
    sc = S.combined_scene(n=args.cpu_n, nstrings=ns)
    s = O.OracleSolver(sc, order_mode=O.ORDER_REFERENCE_TREE, math_mode=O.MATH_LIBM)
    for _ in range(args.warmup):
        s.step()
    first = args.warmup + 1
    done, t_cpu = 0, 0.0
    while done < args.steps and (done == 0 or t_cpu < args.cpu_budget):
        t0 = time.perf_counter()
        s.step()
        t_cpu += time.perf_counter() - t0
        done += 1
    s.close()
    ne_s = sc.nt + sc.nb
    scale = ne_full / ne_s
    sps = done / (t_cpu * scale)
    gpu_same = sum(gpu_ms_by_step[:done]) * 1e-3
    return {
        "value": sps,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle in the reference's order (incremental AABB tree, glibc libm, gcc -O3), "
                   f"one core, steps {first}..{first + done - 1} of the same geometry at "
                   f"{args.cpu_n}x{args.cpu_n} cloth + {ns} strings ({ne_s} elements): "
                   f"{t_cpu:.2f} s, scaled linearly by element count x{scale:.2f} (extrapolated; "
                   f"the reference's tree query grows faster than linearly, so this favours "
                   f"the CPU)"),
        "sample_seconds": t_cpu,
        "steps_timed": done,
        "gpu_over_cpu_same_steps": (t_cpu * scale / gpu_same) if gpu_same > 0 else None,
        "host_cpu": platform.processor() or platform.machine(),
        "host_nproc": os.cpu_count(),
    }


def train_network(args):
    """The network of the learned form, trained on the GPU by the library's own
    kernels on the plain CCD's sub-tests of config 1 (256 x 256 self-collision)."""
    from collision_amd import scenes as S
    from collision_amd.solver import solver_for_scene
    c1 = S.self_collision_cloth(n=256, steps=args.train_steps)
    g = solver_for_scene(c1, pass_timing=0)
    g.learnedInit(7)
    g.learnedCollect(capacity=args.train_samples, stride=args.train_stride)
    for _ in range(args.train_steps):
        g.step()
    _, ex, _ = g.learnedDataset()
    n = len(ex)
    npos = int((ex > 0.5).sum())
    # each epoch repeats the samples with a root round(#without / #with) times
    rep = min(64, max(1, (n - npos + npos // 2) // npos)) if npos else 1
    m = n - npos + rep * npos
    g.synchronize()
    t0 = time.perf_counter()
    losses = g.learnedTrain(epochs=args.epochs, batch=8192, lr=1e-3, seed=1)
    g.synchronize()
    tt = time.perf_counter() - t0
    w = g.learnedWeights()
    g.close()
    flops = MLP_TRAIN_FLOP * m * args.epochs
    info = {"samples": n, "samples_with_root": npos, "samples_per_epoch": m,
            "epochs": args.epochs, "batch": 8192, "lr": 1e-3,
            "source": f"config 1 (256x256) plain CCD sub-tests, steps 1..{args.train_steps}, "
                      f"1 in {args.train_stride} kept",
            "epoch_losses": [float(v) for v in losses], "seconds": tt,
            "samples_per_s": m * args.epochs / tt if tt > 0 else None,
            "tflops": flops / tt / 1e12 if tt > 0 else None,
            "mfma_frac": flops / tt / 1e12 / MFMA_BF16_PEAK_TFLOPS if tt > 0 else None,
            "flop_note": f"{MLP_TRAIN_FLOP} algorithmic FLOP per sample (forward + backward), "
                         "wall time of the whole training call (shuffles, reductions, Adam)"}
    return w, info


def run_batch(args):
    """Config 2: 512 uneven cloth-over-sphere scenes in one mesh, one launch
    per kernel for all of them (cs_step_batch), against one solver handle
    per scene stepped in turn, on the same first steps."""
    from collision_amd import scenes as S
    from collision_amd.solver import solver_for_scene
    sc = S.batched_scenes(count=args.batch_scenes, steps=args.batch_steps)
    parts = S.scene_parts(sc)
    g = solver_for_scene(sc, pass_timing=0, log_pairs=0)
    g.synchronize()
    k0 = args.batch_separate_steps
    t0 = time.perf_counter()
    passes = syncs = 0
    t_first = None
    for k in range(args.batch_steps):
        rep, _ = g.stepBatch()
        passes += rep.n_passes
        syncs += rep.host_syncs
        if k + 1 == k0:
            g.synchronize()
            t_first = time.perf_counter() - t0
    g.synchronize()
    el = time.perf_counter() - t0
    g.close()
    hs = [solver_for_scene(p, pass_timing=0, log_pairs=0) for p in parts]
    for h in hs:
        h.synchronize()
    t1 = time.perf_counter()
    sep_syncs = 0
    for _ in range(k0):
        for h in hs:
            sep_syncs += h.step().host_syncs
    for h in hs:
        h.synchronize()
    sep = time.perf_counter() - t1
    for h in hs:
        h.close()
    res = {"scenes": args.batch_scenes, "vertices": sc.nv, "triangles": sc.nt,
           "steps": args.batch_steps, "steps_per_s": args.batch_steps / el,
           "ms_per_step": 1e3 * el / args.batch_steps,
           "passes_per_step": passes / args.batch_steps,
           "host_syncs_per_step": syncs / args.batch_steps,
           "first_steps": k0,
           "first_steps_per_s": k0 / t_first if t_first else None,
           "separate_handles_first_steps_per_s": k0 / sep,
           "separate_host_syncs_per_step": sep_syncs / k0,
           "batched_over_separate": (sep / t_first) if t_first else None,
           "note": "batched: one mesh, one launch per kernel for every scene; separate: one "
                   "solver handle per scene, stepped in turn; both on steps 1.."
                   f"{k0} (no contact yet in that window)"}
    print(f"[bench] batch {res['steps_per_s']:.2f} steps/s; first {k0}: batched "
          f"{res['first_steps_per_s']:.2f} vs separate {res['separate_handles_first_steps_per_s']:.3f}",
          file=sys.stderr, flush=True)
    return res


def state_digest(solver):
    import hashlib
    st = solver.getState()
    h = hashlib.sha256()
    for k in ("x", "vel", "avg_vel", "x_old", "has_collsn"):
        h.update(np.ascontiguousarray(st[k]).tobytes())
    return h.hexdigest()[:16]


def run_forms(sc, weights):
    """Every form over the whole configuration (steps 1..100) from t = 0, no
    timing points; the state after the last step must be the same in all."""
    from collision_amd.solver import solver_for_scene
    res = {}
    for name, kw in FORMS.items():
        s = solver_for_scene(sc, pass_timing=0, **kw)
        if kw.get("learned"):
            s.setLearnedWeights(weights)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(CONFIG_STEPS):
            s.step()
        s.synchronize()
        el = time.perf_counter() - t0
        r = {"steps_per_s": CONFIG_STEPS / el, "ms_per_step": 1e3 * el / CONFIG_STEPS,
             "state_digest": state_digest(s)}
        if kw.get("learned"):
            ls = s.learnedStats()
            roots = max(1, ls["newton_roots"])
            tp = (ls["predicted_exists"] + ls["exists"] - ls["scored"] + ls["agree"]) // 2
            r["learned"] = dict(ls, accuracy=ls["agree"] / max(1, ls["scored"]),
                                recall=tp / max(1, ls["exists"]),
                                precision=tp / max(1, ls["predicted_exists"]),
                                newton_iters_per_root=ls["newton_iters"] / roots,
                                midpoint_iters_per_root=ls["newton_iters_mid"] / roots)
        s.close()
        res[name] = r
        print(f"[bench] form {name}: {r['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    base = res["plain"]["steps_per_s"]
    for r in res.values():
        r["speedup_vs_plain"] = r["steps_per_s"] / base
    return res


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
    nv, nt, nb = sc.nv, sc.nt, sc.nb
    ne = nt + nb
    # timed steps keep only the kernel timing points (traversal, query, narrow
    # phase, strain sweep); the phase breakdown comes from a few extra steps
    weights, training = None, None
    if not args.no_forms or "learned" in args.form:
        weights, training = train_network(args)
        print(f"[bench] trained on {training['samples']} sub-tests in {training['seconds']:.2f} s",
              file=sys.stderr, flush=True)
    solver = solver_for_scene(sc, pass_timing=1, **FORMS[args.form])
    if "learned" in args.form:
        solver.setLearnedWeights(weights)
    print(f"[bench] scene nv={nv} nt={nt} nb={nb}; warmup {args.warmup}", file=sys.stderr,
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

    forms = None if args.no_forms else run_forms(sc, weights)
    batch = None if args.no_batch else run_batch(args)
    config = None
    if forms is not None and args.form in forms:
        config = {"steps_per_s": forms[args.form]["steps_per_s"],
                  "ms_per_step": forms[args.form]["ms_per_step"], "steps": f"1..{CONFIG_STEPS}"}
    elif not args.no_config_run and (args.warmup, args.steps) != (0, CONFIG_STEPS):
        cs = solver_for_scene(sc, pass_timing=0, **FORMS[args.form])
        if "learned" in args.form:
            cs.setLearnedWeights(weights)
        cs.synchronize()
        tc = time.perf_counter()
        for k in range(CONFIG_STEPS):
            cs.step()
        cs.synchronize()
        tcw = time.perf_counter() - tc
        cs.close()
        config = {"steps_per_s": CONFIG_STEPS / tcw, "ms_per_step": 1e3 * tcw / CONFIG_STEPS,
                  "steps": f"1..{CONFIG_STEPS}"}
        print(f"[bench] config run {CONFIG_STEPS} steps {tcw:.2f} s", file=sys.stderr, flush=True)

    full = [p for p in passes if not p["incremental"]]
    incr = [p for p in passes if p["incremental"]]
    tested = sum(p["tested_pairs"] for p in passes)
    t_narrow_all = sum(p["ms_narrow_kernel"] for p in passes) * 1e-3

    def avg(lst, key):
        return sum(p[key] for p in lst) / max(len(lst), 1)

    def narrow_bytes(p):
        # each distinct vertex (x_old, avgVel: 48 B) and element (ids: 12 B tri, 8 B bond)
        # once, but no more than the pairs tested can touch (6 vertices, 2 elements each)
        P = p["tested_pairs"]
        return (8 * P + min(48 * nv, 288 * P) + min(12 * nt + 8 * nb, 24 * P)
                + HIT_RECORD_BYTES * p["hit_subtests"])

    # Algorithmic (compulsory) bytes per launch, s = 8 B FP64 state:
    #   traverse : 6 s Nv + 12 Nt + 8 Nb + 8 C           (C = box-overlap candidates)
    #   query    : not stated (the dirty-element counts are not reported); time only
    #   narrow   : 8 P + min(48 Nv, 288 P) + min(12 Nt + 8 Nb, 24 P) + R H per launch
    #              (P pairs evaluated, H hit sub-tests), averaged over the launches
    #   strain   : 72 Nv + 36 Nt + 16 Nb                  (x_old and avgVel of every vertex read,
    #              avgVel written, the rest lengths and ids of every element read once)
    b_broad = 48 * nv + 12 * nt + 8 * nb + 8 * avg(full, "candidates")
    kern = {
        "traverse": dict(launches=len(full), total=sum(p["ms_traverse_kernel"] for p in full),
                         bytes=b_broad),
        "query": dict(launches=len(incr), total=sum(p["ms_traverse_kernel"] for p in incr),
                      bytes=None),
        "narrow": dict(launches=len(passes), total=sum(p["ms_narrow_kernel"] for p in passes),
                       bytes=sum(narrow_bytes(p) for p in passes) / max(len(passes), 1)),
        "strain": dict(launches=args.steps, total=sum(r["ms_strain_kernel"] for r in reports),
                       bytes=72 * nv + 36 * nt + 16 * nb),
    }
    for v in kern.values():
        v["avg_launch_ms"] = v["total"] / max(v["launches"], 1)
        t = v["avg_launch_ms"] * 1e-3
        v["achieved"] = v["bytes"] / t / 1e9 if (t > 0 and v["bytes"]) else None
        v["ms_per_step"] = v["total"] / args.steps
    # the roofline names the kernel with the most time among those with stated bytes
    dom_name = max((k for k in kern if kern[k]["bytes"]), key=lambda k: kern[k]["total"])
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
            "form": args.form,
        },
        "config_run": config,
        "forms": forms,
        "forms_identical_state": (len({f["state_digest"] for f in forms.values()}) == 1
                                  if forms else None),
        "training": training,
        "batch": batch,
        "pairs_per_s": tested / t_narrow_all if t_narrow_all > 0 else None,
        "pairs_per_s_note": "element pairs evaluated by the narrow-phase kernel per second of "
                            "its run time (full and incremental passes)",
        "passes_per_step": len(passes) / args.steps,
        "host_syncs_per_step": sum(r["host_syncs"] for r in reports) / args.steps,
        "full_passes_per_step": len(full) / args.steps,
        "candidates_per_full_pass": avg(full, "candidates"),
        "narrow_pairs_per_full_pass": avg(full, "narrow_pairs"),
        "tested_pairs_per_incremental_pass": avg(incr, "tested_pairs"),
        "strain_sweeps_per_step": sum(r["strain_sweeps"] for r in reports) / args.steps,
        "strain_levels_per_step": sum(r["strain_levels"] for r in reports) / args.steps,
        "phase_ms_per_step": {
            k: sum(r[k] for r in phase_reports) / phase_steps
            for k in ("ms_total", "ms_broad", "ms_narrow", "ms_response", "ms_zone", "ms_strain")
        },
        "phase_ms_note": f"steps {args.warmup + args.steps + 1}.."
                         f"{args.warmup + args.steps + phase_steps} with every phase boundary timed",
        "kernels": {k: {"ms_per_step": v["ms_per_step"], "avg_launch_ms": v["avg_launch_ms"],
                        "launches_per_step": v["launches"] / args.steps,
                        "algorithmic_bytes_per_launch": v["bytes"],
                        "achieved_GBs": v["achieved"],
                        "frac": v["achieved"] / HBM_PEAK_GBS if v["achieved"] else None}
                    for k, v in kern.items()},
        "roofline": {
            "kernel": dom_name,
            "bound": "hbm",
            "achieved": dom["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["achieved"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom["bytes"],
            "avg_launch_ms": dom["avg_launch_ms"],
            "note": ROOFLINE_NOTES.get(dom_name),
        },
    }
    if not args.no_cpu:
        cb = cpu_baseline(args, ne, [r["ms_total"] for r in reports])
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = steps_per_s / cb["value"] if cb["value"] > 0 else None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
