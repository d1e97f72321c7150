"""Headline benchmark: particle-steps/sec of the Barnes-Hut substep loop at N = 1,048,576.

Workload (default): 2^20 particles with the relativistic dynamics of the
N=65536 configuration scaled to keep its density (1024 x 1024 box): gravity +
Coulomb + Yukawa (mu = 1) + magnetic dipole, softening 8, spins uniform in
[-1, 1], plain monopole BH at theta = 0.3.  The box is open (despawn) by
default, where BH converges to the FP64 direct sum the force error is judged
against; --boundary torus runs the min-image torus of the reference instead.  One step = one
Physics.update(dt = 1/128) (adaptive substeps); particle-steps = sum over
substeps of live particles.  Inputs are resident in HBM before timing.

Also reported: the force walk's roofline (FP32 vector FLOP/s from the kernel's
own interaction counters x flops per interaction, over its HIP-event duration),
the global relative L2 force error of the GPU forces against an FP64 direct sum
on a subsample, and the reference CPU algorithm (the FP64 oracle) timed on this
host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP32_PEAK_TFLOPS = 157.3


def flops_per_interaction(toggles):
    """Flops of one pair_force call by the GPU-Gems rule, for the feature set
    (counted from physsim_amd/csrc/geon_dev.h; table in physsim_amd/flops.json)."""
    with open(os.path.join(ROOT, "physsim_amd", "flops.json")) as f:
        tab = json.load(f)
    from physsim_amd import _abi as A
    c = tab["base"]
    if toggles & A.GRAVITY:
        c += tab["gravity"] + tab["tidal"]
    if toggles & A.COULOMB:
        c += tab["coulomb"] + (tab["tidal_coulomb"] if toggles & A.GRAVITY else 0)
    if toggles & (A.MAGNETIC | A.GRAVITOMAG | A.YUKAWA):
        c += tab["extra_base"]
    if toggles & A.MAGNETIC:
        c += tab["magnetic"]
    if toggles & A.YUKAWA:
        c += tab["yukawa"]
    return c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--err-sample", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--boundary", choices=("open", "torus"), default="open")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the Barnes-Hut path here is single-GPU "
              "(see DESIGN.md for how it would shard); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from physsim_amd import HipPhysics, default_params, workloads as W
    st, prm, dt = W.config2(args.n, open_box=(args.boundary == "open"))
    prm["theta"] = args.theta
    n = args.n
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    print("bench: n=%d theta=%.2f warmup=%d steps=%d" % (n, args.theta, args.warmup, args.steps), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        g.update(dt)
    g.timing(True)
    alive0 = g.getParticleCount()
    t0 = time.perf_counter()
    substeps = 0
    for k in range(args.steps):
        s = g.update(dt)
        substeps += s
        print("bench: step %d substeps %d" % (k, s), file=sys.stderr, flush=True)
    g.getParticleCount()       # synchronizes the stream
    t1 = time.perf_counter()
    wall = t1 - t0
    alive1 = g.getParticleCount()
    psteps = substeps * alive1  # torus: no particle leaves
    value = psteps / wall
    walk_ms, walk_n = g.kernel_timing(0)
    tree_ms, tree_n = g.kernel_timing(1)
    sort_ms, sort_n = g.kernel_timing(2)
    kick_ms, kick_n = g.kernel_timing(3)
    g.timing(False)
    cnt = g.counters()
    inter = cnt["pair_interactions"] + cnt["node_interactions"]
    c_f = flops_per_interaction(prm["toggles0"])
    walk_avg_s = walk_ms / max(walk_n, 1) / 1e3
    achieved_tflops = inter * c_f / walk_avg_s / 1e12 if walk_avg_s > 0 else 0.0

    # ---- force error of the GPU forces vs the FP64 direct sum (checker) ----
    result_err = None
    cpu = None
    if not args.no_cpu:
        from oracle import bind
        threads = min(16, os.cpu_count() or 1)
        bind.set_threads(threads)
        d = g.download()
        g.compute_forces()
        fg = g.forces(n)
        orc = bind.Oracle(bind.default_params(**prm), n)
        orc.upload(d["x"], d["y"], d["wx"], d["wy"], d["mass"], d["charge"], d["angw"])
        rs = np.random.Generator(np.random.Philox(5000)).choice(n, args.err_sample, replace=False)
        rs = np.sort(rs).astype(np.int32)
        te = time.perf_counter()
        fx, fy = orc.direct_forces(rs)
        te = time.perf_counter() - te
        num = np.sum((fg["fx"][rs] - fx) ** 2 + (fg["fy"][rs] - fy) ** 2)
        den = np.sum(fx ** 2 + fy ** 2)
        rel = np.sqrt(((fg["fx"][rs] - fx) ** 2 + (fg["fy"][rs] - fy) ** 2) / np.maximum(fx ** 2 + fy ** 2, 1e-300))
        result_err = {"rel_l2": float(np.sqrt(num / den)), "median": float(np.median(rel)),
                      "p99": float(np.percentile(rel, 99)), "sample": int(args.err_sample),
                      "reference": "FP64 direct sum (forces.js:93-126, min-image)", "seconds": round(te, 2)}
        del orc
        # ---- CPU baseline: the reference algorithm in FP64 on this host ----
        cs, cprm, _ = W.config2(args.cpu_sample, open_box=(args.boundary == "open"))
        cprm["theta"] = args.theta
        oc = bind.Oracle(bind.default_params(**cprm), args.cpu_sample)
        oc.upload(cs["x"], cs["y"], cs["wx"], cs["wy"], cs["mass"], cs["charge"], cs["angw"])
        tc = time.perf_counter()
        oc.compute_forces()                     # bootstrap force pass (part of the first update)
        tb = time.perf_counter()
        steps_c = oc.update(dt)
        tc2 = time.perf_counter()
        cpu = {"value": steps_c * oc.count() / (tc2 - tb), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": "FP64 oracle (integrator.js/forces.js/quadtree.js restated in C, OpenMP over "
                         "particles) on the same dynamics at N=%d (same density), one update of %d substeps, "
                         "%.1f s; bootstrap force pass %.1f s untimed" % (args.cpu_sample, steps_c, tc2 - tb, tb - tc)}

    line = {
        "metric": "particle-steps/sec at N=1M, plain BH (monopole, theta=%.2f)" % args.theta,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox seed 2000: uniform torus box, lognormal masses, 0.1-quantized charges)",
        "config": {"workload": "%s 1024x1024 box, N=%d uniform, gravity+Coulomb+Yukawa(mu=1)+magnetic dipole, "
                               "relativistic, spins U[-1,1], softening 8, theta %.2f, dt 1/128 per update"
                               % ("open (despawn)" if args.boundary == "open" else "torus min-image", n, args.theta),
                   "n": n, "theta": args.theta, "substeps": substeps},
        "roofline": {"bound": "valu", "kernel": "k_walk", "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / FP32_PEAK_TFLOPS, "traffic": None,
                     "interactions_per_launch": inter, "flops_per_interaction": c_f,
                     "walk_ms_avg": walk_avg_s * 1e3},
        "phases_ms_avg": {"walk": walk_ms / max(walk_n, 1), "tree": tree_ms / max(tree_n, 1),
                          "sort_gather": sort_ms / max(sort_n, 1), "kick_drift": kick_ms / max(kick_n, 1)},
        "force_error": result_err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
