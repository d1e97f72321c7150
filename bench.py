"""Headline benchmark: particle-steps/sec at N = 1,048,576, plain vs
learned-correction Barnes-Hut at force error <= 1e-4.

Workload: 2^20 particles with the relativistic dynamics of the N=65536
configuration scaled to keep its density (1024 x 1024 open box): gravity +
Coulomb + Yukawa (mu = 1) + magnetic dipole, softening 8,
spins uniform in [-1, 1].  The box is open (despawn), where Barnes-Hut
converges to the FP64 direct sum the force error is judged against.  One step
= one Physics.update(dt = 1/128) (adaptive substeps); particle-steps = sum over
substeps of live particles.  Inputs are resident in HBM before timing.

Three force methods run the same updates from the same initial state:
  plain      monopole nodes (the reference's walk) at --theta-plain, the
             coarsest opening angle whose error is <= 1e-4;
  learned    the same walk at --theta-learned plus the 2 x 64 hard-swish network's
             residual for every accepted node; the network is trained here,
             on the GPU, on exact residuals of 32 Plummer spheres of 65536
             (untimed setup; time and sample count reported);
  multipole  the same walk at --theta-multipole plus the analytic
             quadrupole / dipole terms (reported beside the two).
The global relative L2 force error of each against the FP64 direct sum on a
4096-particle subsample is measured in the same run; `value` is the learned
method's throughput.

Also reported: the dominant kernel's roofline (HIP events around its
launches; algorithmic flops from the kernel's own counters), the walk's, and
the reference algorithm (the FP64 oracle) timed on this host on a bounded
sample.
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
# without packed-FP32 instructions (the walks are built without them): one FMA per lane per cycle
FP32_NONPACKED_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
# network flops per (particle, node) pair: 2 (11 x 64 + 64 x 64 + 64 x 6)
MLP_FLOPS_PER_PAIR = 2 * (11 * 64 + 64 * 64 + 64 * 6)
# HBM bytes of one k_far_mlp launch at the default workload (N=2^20, theta 0.30), from the
# FETCH_SIZE / WRITE_SIZE counter passes scaled by a 256 MiB copy (profiles/k_far_mlp_open1M_theta0.30.md)
MLP_TRAFFIC_DEFAULT = 2696.0e6
TIMER_WALK, TIMER_TREE, TIMER_SORT, TIMER_KICK, TIMER_MLP = 0, 1, 2, 3, 4
TIMER_DEPOSIT, TIMER_KDK, TIMER_FIELD_PART = 5, 6, 7
HBM_PEAK_GBS = 8000.0


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


def host_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    limited by OMP_NUM_THREADS where that is set (the host share of one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_available": avail,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "threads_used": host_threads()}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def train_far_model(theta, seeds, n, stride, steps, batch):
    """The learned correction's network, trained on the GPU from the exact
    residuals of `seeds` Plummer spheres of n particles."""
    from physsim_amd import HipPhysics, default_params, workloads as W
    _, tprm = W.plummer_charged(n, 0)
    T = HipPhysics(default_params(**dict(tprm, theta=theta, far_list_cap=1 << 18)), max_particles=n)
    t0 = time.perf_counter()
    ns = 0
    for seed in range(seeds):
        st, _ = W.plummer_charged(n, seed)
        T.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        ns = T.far_collect(stride)
    t1 = time.perf_counter()
    loss = T.far_train(steps, batch=batch, lr=2e-3, lr_final=1e-5, seed=3)
    t2 = time.perf_counter()
    model = T.far_get_model()
    T.destroy()
    info = {"systems": "%d Plummer spheres (scale radius 20) of N=%d, gravity + Coulomb" % (seeds, n),
            "theta": theta, "samples": int(ns), "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
            "steps": steps, "batch": batch, "loss_first": float(loss[:50].mean()), "loss_last": float(loss[-50:].mean())}
    return model, info


def run_drift(n, steps, theta, model):
    """Relative energy drift |E(t) - E(0)| / |E(0)| (E = KE + spin KE + PE,
    energy.js:11-62, FP64 reductions) of the learned method over `steps`
    updates of a virialised projected Plummer sphere (scale radius 20,
    gravity, softening 8, dt 1/128), E(0) after the first update."""
    from physsim_amd import HipPhysics, default_params, workloads as W, _abi as A
    st, prm, dt, _ = W.config3(n, side=1024.0, total_mass=16.0)
    prm = dict(prm, toggles0=A.GRAVITY | A.BARNES_HUT, theta=theta, far_field=A.FAR_LEARNED, far_list_cap=1 << 17)
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.far_set_model(model)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    g.update(dt)
    energy = lambda: (lambda s: s["linearKE"] + s["spinKE"] + s["pe"])(g.stats())
    e0 = energy()
    t0 = time.perf_counter()
    worst, trace = 0.0, []
    for k in range(1, steps + 1):
        g.update(dt)
        if k % max(1, steps // 10) == 0:
            d = (energy() - e0) / abs(e0)
            worst = max(worst, abs(d))
            trace.append([k, d])
            log("drift step %d: %.3e" % (k, d))
    g.destroy()
    return {"max_abs": worst, "trace": trace, "steps": steps, "n": n, "theta": theta, "method": "learned",
            "seconds": round(time.perf_counter() - t0, 1),
            "workload": "projected Plummer sphere (scale radius 20, truncated at 0.45 W, W = 1024), lognormal masses "
                        "(total 16), virialised Gaussian velocities, gravity, softening 8, dt 1/128 per update"}


def run_batch(n_systems, steps):
    """Batched ensemble (config 4): n_systems config-1 disks, N log-uniform in
    [256, 16384], gravity + Coulomb, theta 0.5, dt 1e-3 per update, stepped
    together (captured substep graphs, persistent walk with autotuned tile and
    stack).  systems-steps/s = systems x updates / wall time.  The reported
    value replays graphs of 2 substep bodies, as many back to back as the last
    update needed; graphs of 4 bodies and the whole 32-substep loop as one
    graph per update (finished systems predicated off inside it) are timed
    beside it."""
    from physsim_amd import HipBatch, default_params, workloads as W
    states, prm, dt = W.config4(n_systems)
    sizes = np.array([len(s["x"]) for s in states])
    b = HipBatch(default_params(**prm), sizes)
    for k, st in enumerate(states):
        b.upload(k, st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    b.compute_forces()
    b.autotune()
    timed = {}
    for bodies in (32, 4, 2):
        tune = b.tune(graph_bodies=bodies)
        b.update(dt)                              # warmup (captures this loop shape's graphs)
        t0 = time.perf_counter()
        subs = np.array([b.update(dt) for _ in range(steps)])
        wall = time.perf_counter() - t0
        timed[bodies] = (wall, subs, tune)
        log("batch: %d systems, %d bodies per graph: %.3e systems-steps/s"
            % (n_systems, bodies, n_systems * steps / wall))
    wall, subs, tune = timed[2]
    w4, w32 = timed[4][0], timed[32][0]
    return {"metric": "systems-steps/sec", "value": n_systems * steps / wall, "systems": n_systems,
            "particles": int(sizes.sum()), "steps": steps, "ms_per_step": wall / steps * 1e3,
            "particle_steps_per_s": float((subs * sizes[None, :]).sum() / wall),
            "substeps_mean": float(subs.mean()), "substeps_max": int(subs.max()), "tune": tune,
            "graphs_of_4_bodies": {"value": n_systems * steps / w4, "ms_per_step": w4 / steps * 1e3},
            "one_graph_of_32_bodies": {"value": n_systems * steps / w32, "ms_per_step": w32 / steps * 1e3},
            "workload": "config-1 disks, N = floor(256 * 64^U), gravity + Coulomb, softening 0.01, theta 0.5, "
                        "dt 1e-3 per update, despawn"}


def _cpu_update_rate(prm, st, dt, warm, threads):
    """The FP64 oracle on this host: particle-steps/s of one update after
    `warm` untimed ones (same parameters, smaller N)."""
    from oracle import bind
    bind.set_threads(threads)
    o = bind.Oracle(bind.default_params(**prm), len(st["x"]))
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    if "anti" in st:
        o.set_antimatter(st["anti"])
    for _ in range(warm):
        o.update(dt)
    n0 = o.count()
    t0 = time.perf_counter()
    steps = o.update(dt)
    return {"value": steps * n0 / (time.perf_counter() - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "n": len(st["x"])}


def run_legs(names, with_cpu, threads, model=None):
    """The rows beyond the headline, each timed on the GPU with the FP64 oracle
    beside it on a smaller N: signal-delayed leaf forces, collision mode merge,
    radiation reaction.  Rates are particle-steps/s over whole updates."""
    from physsim_amd import HipPhysics, GeonError, default_params, workloads as W, _abi as A
    from oracle import bind
    out = {}
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)

    def state(seed, n, side, speed, mass=(0.5, 3.0), spin=0.3, anti=0.0):
        # each leg draws from its own seed: a leg's workload does not depend on which legs ran before it
        g0 = np.random.default_rng(seed)
        return dict(x=f32(g0.uniform(0, side, n)), y=f32(g0.uniform(0, side, n)), wx=f32(g0.normal(0, speed, n)),
                    wy=f32(g0.normal(0, speed, n)), mass=f32(g0.uniform(*mass, n)),
                    charge=g0.integers(-10, 11, n) * 0.1, angw=f32(g0.uniform(-spin, spin, n)),
                    anti=(g0.random(n) < anti).astype(np.int32))

    def gpu_rate(prm, st, dt, warm, steps, cap=None, timers=True):
        g = HipPhysics(default_params(**prm), max_particles=cap or len(st["x"]))
        g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"],
                 antimatter=st.get("anti"))
        for _ in range(warm):
            g.update(dt)
        g.timing(timers)
        ps, t0 = 0, time.perf_counter()
        for _ in range(steps):
            n0 = g.getParticleCount()
            ps += g.update(dt) * n0
        wall = time.perf_counter() - t0
        walk_ms, walk_n = g.kernel_timing(TIMER_WALK)
        r = {"particle_steps_per_s": ps / wall, "ms_per_update": wall / steps * 1e3,
             "walk_ms_avg": walk_ms / max(walk_n, 1), "particles": g.getParticleCount()}
        g.destroy()
        return r

    if "delay" in names:
        # light-cone sources for every leaf pair (signal-delay.js:16-221): config-2 forces, relativity
        n, side, dt = 1 << 16, 256.0, 1.0 / 128
        prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.MAGNETIC | A.RELATIVITY | A.BARNES_HUT, domain_w=side,
                   domain_h=side, boundary_mode=A.BOUND_DESPAWN, theta=0.5, softening=8.0,
                   signal_delay=A.DELAY_ON, history_stride=1)
        st = state(4242, n, side, 0.3)
        r = gpu_rate(prm, st, dt, 8, 4)
        r_off = gpu_rate(dict(prm, signal_delay=A.DELAY_OFF), st, dt, 8, 4)
        r.update({"workload": "N=%d uniform %gx%g open box, gravity+Coulomb+magnetic, relativistic, BH theta 0.5, "
                              "signal delay on, a history sample per update, 8 updates of history first" % (n, side, side),
                  "current_time_sources": r_off, "slowdown_vs_current_time": r_off["particle_steps_per_s"] /
                  r["particle_steps_per_s"]})
        if with_cpu:
            cs = state(4243, 16384, side * 0.5, 0.3)
            r["cpu_baseline"] = _cpu_update_rate(prm, cs, dt, 8, threads)
        out["signal_delay"] = r
        log("leg delay: %.3e particle-steps/s" % r["particle_steps_per_s"])
    if "merge" in names:
        # collision mode merge (collisions.js:24-158): about one body in ten overlaps another at the start
        n, side, dt = 1 << 20, 16384.0, 1.0 / 64
        prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.BARNES_HUT, domain_w=side, domain_h=side,
                   boundary_mode=A.BOUND_DESPAWN, theta=0.5, softening=8.0, collision_mode=A.COL_MERGE,
                   signal_delay=A.DELAY_OFF)
        st = state(4244, n, side, 0.5, anti=0.2)
        g = HipPhysics(default_params(**prm), max_particles=2 * n)
        g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"], antimatter=st["anti"])
        g.compute_forces()
        t0 = time.perf_counter()
        g.update(dt)
        first_ms = (time.perf_counter() - t0) * 1e3
        events = len(g.collision_events())
        left = g.getParticleCount()
        g.destroy()
        steady = gpu_rate(prm, st, dt, 2, 4, cap=2 * n)
        passing = gpu_rate(dict(prm, collision_mode=A.COL_PASS), st, dt, 2, 4, cap=2 * n)
        r = {"workload": "N=%d uniform %gx%g open box, masses U[0.5, 3] (radii 0.8-1.4), 20%% antimatter, "
                         "gravity+Coulomb, BH theta 0.5, dt 1/64" % (n, side, side),
             "first_update_ms": first_ms, "first_update_events": events, "particles_after_first": left,
             "steady": steady, "pass_mode": passing,
             "merge_pass_ms_per_update": steady["ms_per_update"] - passing["ms_per_update"],
             "particle_steps_per_s": steady["particle_steps_per_s"]}
        if with_cpu:
            cs = state(4245, 65536, side * 0.25, 0.5, anti=0.2)
            r["cpu_baseline"] = _cpu_update_rate(prm, cs, dt, 0, threads)
        out["merge"] = r
        log("leg merge: first update %.1f ms (%d events), steady %.3e particle-steps/s"
            % (first_ms, events, steady["particle_steps_per_s"]))
    if "bounce" in names:
        # collision mode bounce (integrator.js:669-724) and bounce walls (606-667): bodies packed at about half
        # the box's area, so most touch a neighbour
        n, side, dt = 1 << 20, 2896.0, 1.0 / 64
        prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.BARNES_HUT, domain_w=side, domain_h=side,
                   boundary_mode=A.BOUND_BOUNCE, theta=0.5, softening=8.0, collision_mode=A.COL_BOUNCE,
                   signal_delay=A.DELAY_OFF)
        st = state(4248, n, side, 0.3)
        contacts = gpu_rate(prm, st, dt, 2, 4)
        passing = gpu_rate(dict(prm, collision_mode=A.COL_PASS), st, dt, 2, 4)
        r = {"workload": "N=%d uniform %gx%g box with bounce walls, masses U[0.5, 3] (radii 0.8-1.4, about half the "
                         "area covered), gravity+Coulomb, BH theta 0.5, dt 1/64, collision mode bounce"
                         % (n, side, side),
             "particle_steps_per_s": contacts["particle_steps_per_s"], "contacts": contacts, "pass_mode": passing,
             "contact_ms_per_update": contacts["ms_per_update"] - passing["ms_per_update"]}
        if with_cpu:
            cs = state(4249, 65536, side * 0.25, 0.3)
            r["cpu_baseline"] = _cpu_update_rate(prm, cs, dt, 0, threads)
        out["bounce"] = r
        log("leg bounce: %.3e particle-steps/s (pass mode %.3e)"
            % (contacts["particle_steps_per_s"], passing["particle_steps_per_s"]))
    if "config3" in names:
        # configuration 3: gravity + 1PN EIH (with gravitomagnetism) + Higgs + axion on the 128^2 grid
        n = 1 << 20
        st, prm, dt, fields = W.config3(n, grid=128, side=1024.0, total_mass=16.0)
        g = HipPhysics(default_params(**prm), max_particles=n)
        g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        g.set_field(0, fields["higgs"])
        g.set_field(1, fields["axion"])
        g.update(dt)
        steps = 8
        g.timing(True)
        ps, t0 = 0, time.perf_counter()
        for _ in range(steps):
            n0 = g.getParticleCount()
            ps += g.update(dt) * n0
        wall = time.perf_counter() - t0
        ph = {}
        for key, tid in (("walk", TIMER_WALK), ("tree", TIMER_TREE), ("sort_gather", TIMER_SORT),
                         ("kick_drift", TIMER_KICK), ("field_deposit", TIMER_DEPOSIT), ("field_kdk", TIMER_KDK),
                         ("field_to_particle", TIMER_FIELD_PART)):
            ms, cnt = g.kernel_timing(tid)
            ph[key] = {"ms_avg": ms / cnt if cnt else 0.0, "launches": cnt}
        g.timing(False)
        G = 128
        # algorithmic bytes (per deposit: positions + one value per particle, the grid written once;
        # per KDK: two FFT self-gravity convolutions of 2 x G^2 x 8 B per transform pair and five
        # stencil sweeps of ~40 B per cell)
        dep_bytes = n * 12 + G * G * 4
        kdk_bytes = 2 * 2 * 2 * G * G * 8 + 5 * G * G * 40
        for key, b in (("field_deposit", dep_bytes), ("field_kdk", kdk_bytes)):
            t = ph[key]["ms_avg"] * 1e-3
            ph[key]["algorithmic_bytes"] = b
            ph[key]["achieved_gbs"] = b / t / 1e9 if t > 0 else None
            ph[key]["frac_of_hbm"] = b / t / 1e9 / HBM_PEAK_GBS if t > 0 else None
        r = {"particle_steps_per_s": ps / wall, "ms_per_update": wall / steps * 1e3, "particles": g.getParticleCount(),
             "updates": steps, "phases": ph,
             "workload": "N=%d projected Plummer sphere (scale radius 20), gravity + gravitomagnetism + 1PN EIH "
                         "(velocity-Verlet pass), Higgs + axion on a 128^2 PQS grid with FFT self-gravity, "
                         "Gaussian random initial fields (index -2, rms 0.01), BH theta 0.5, dt 1/128" % n,
             "note": "the 128^2 field work is latency-bound: one grid is 64 KiB, so a field kernel moves a few "
                     "hundred KiB and its time is launch and dependency latency, not HBM bandwidth; the deposit "
                     "reads 12 B per particle and sums each block's stencils in an LDS tile, so its time is the "
                     "stencil arithmetic and one global atomic per touched cell"}
        g.destroy()
        if with_cpu:
            cs, cprm, cdt, cf = W.config3(16384, grid=128, side=256.0, total_mass=1.0)
            bind.set_threads(threads)
            o = bind.Oracle(bind.default_params(**cprm), 16384)
            o.upload(cs["x"], cs["y"], cs["wx"], cs["wy"], cs["mass"], cs["charge"], cs["angw"])
            o.field_set(0, cf["higgs"])
            o.field_set(1, cf["axion"])
            o.update(cdt)
            t0 = time.perf_counter()
            sc = o.update(cdt)
            r["cpu_baseline"] = {"value": sc * o.count() / (time.perf_counter() - t0), "unit": "particle-steps/s",
                                 "cores": threads, "kind": "port", "n": 16384}
        out["config3"] = r
        log("leg config3: %.3e particle-steps/s" % r["particle_steps_per_s"])
    if "torus" in names:
        # configuration 5 on the torus: the learned correction in windowed periodic mode (nodes straddling
        # the observer's min-image window are opened, so Barnes-Hut converges to the min-image direct sum)
        n = 1 << 20
        st, prm, dt = W.config2(n)
        prm = dict(prm, periodic_mode=A.PERIODIC_WINDOWED, far_list_cap=1 << 18)
        g = HipPhysics(default_params(**prm), max_particles=n)
        g.far_set_model(model)
        g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        rs = np.sort(np.random.Generator(np.random.Philox(5001)).choice(n, 4096, replace=False)).astype(np.int32)
        o = bind.Oracle(bind.default_params(**prm), n)
        o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        fx, fy = o.direct_forces(rs)
        del o
        r = {"workload": "N=%d uniform %.0fx%.0f torus, windowed min-image Barnes-Hut, config-2 forces "
                         "(relativistic, spins), dt 1/128; force error against the FP64 min-image direct sum "
                         "on 4096 particles" % (n, prm["domain_w"], prm["domain_h"]), "methods": {}}
        for name, theta, far in (("plain", 0.3, A.FAR_NONE), ("learned", 0.3, A.FAR_LEARNED),
                                 ("multipole", 0.3, A.FAR_MULTIPOLE)):
            p = A.GeonParams.from_buffer_copy(g.params)
            p.theta, p.far_field = theta, far
            g.set_params(p)
            g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
            while True:
                try:
                    g.compute_forces()
                    break
                except GeonError as e:
                    # lanes near the window edge accept more nodes than the open box's: grow each
                    # particle's far queue until the longest one fits (the overflow changes no state)
                    if e.code != A.EOVERFLOW or p.far_list_cap >= (1 << 21):
                        raise
                    p.far_list_cap *= 2
                    g.set_params(p)
                    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
            r["far_list_cap"] = int(p.far_list_cap)
            f = g.forces(n)
            err = float(np.sqrt(np.sum((f["fx"][rs] - fx) ** 2 + (f["fy"][rs] - fy) ** 2) / np.sum(fx ** 2 + fy ** 2)))
            g.update(dt)
            ps, t0 = 0, time.perf_counter()
            for _ in range(4):
                n0 = g.getParticleCount()
                ps += g.update(dt) * n0
            wall = time.perf_counter() - t0
            r["methods"][name] = {"theta": theta, "force_rel_l2": err, "particle_steps_per_s": ps / wall,
                                  "ms_per_update": wall / 4 * 1e3}
            log("leg torus %s theta %.2f: error %.3e, %.3e particle-steps/s" % (name, theta, err, ps / wall))
        g.destroy()
        out["torus_windowed"] = r
    if "radiation" in names:
        # Landau-Lifshitz reaction every substep, quadrupole radiation every update
        n, side, dt = 1 << 20, 1024.0, 1.0 / 128
        prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.RELATIVITY | A.RADIATION | A.BARNES_HUT, domain_w=side,
                   domain_h=side, boundary_mode=A.BOUND_DESPAWN, theta=0.5, softening=8.0, signal_delay=A.DELAY_OFF)
        st = state(4246, n, side, 0.3)
        r = gpu_rate(prm, st, dt, 1, 4)
        base = gpu_rate(dict(prm, toggles0=prm["toggles0"] & ~A.RADIATION), st, dt, 1, 4)
        r.update({"workload": "N=%d uniform %gx%g open box, gravity+Coulomb, relativistic, BH theta 0.5, radiation "
                              "reaction (jerk, Landau-Lifshitz, quadrupole)" % (n, side, side),
                  "without_radiation": base,
                  "overhead": base["particle_steps_per_s"] / r["particle_steps_per_s"] - 1.0})
        if with_cpu:
            cs = state(4247, 65536, side * 0.25, 0.3)
            r["cpu_baseline"] = _cpu_update_rate(prm, cs, dt, 0, threads)
        out["radiation"] = r
        log("leg radiation: %.3e particle-steps/s (%.1f%% over the same run without)"
            % (r["particle_steps_per_s"], 100 * r["overhead"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    # opening-angle ladders, coarsest first: each method runs at the coarsest
    # theta whose force error (measured below) is <= --err-target
    ap.add_argument("--theta-plain", default="0.05,0.045,0.04,0.035,0.03")
    ap.add_argument("--theta-learned", default="0.3,0.29,0.28,0.26,0.24,0.22")
    ap.add_argument("--theta-multipole", default="0.35,0.32,0.3,0.28,0.25")
    ap.add_argument("--theta-train", type=float, default=0.3)
    ap.add_argument("--err-target", type=float, default=1e-4)
    ap.add_argument("--train-steps", type=int, default=80000)
    ap.add_argument("--train-seeds", type=int, default=32)
    ap.add_argument("--err-sample", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--cpu-sample-1core", type=int, default=1 << 14)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="energy-drift leg: updates of a virialised Plummer sphere with the learned method; 0 skips")
    ap.add_argument("--drift-n", type=int, default=1 << 20)
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="batched-ensemble leg (config 4): systems; 0 skips it")
    ap.add_argument("--legs", default="config3,torus,delay,merge,bounce,radiation",
                    help="rows beyond the headline to time ('' for none)")
    ap.add_argument("--methods", default="plain,learned,multipole",
                    help="comma list of force methods to run (the learned one is always run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the Barnes-Hut path here is single-GPU; got --gpus %d"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from physsim_amd import HipPhysics, default_params, workloads as W, _abi as A
    from oracle import bind
    threads = host_threads()
    bind.set_threads(threads)
    n = args.n
    st, prm, dt = W.config2(n, open_box=True)
    g = HipPhysics(default_params(**dict(prm, far_list_cap=1 << 17)), max_particles=n)

    def upload():
        g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])

    def set_method(theta, far):
        p = A.GeonParams.from_buffer_copy(g.params)
        p.theta, p.far_field = theta, far
        g.set_params(p)

    log("n=%d: training the far-field network" % n)
    model, train_info = train_far_model(args.theta_train, args.train_seeds, 65536, 8, args.train_steps, 16384)
    g.far_set_model(model)
    upload()

    # ---- force error of each method vs the FP64 direct sum (the checker) ----
    rs = np.sort(np.random.Generator(np.random.Philox(5000)).choice(n, args.err_sample, replace=False)).astype(np.int32)
    orc = bind.Oracle(bind.default_params(**prm), n)
    orc.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    te = time.perf_counter()
    fx, fy = orc.direct_forces(rs)
    te = time.perf_counter() - te
    del orc
    ladders = {"plain": ([float(t) for t in args.theta_plain.split(",")], A.FAR_NONE),
               "learned": ([float(t) for t in args.theta_learned.split(",")], A.FAR_LEARNED),
               "multipole": ([float(t) for t in args.theta_multipole.split(",")], A.FAR_MULTIPOLE)}
    keep = set(args.methods.split(",")) | {"learned"}
    res, methods = {}, {}
    for name, (ladder, far) in ladders.items():
        if name not in keep:
            continue
        tried = []
        for theta in ladder:
            set_method(theta, far)
            g.compute_forces()
            f = g.forces(n)
            dx, dy = f["fx"][rs] - fx, f["fy"][rs] - fy
            rel = np.sqrt((dx ** 2 + dy ** 2) / np.maximum(fx ** 2 + fy ** 2, 1e-300))
            r = {"theta": theta, "force_rel_l2": float(np.sqrt(np.sum(dx ** 2 + dy ** 2) / np.sum(fx ** 2 + fy ** 2))),
                 "force_median": float(np.median(rel)), "force_p99": float(np.percentile(rel, 99))}
            tried.append((theta, r["force_rel_l2"]))
            log("%s theta %.3f: force error %.3e" % (name, theta, r["force_rel_l2"]))
            if r["force_rel_l2"] <= args.err_target or theta == ladder[-1]:
                break
        r["ladder"] = tried
        r["meets_target"] = r["force_rel_l2"] <= args.err_target
        res[name] = r
        methods[name] = (r["theta"], far)

    def end_state_error(gg, nn, sample):
        """Relative L2 force error of the sampled live particles against the
        oracle's FP64 direct sum, at the current state."""
        gg.compute_forces()
        f = gg.forces(nn)
        d = gg.download()
        live = np.nonzero(d["alive"])[0]
        o = bind.Oracle(bind.default_params(**prm), len(live))
        o.upload(d["x"][live], d["y"][live], d["wx"][live], d["wy"][live], d["mass"][live], d["charge"][live],
                 d["angw"][live])
        pos = np.full(len(d["alive"]), -1, dtype=np.int64)
        pos[live] = np.arange(len(live))
        keep = sample[pos[sample] >= 0]
        ex, ey = o.direct_forces(pos[keep].astype(np.int32))
        del o
        dx, dy = f["fx"][keep] - ex, f["fy"][keep] - ey
        return float(np.sqrt(np.sum(dx ** 2 + dy ** 2) / np.sum(ex ** 2 + ey ** 2)))

    # ---- timed updates, each method from the same initial state ----
    c_f = flops_per_interaction(prm["toggles0"])
    for name, (theta, far) in methods.items():
        upload()
        set_method(theta, far)
        for _ in range(args.warmup):
            g.update(dt)
        upload()                              # timed steps start from the initial state again
        g.compute_forces()
        g.timing(True)
        t0 = time.perf_counter()
        substeps = 0
        for k in range(args.steps):
            s = g.update(dt)
            substeps += s
            log("%s step %d substeps %d" % (name, k, s))
        g.getParticleCount()                  # synchronizes the stream
        wall = time.perf_counter() - t0
        alive = g.getParticleCount()
        ph = {}
        for key, tid in (("walk", TIMER_WALK), ("mlp", TIMER_MLP), ("tree", TIMER_TREE), ("sort_gather", TIMER_SORT),
                         ("kick_drift", TIMER_KICK)):
            ms, cnt = g.kernel_timing(tid)
            ph[key] = ms / cnt if cnt else 0.0
        g.timing(False)
        cnt = g.counters()                    # counters of the last force pass
        r = res[name]
        r.update({"particle_steps_per_s": substeps * alive / wall, "ms_per_step": wall / args.steps * 1e3,
                  "substeps": substeps, "phases_ms_avg": ph,
                  "pair_interactions_per_particle": cnt["pair_interactions"] / alive,
                  "node_interactions_per_particle": cnt["node_interactions"] / alive})
        walk_s = ph["walk"] * 1e-3
        inter = cnt["pair_interactions"] + cnt["node_interactions"]
        r["walk_tflops"] = inter * c_f / walk_s / 1e12 if walk_s > 0 else 0.0
        if name == "learned" and ph["mlp"] > 0:
            r["mlp_tflops"] = cnt["node_interactions"] * MLP_FLOPS_PER_PAIR / (ph["mlp"] * 1e-3) / 1e12
        log("%s: %.3e particle-steps/s" % (name, r["particle_steps_per_s"]))
        # the same error measure at the state the timed updates ended in
        r["force_rel_l2_end"] = end_state_error(g, n, rs)
        log("%s: force error %.3e at the end of the timed updates" % (name, r["force_rel_l2_end"]))

    L, P = res["learned"], res.get("plain")
    dominant_mlp = L["phases_ms_avg"]["mlp"] >= L["phases_ms_avg"]["walk"]
    if dominant_mlp:
        default = n == 1 << 20 and abs(L["theta"] - 0.30) < 1e-9
        roof = {"bound": "mfma", "kernel": "k_far_mlp", "achieved": L["mlp_tflops"], "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": L["mlp_tflops"] / BF16_PEAK_TFLOPS,
                "traffic": MLP_TRAFFIC_DEFAULT if default else None,
                "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch (copy-calibrated), "
                                  "profiles/k_far_mlp_open1M_theta0.30.md" if default else None,
                "note": "vector work (the hard-swish of 128 hidden values per lane, features, output rotation) and the MFMAs "
                        "barely overlap on a SIMD; their sum, not MFMA throughput, bounds this kernel",
                "pairs_per_launch": L["node_interactions_per_particle"] * n, "flops_per_pair": MLP_FLOPS_PER_PAIR,
                "ms_avg": L["phases_ms_avg"]["mlp"]}
    else:
        roof = {"bound": "valu", "kernel": "k_walk", "achieved": L["walk_tflops"], "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": L["walk_tflops"] / FP32_PEAK_TFLOPS, "traffic": None,
                "ms_avg": L["phases_ms_avg"]["walk"]}
    walk_roof = {"bound": "valu", "kernel": "k_walk_list (learned run)", "achieved": L["walk_tflops"],
                 "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": L["walk_tflops"] / FP32_PEAK_TFLOPS,
                 "frac_nonpacked": L["walk_tflops"] / FP32_NONPACKED_TFLOPS,
                 "flops_per_interaction": c_f, "ms_avg": L["phases_ms_avg"]["walk"]}
    walk_roof_plain = None
    if P is not None and "walk_tflops" in P:
        walk_roof_plain = {"bound": "valu", "kernel": "k_walk (plain run)", "achieved": P["walk_tflops"],
                           "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": P["walk_tflops"] / FP32_PEAK_TFLOPS,
                           "frac_nonpacked": P["walk_tflops"] / FP32_NONPACKED_TFLOPS,
                           "flops_per_interaction": c_f, "ms_avg": P["phases_ms_avg"]["walk"]}

    batch = run_batch(args.batch_systems, args.steps) if args.batch_systems > 0 else None
    drift = run_drift(args.drift_n, args.drift_steps, L["theta"], model) if args.drift_steps > 0 else None

    legs = run_legs([x for x in args.legs.split(",") if x], not args.no_cpu, threads, model) if args.legs else None

    cpu = None
    if not args.no_cpu:
        # ---- CPU baseline: the reference algorithm in FP64 on this host ----
        theta_c = P["theta"] if P else 0.04

        def cpu_rate(nc, nthreads):
            cs, cprm, _ = W.config2(nc, open_box=True)
            cprm["theta"] = theta_c
            bind.set_threads(nthreads)
            oc = bind.Oracle(bind.default_params(**cprm), nc)
            oc.upload(cs["x"], cs["y"], cs["wx"], cs["wy"], cs["mass"], cs["charge"], cs["angw"])
            tc = time.perf_counter()
            oc.compute_forces()                # bootstrap force pass (part of the first update)
            tb = time.perf_counter()
            steps_c = oc.update(dt)
            tc2 = time.perf_counter()
            return steps_c * oc.count() / (tc2 - tb), steps_c, tc2 - tb, tb - tc

        rate, steps_c, t_upd, t_boot = cpu_rate(args.cpu_sample, threads)
        rate1, steps1, t_upd1, _ = cpu_rate(args.cpu_sample_1core, 1)
        bind.set_threads(threads)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "FP64 oracle (integrator.js/forces.js/quadtree.js restated in C, OpenMP over particles), "
                         "plain monopole BH at theta %.2f, same dynamics at N=%d (same density), one update of %d "
                         "substeps, %.1f s; bootstrap force pass %.1f s untimed"
                         % (theta_c, args.cpu_sample, steps_c, t_upd, t_boot),
               "host": host_info(),
               "one_core": {"value": rate1, "unit": "particle-steps/s", "cores": 1, "n": args.cpu_sample_1core,
                            "seconds": round(t_upd1, 2), "substeps": steps1}}

    line = {
        "metric": "particle-steps/sec at N=1M, plain vs learned-correction BH at force err <=1e-4",
        "value": L["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": L["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (forces, Kahan-compensated sums); network: bf16 layer 1, fp16 layers 2-3, f32 accumulation",
        "data": "synthetic (Philox seed 2000: uniform open box, lognormal masses, 0.1-quantized charges)",
        "config": {"workload": "open 1024x1024 box, N=%d uniform, gravity+Coulomb+Yukawa(mu=1)+magnetic dipole, "
                               "relativistic, spins U[-1,1], softening 8, dt 1/128 per update; learned BH at "
                               "theta %.3f vs plain BH at theta %.3f, each the coarsest of its ladder with force "
                               "error <= %.0e" % (n, L["theta"], P["theta"] if P else float("nan"), args.err_target),
                   "n": n, "err_target": args.err_target},
        "speedup_learned_vs_plain": L["particle_steps_per_s"] / P["particle_steps_per_s"] if P else None,
        "methods": res,
        "force_error_reference": {"kind": "FP64 direct sum (forces.js:93-126)", "sample": int(args.err_sample),
                                  "seconds": round(te, 2)},
        "training": train_info,
        "roofline": roof,
        "roofline_walk": walk_roof,
        "roofline_walk_plain": walk_roof_plain,
        "batch": batch,
        "energy_drift": drift,
        "legs": legs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
