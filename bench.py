"""Throughput of the Brownian-dynamics step on one MI355X.

Workload (the aging-quench configuration): N = 32768 bidisperse WCA particles
(16384 sigma=1.4 + 16384 sigma=1.0), phi = 0.58, box L = 38.1161, FP64,
Ermak-McCammon dt = 1e-4, equilibrated at T = 1.0 during the warm-up steps
and instantly quenched to T = 0.1 before the timed steps.  One step =
update + cell binning + forces for every particle (main.cu:494-586).

Also reported: the 256-replica uneven batch (N_r log-uniform in [1000, 16000],
phi_r in [0.52, 0.60], quench T_r in [0.05, 0.5], distinct seeds) in one
engine, and the oracle's single-core CPU rate on a bounded sample.

Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bd_aging_amd.engine import (Contraction, Engine, make_params, box_side,  # noqa: E402
                                 grid_barrier_cost, launch_floor, stream_bandwidth)
from bd_aging_amd.init import relax_overlaps, relax_batch  # noqa: E402

HBM_PEAK = 8.0e12          # B/s, MI355X
B_FORCES = 72              # algorithmic bytes per particle, force pass
B_BIN = 36                 # algorithmic bytes per particle, binning pass
B_UPDATE = 120             # algorithmic bytes per particle, update pass (ring excluded)
B_STEP = 228               # algorithmic bytes per particle-step (bin 36 + forces 72 + update 120)
# g(r) binnings timed: config 1's (200 bins to L/2), the batch's (128 bins to
# 5 sigma) and the reference script's (400 bins to 15)
GDER_CASES = (("200 bins to L/2", 200, None), ("128 bins to 5", 128, 5.0),
              ("400 bins to 15", 400, 15.0))
CV3D_BINS, CV3D_RANGE = 400, 5.0   # force-contraction sample: bins to range
RNG_BYTES = 112            # rannew64 ring and LCG bytes per particle-step (8 x 8 B read,
                           # 4 x 8 B written, LCG 16 B), beside the update's 120 B
# HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs in
# profiles/k_forces_traffic_pmc.txt, scaled by a copy kernel of known size:
# plain N=32768 with the four-wavefront chunked walk (forces only, and forces
# with the fused update), batch with the slab walk (both at deferral depth 16)
TRAFFIC_FORCES_PLAIN = 6.77e6
TRAFFIC_FUSED_PLAIN = 1.487e7
TRAFFIC_FORCES_BATCH = 1.489e8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch_specs(R=256, seed=20251203):
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(1000), np.log(16000), R))).astype(int)
    phi = rng.uniform(0.52, 0.60, R)
    T = rng.uniform(0.05, 0.5, R)
    return n, phi, T


def timed_steps(eng, steps):
    """Seconds for `steps` BD steps replayed as graphs of steps - 1 (built
    beforehand), as the timed runs replay them."""
    G = eng.graph_steps()
    eng.set_graph_steps(max(1, min(256, steps - 1)))
    eng.prepare_step(steps)
    t0 = time.perf_counter()
    eng.step(steps)
    dt = time.perf_counter() - t0
    eng.set_graph_steps(G)
    return dt


def confirm_flat(eng, steps):
    """The autotuner times forces-only launches, so it cannot see what fusing
    the next step's move into the flat walk saves.  Time `steps` real steps
    with the autotuned shape and with the flat walk (which fuses the move),
    keep the faster.  Every shape gives the same bits, so the state advances
    the same either way.  Returns the seconds of both (None: no flat walk)."""
    tuned = (eng.brick(), eng.defer_depth(), eng.walk(), eng.pass_split())
    if tuned[2] == 5:
        return None
    t_tuned = timed_steps(eng, steps)
    try:
        eng.set_brick(0)
        eng.set_defer_depth(0)
        eng.set_walk(5)
        eng.set_pass_split(0)
    except Exception:  # the flat walk does not apply to this engine
        eng.set_brick(tuned[0])
        eng.set_defer_depth(tuned[1])
        eng.set_walk(tuned[2])
        eng.set_pass_split(tuned[3])
        return None
    t_flat = timed_steps(eng, steps)
    if t_flat >= t_tuned:
        eng.set_brick(tuned[0])
        eng.set_defer_depth(tuned[1])
        eng.set_walk(tuned[2])
        eng.set_pass_split(tuned[3])
    return {"steps": steps, "autotuned_s": t_tuned, "flat_fused_s": t_flat}


def pick_shape(eng, shape, trials, confirm_steps=0):
    """Autotune the force kernel, or set the given (brick, deferral depth);
    with confirm_steps, check the autotuner's pick against the fused flat walk
    on real steps."""
    confirm = None
    if shape is None:
        brick = eng.autotune(trials)
        if confirm_steps > 0:
            confirm = confirm_flat(eng, confirm_steps)
            brick = eng.brick()
    else:
        eng.set_brick(shape[0])
        eng.set_defer_depth(shape[1])
        eng.set_walk(shape[2] if len(shape) > 2 else 0)
        eng.set_pass_split(shape[3] if len(shape) > 3 else 0)
        brick = shape[0]
    return {"brick": brick,
            "walk": ["rows", "slabs", "chunks2", "chunks4", "list", "flat"][eng.walk()],
            "pass_split": eng.pass_split(), "fused_update": bool(eng.fused_update()),
            "defer_depth": eng.defer_depth(),
            "chosen_by": ("autotuner" if confirm is None else "autotuner, then timed steps against "
                          "the fused flat walk") if shape is None else "--shape",
            "confirm": confirm}


def timed_launch_mode(eng, steps, graph_steps):
    """Graphs of G = min(graph_steps, steps - 1) steps, built before the timed
    region: a timed call of any length >= 2 replays at least one graph.
    Returns the launch mode as the timed call runs it."""
    G = max(1, min(graph_steps, steps - 1)) if graph_steps > 0 else 0
    eng.set_graph_steps(G)
    eng.prepare_step(steps)
    fused = bool(eng.fused_update())
    if G == 0:
        return {"launch": "eager", "graph_steps": 0, "fused_update": fused,
                "graph_replays": 0, "eager_steps": steps}
    if fused and steps > G:
        replays, eager = (steps - 1) // G, (steps - 1) % G
        note = "first update and last forces eager; (forces + move, binning) pairs in graphs"
    else:
        replays, eager = steps // G, steps % G
        note = "whole steps in graphs"
    return {"launch": "graph" if replays > 0 else "eager", "graph_steps": G,
            "fused_update": fused, "graph_replays": replays, "eager_steps": eager, "note": note}


def run_plain(steps, warmup, graph_steps=16, shape=None, persistent_steps=0):
    nb = ns = 16384
    side = box_side(nb, ns, 1.4, 1.0, 0.58)
    log(f"[bench] relaxing N=32768 phi=0.58 L={side:.4f}")
    x = relax_overlaps(nb, ns, side, seed=1, graph_steps=graph_steps)
    p = make_params(nb, ns, side=side, dt=1e-4, T0=1.0, Tf=0.1, t0=0.0, tf=0.0,
                    first_step=-warmup, seed=777)
    eng = Engine(p)
    eng.set_graph_steps(graph_steps)
    eng.set_positions(x)
    log(f"[bench] warmup {warmup} steps at T=1.0")
    eng.step(warmup)
    shape = pick_shape(eng, shape, 20)
    log(f"[bench] force kernel {shape}")
    eng.mark_origin()
    mode = timed_launch_mode(eng, steps, graph_steps)
    log(f"[bench] timed {steps} steps at T=0.1, {mode}")
    t0 = time.perf_counter()
    eng.step(steps)
    dt = time.perf_counter() - t0
    ksteps = min(200, steps)
    ms = eng.kernel_times(ksteps) / ksteps
    fused_ms = None
    if eng.fused_update():
        fused_ms = (eng.kernel_times_fused(ksteps + 1) / [ksteps + 1, ksteps]).tolist()
    obs = eng.observables(7.0)
    gder = time_gder(eng, side, nb + ns, reps=3)
    reps = 5
    # force contraction (Contracting_Vector3D): one sample of the current state
    cv = Contraction(CV3D_BINS, CV3D_RANGE)
    cv.add(eng)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        cv.add(eng)
    cv3d_s = (time.perf_counter() - t0) / reps
    cv.close()
    persistent = None
    if persistent_steps > 0:
        persistent = persistent_compare(eng, persistent_steps, graph_steps)
    eng.close()
    return {"N": nb + ns, "side": side, "seconds": dt, "rate": (nb + ns) * steps / dt,
            "kernel_ms": ms.tolist(), "obs": obs, "x": x, "params": p, "shape": shape,
            "mode": mode,
            "gder": gder, "fused_ms": fused_ms, "cv3d_s": cv3d_s, "persistent": persistent}


def persistent_compare(eng, steps, graph_steps):
    """The same system stepped `steps` times by one launch of the persistent
    cooperative step kernel (three grid barriers per step) and by graph
    replays, and the cost of one grid barrier alone at several grid sizes."""
    eng.set_persistent(steps)
    eng.step(1)  # warm-up launch
    t0 = time.perf_counter()
    eng.step(steps)
    t_p = time.perf_counter() - t0
    eng.set_persistent(0)
    timed_launch_mode(eng, steps, graph_steps)
    t0 = time.perf_counter()
    eng.step(steps)
    t_g = time.perf_counter() - t0
    _, most = grid_barrier_cost(64, 10)
    barrier = {str(g): round(grid_barrier_cost(g)[0], 2) for g in (64, 256, 512, 1024, 2048, most)
               if g <= most}
    log(f"[bench] persistent {t_p / steps * 1e6:.1f} us/step, graphs {t_g / steps * 1e6:.1f} "
        f"us/step; grid barrier us by workgroups {barrier}")
    return {"steps": steps, "us_per_step_persistent": t_p / steps * 1e6,
            "us_per_step_graphs": t_g / steps * 1e6, "grid_barrier_us": barrier,
            "coresident_barrier_workgroups": most}


def time_gder(eng, side, N, reps=3):
    """Pair-histogram rate at each binning of GDER_CASES: ordered pairs per
    second (N (N - 1) per call), the fraction of pairs within range (each one
    square root, one bin and one LDS increment) and the LDS increments per
    second."""
    out = []
    for name, nbins, rng in GDER_CASES:
        r = 0.5 * side if rng is None else rng
        eng.gder_counts(nbins, r)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            c = eng.gder_counts(nbins, r)
        dt = (time.perf_counter() - t0) / reps
        hits = int(np.asarray(c).sum())
        pairs = N * (N - 1)
        out.append({"binning": name, "bins": nbins, "range": round(r, 4), "ms": dt * 1e3,
                     "pairs_per_s": pairs / dt, "in_range_fraction": hits / pairs,
                     "lds_atomics_per_s": hits / dt})
    return out


def run_batch(steps, warmup, graph_steps=16, shape=None):
    n, phi, T = batch_specs()
    specs, params = [], []
    for r in range(len(n)):
        nb = int(n[r]) // 2
        ns = int(n[r]) - nb
        side = box_side(nb, ns, 1.4, 1.0, float(phi[r]))
        specs.append((nb, ns, side, 1.4, 1.0, 5000 + r))
        params.append(make_params(nb, ns, side=side, dt=1e-4, T0=1.0, Tf=float(T[r]), t0=0.0,
                                  tf=0.0, first_step=-warmup, seed=90000 + r))
    log(f"[bench] relaxing 256 replicas, sum N = {int(n.sum())}")
    xs = relax_batch(specs, graph_steps=graph_steps)
    eng = Engine(params)
    eng.set_graph_steps(graph_steps)
    for r, x in enumerate(xs):
        eng.set_positions(x, r)
    eng.step(warmup)
    shape = pick_shape(eng, shape, 10, confirm_steps=100)
    log(f"[bench] batch force kernel {shape}")
    eng.mark_origin()
    mode = timed_launch_mode(eng, steps, graph_steps)
    log(f"[bench] batch timed {steps} steps, {mode}")
    t0 = time.perf_counter()
    eng.step(steps)
    dt = time.perf_counter() - t0
    ksteps = min(50, steps)
    ms = eng.kernel_times(ksteps) / ksteps
    eng.close()
    tot = int(n.sum())
    return {"replicas": len(n), "sum_N": tot, "seconds": dt, "rate": tot * steps / dt,
            "kernel_ms": ms.tolist(), "shape": shape, "specs": specs, "T": T, "x": xs,
            "mode": mode}


def run_stress(steps, warmup, graph_steps=16, shape=None):
    """N=262144 at phi=0.58, T=0.5, dt=5e-5: per-kernel times and HBM fractions."""
    nb = ns = 131072
    side = box_side(nb, ns, 1.4, 1.0, 0.58)
    log(f"[bench] stress: relaxing N=262144 phi=0.58 L={side:.4f}")
    x = relax_overlaps(nb, ns, side, seed=3, graph_steps=graph_steps)
    p = make_params(nb, ns, side=side, dt=5e-5, T0=0.5, Tf=0.5, first_step=0, seed=4242)
    eng = Engine(p)
    eng.set_graph_steps(graph_steps)
    eng.set_positions(x)
    eng.step(warmup)
    shape = pick_shape(eng, shape, 10, confirm_steps=200)
    timed_launch_mode(eng, steps, graph_steps)
    t0 = time.perf_counter()
    eng.step(steps)
    dt = time.perf_counter() - t0
    ms = eng.kernel_times(100) / 100
    geo = eng.geometry()
    gder = time_gder(eng, side, nb + ns, reps=1)
    eng.close()
    N = nb + ns
    kern = {}
    for name, b, m in (("update", B_UPDATE, ms[0]), ("binning", B_BIN, ms[1]),
                       ("forces", B_FORCES, ms[2])):
        gbs = b * N / (m * 1e-3)
        kern[name] = {"ms": m, "bytes_per_particle": b, "GB_s": gbs / 1e9, "hbm_frac": gbs / HBM_PEAK}
    kern["update"]["rng_bytes_per_particle"] = RNG_BYTES
    kern["update"]["GB_s_with_rng"] = (B_UPDATE + RNG_BYTES) * N / (ms[0] * 1e-3) / 1e9
    return {"N": N, "side": round(side, 4), "steps": steps, "rate": N * steps / dt,
            "cells": f"{geo.ncell_big}^3 big + {geo.ncell_sml}^3 small",
            "force_kernel": shape, "kernels": kern, "gder": gder}


def host_info():
    model = "?"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(plain, budget_s=10.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    host = host_info()
    allowed = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(allowed)})  # one pinned core
    try:
        out = _cpu_baseline(orc, plain, budget_s)
    finally:
        os.sched_setaffinity(0, allowed)
    out["host"] = host
    return out


def _cpu_baseline(orc, plain, budget_s):
    o = orc.System(plain["params"])
    o.set_positions(plain["x"])
    o.cell_locate()
    o.compute_forces()
    k = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.step(1)
        k += 1
    dt = time.perf_counter() - t0
    o.close()
    return {"value": plain["N"] * k / dt, "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"FP64 oracle, N=32768 phi=0.58, {k} steps after relaxation "
                      f"({dt:.1f} s, one thread pinned to one core)"}


def _oracle_replica(task):
    """One replica of the batch on the CPU oracle (a pool worker)."""
    nb, ns, side, T, seed, x, steps = task
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    o = orc.System(make_params(nb, ns, side=side, dt=1e-4, T0=T, Tf=T, seed=seed))
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return (nb + ns) * steps, dt


def cpu_baseline_batch(pool, workers, batch, replicas=64, steps=150):
    """The oracle over the first `replicas` replicas of the batch, one replica
    per task on a pool of host processes (all the cores given to the job)."""
    tasks = [(nb, ns, side, float(batch["T"][r]), 90000 + r, batch["x"][r], steps)
             for r, (nb, ns, side, _, _, _) in enumerate(batch["specs"][:replicas])]
    t0 = time.perf_counter()
    out = pool.map(_oracle_replica, tasks, chunksize=1)
    wall = time.perf_counter() - t0
    work = sum(w for w, _ in out)
    return {"value": work / wall, "unit": "particle-steps/s", "cores": workers, "kind": "port",
            "sample": f"FP64 oracle, first {replicas} replicas of the 256-replica batch "
                      f"(sum N = {sum(t[0] + t[1] for t in tasks)}), {steps} steps each, one "
                      f"process per replica on {workers} host processes ({wall:.1f} s)",
            "host": host_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stress-steps", type=int, default=1000,
                    help="steps of the N=262144 stress configuration (0 skips it)")
    ap.add_argument("--persistent-steps", type=int, default=200,
                    help="steps of the persistent-kernel comparison on the plain system "
                         "(0 skips it)")
    ap.add_argument("--shape", default=None,
                    help="force kernel as BRICK,DEPTH[,WALK[,SPLIT]] (e.g. 0,16,3,2) instead of "
                         "autotuning")
    ap.add_argument("--graph-steps", type=int, default=256,
                    help="largest number of steps per captured HIP graph (the engine's "
                         "default is 16); a timed call of K steps uses graphs of "
                         "min(this, K - 1) steps, built before the timed region; 0 launches "
                         "every kernel eagerly (same kernels, for kernel tracing)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the runs, copy a 1 GiB buffer 3 times (a plain copy kernel of "
                         "known traffic, the scale for FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: replicas shard across GPUs with no per-step "
              "exchange, so N>1 is not run here", file=sys.stderr)
        sys.exit(2)
    shape = tuple(int(v) for v in a.shape.split(",")) if a.shape else None
    pool, workers = None, 0
    if not a.no_cpu and not a.no_batch:
        # forked before the GPU is touched: the workers run only the CPU oracle
        import multiprocessing as mp
        workers = max(1, min(16, os.cpu_count() or 1))
        pool = mp.get_context("fork").Pool(workers)
    plain = run_plain(a.steps, a.warmup, a.graph_steps, shape, a.persistent_steps)
    log(f"[bench] plain: {plain['rate']:.4e} particle-steps/s, kernel ms {plain['kernel_ms']}")
    batch = None
    if not a.no_batch:
        batch = run_batch(a.batch_steps, min(a.warmup, 500), a.graph_steps, shape)
        log(f"[bench] batch: {batch['rate']:.4e} particle-steps/s")
    stress = None
    if a.stress_steps > 0:
        stress = run_stress(a.stress_steps, 100, a.graph_steps, shape)
        log(f"[bench] stress: {stress['rate']:.4e} particle-steps/s")
    cpu = None if a.no_cpu else cpu_baseline(plain)
    cpu_batch = None
    if pool is not None and batch:
        cpu_batch = cpu_baseline_batch(pool, workers, batch)
        pool.close()
        pool.join()
        log(f"[bench] cpu batch baseline: {cpu_batch['value']:.4e} particle-steps/s")
    N = plain["N"]
    # the dominant kernel as the timed steps run it: the force kernel, which
    # with the fused update also moves the particles (forces + update bytes)
    kname = {0: "k_forces", 1: "k_forces_pair", 2: "k_forces_quad"}[plain["shape"]["pass_split"]]
    if plain["fused_ms"] is not None:
        f_ms, f_bytes, traffic = plain["fused_ms"][1], B_FORCES + B_UPDATE, TRAFFIC_FUSED_PLAIN
        kname += " (forces + fused update)"
    else:
        f_ms, f_bytes, traffic = plain["kernel_ms"][2], B_FORCES, TRAFFIC_FORCES_PLAIN
    achieved = f_bytes * N / (f_ms * 1e-3)
    copy_gbs = stream_bandwidth()
    log(f"[bench] device copy kernel: {copy_gbs:.0f} GB/s")
    # an almost empty kernel, timed as the kernel times are: the per-launch
    # part of every kernel time (one workgroup; the plain force kernel's shape)
    floor = {"1x64": launch_floor(1, 64), "512x256": launch_floor(512, 256)}
    log(f"[bench] launch floor us: {floor}")
    # the normal equations' multiply-adds: 3 axes x the (bins + 1)-column
    # augmented lower triangle x N particles
    cv3d_macs = 3 * (CV3D_BINS + 1) * (CV3D_BINS + 2) // 2 * N
    out = {
        "metric": "particle-steps/s at N=32768 phi=0.58 (plain single system)",
        "value": plain["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": plain["seconds"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded random placement + overlap removal, then BD",
        "config": {"workload": "aging quench N=32768 phi=0.58 WCA sigma 1.4/1.0, dt=1e-4, "
                               "T 1.0 -> 0.1", "N": N, "side": round(plain["side"], 4),
                   "potential": "wca", "cells": "23^3 big + 32^3 small, ntags 8",
                   "force_kernel": plain["shape"], "launch_mode": plain["mode"]},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "traffic_unit": "B per launch (rocprofv3)",
                     "bytes_per_particle": f_bytes,
                     "avg_launch_ms": f_ms,
                     "peak_measured_copy": copy_gbs, "frac_of_measured": achieved / 1e9 / copy_gbs},
        "kernel_ms": {"update": plain["kernel_ms"][0], "cells": plain["kernel_ms"][1],
                      "forces": plain["kernel_ms"][2],
                      "forces_with_update": plain["fused_ms"][1] if plain["fused_ms"] else None},
        "launch_floor_us": floor,
        "step_hbm_frac": B_STEP * plain["rate"] / HBM_PEAK,
        "gder": plain["gder"],
        "persistent": plain["persistent"],
        "cv3d": {"bins": CV3D_BINS, "range": CV3D_RANGE, "ms_per_sample": plain["cv3d_s"] * 1e3,
                 "pairs_per_s": N * (N - 1) / plain["cv3d_s"],
                 "normal_equation_macs": cv3d_macs},
        "cpu_baseline": cpu,
    }
    if stress:
        out["stress"] = stress
    if batch:
        out["batch"] = {"value": batch["rate"], "unit": "particle-steps/s",
                        "replicas": batch["replicas"], "sum_N": batch["sum_N"],
                        "steps": a.batch_steps, "ratio_vs_plain": batch["rate"] / plain["rate"],
                        "step_hbm_frac": B_STEP * batch["rate"] / HBM_PEAK,
                        "kernel_ms": dict(zip(["update", "cells", "forces"], batch["kernel_ms"])),
                        "forces_hbm_GB_s": B_FORCES * batch["sum_N"] / (batch["kernel_ms"][2] * 1e-3) / 1e9,
                        "forces_traffic": TRAFFIC_FORCES_BATCH,
                        "cpu_baseline": cpu_batch,
                        "force_kernel": batch["shape"], "launch_mode": batch["mode"]}
    print(json.dumps(out), flush=True)
    if a.copy_calibration:
        copy_calibration()


def copy_calibration(nbytes=1 << 30, reps=3):
    """Device-to-device hipMemcpy of a buffer larger than the L2 and the
    Infinity Cache (the runtime's blit kernel reads nbytes and writes nbytes
    per call), through the HIP runtime the engine library itself links."""
    import ctypes as C
    from bd_aging_amd.abi import load
    lib = load()
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if lib.hipMalloc(C.byref(src), nbytes) or lib.hipMalloc(C.byref(dst), nbytes):
        raise RuntimeError("hipMalloc failed for the copy calibration")
    rc = lib.hipMemset(src, 1, nbytes)
    for _ in range(reps):
        rc = rc or lib.hipMemcpy(dst, src, nbytes, 3)  # hipMemcpyDeviceToDevice
    rc = rc or lib.hipDeviceSynchronize()
    lib.hipFree(src)
    lib.hipFree(dst)
    if rc:
        raise RuntimeError(f"copy calibration failed (hip error {rc})")
    log(f"[bench] copy calibration: {reps} x {nbytes} B read + {nbytes} B written")

if __name__ == "__main__":
    main()
