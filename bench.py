#!/usr/bin/env python3
"""bench.py — Lennard-Jones pair-interaction throughput on one MI355X.

Workload (the headline size of the reference's force path): FCC lattice of
4*31^3 = 119,164 particles at density 1.0 in a periodic cubic box
(L = 49.21), seeded jitter U(-0.05, 0.05), cutoff 3.0, search length 3.3,
full Verlet list (16,895,170 entries = 8,447,585 unique pairs), dt = 0.01,
AoS double4 positions and momenta resident in HBM. One "step" is one force
evaluation p += F(q) * dt over the whole list (the body of the reference's
measure() loop, cuda/force_cuda.cu:333-335).

Prints ONE JSON line on stdout:
  value          unique pairs * steps / GPU time of the FP64 kernel (Gpair/s):
                 the median of R timed calls of exactly K steps, each between
                 HIP events on the library's stream, after W warm-up steps, one
                 untimed call of K steps and a short untimed settling period
  mixed          the same for the mixed-precision kernel, with its force error
                 against the FP64 kernel (max and RMS, normwise)
  roofline       algorithmic bytes per launch / median launch time, vs 8 TB/s
  cpu_baseline   the reference's force_sorted (half list, j scatter) restated
                 in oracle/cpu_baseline.cpp (-O3, AVX2/FMA), on one host core,
                 a bounded sample of the same list; cpu_baselines_other holds
                 the restated force_intrin_mat_transpose (one core) and an
                 all-core OpenMP full-list loop

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--repeats R]
       python bench.py --workload {large,replicas,md} [--steps K]
The other workloads print one JSON line each in the same format: "large" the
same evaluation on FCC 4*64^3 = 1,048,576 particles with FP64, mixed and
Kahan-mixed kernels; "replicas" 256 periodic boxes of uneven size in one
autotuned launch per step; "md" K velocity-Verlet steps of 8,388,608
particles at density 0.8 with device list rebuilds whenever a particle has
moved half the skin.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpair-interactions/s FP64 vs mixed-FP32 at ≤1e-6 rel force err, N=119164"
HBM_PEAK_GBS = 8000.0
FLOP_PER_ENTRY = 22  # sub 3, r2 5, r6 2, r14 2, numerator 2, divide 1, dt 1, accumulate 6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lj_force_f64_N119164_rho1_traffic.json")


def measured_traffic(info, tile_order):
    """HBM bytes per launch of the FP64 force kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs (copy-calibrated) when they were
    taken with the kernel configuration in use; None otherwise."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    cfg = t.get("kernel_config", {})
    if (cfg.get("lanes_per_particle"), cfg.get("unroll"), cfg.get("tile_particles"),
            cfg.get("tile_order")) != (info.lanes_per_particle, info.unroll,
                                       info.tile_particles, tile_order):
        return None
    return t.get("hbm_bytes_per_launch")


def cpu_baselines(q, box, half, full, dt, budget_s: float) -> list[dict]:
    """The reference's CPU force path restated (oracle/cpu_baseline.cpp, -O3
    AVX2/FMA), timed on this host over the same periodic lists, each on a
    bounded sample of steps (about `budget_s` seconds): force_sorted on one
    core (the reference's "Reference" row), force_intrin_mat_transpose on
    one core (its fastest CPU routine) and the full list over all cores with
    OpenMP (labelled; the reference has no threaded CPU path)."""
    import numpy as np

    import oracle as O

    h = O.CSR(half.number_of_partners, half.pointer, half.sorted_list, "half")
    f = O.CSR(full.number_of_partners, full.pointer, full.sorted_list, "full")
    qa = O.aligned_zeros(q.shape)
    qa[...] = q
    L = O.baseline_lib()
    build = L.cb_build_info().decode()
    threads = int(L.cb_threads())
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    out = []
    for kind, csr, name, cores, ref in (
            (O.BASELINE_SORTED, h, "force_sorted", 1, "cpu_ref/force_aos.cpp:192-224"),
            (O.BASELINE_MAT_TRANSPOSE, h, "force_intrin_mat_transpose", 1,
             "cpu_ref/force_aos.cpp:413-500"),
            (O.BASELINE_FULL_OMP, f, "full list, OpenMP over rows", threads,
             "cpu_ref/force_aos.cpp:192-224 on the full list")):
        p = O.aligned_zeros(q.shape)
        t1 = O.run_baseline(kind, csr, qa, p, dt, O.CL2, box, 1)
        steps = max(1, min(1000, int(budget_s / max(t1, 1e-6))))
        t = O.run_baseline(kind, csr, qa, p, dt, O.CL2, box, steps)
        rate = half.n_entries * steps / t / 1e9
        out.append({
            "value": round(rate, 6), "unit": "Gpair-interactions/s", "cores": cores,
            "kind": "port", "routine": name,
            "sample": f"{name} ({ref} restated, minimum image) over the same N={half.n} "
                      f"{'half' if csr is h else 'full'} list ({half.n_entries} unique pairs), "
                      f"{steps} steps in {t:.2f} s, {cores} thread(s)",
            "cpu": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "build": build, "seconds": round(t, 3), "steps": steps})
        log(f"cpu baseline {name}: {rate:.4f} Gpair/s ({cores} thread(s))")
    return out


def time_leg(ctx, fn, steps: int, warmup: int, repeats: int, settle_ms: float) -> dict:
    """Time `fn(steps)` (one force call of `steps` launches) the way the
    reference's measure() times its launch loop (cuda/force_cuda.cu:330-336):
    only the launches, after every one-time cost. `warmup` untimed steps,
    then one untimed call of exactly `steps` steps (so the graphs and buffers
    the timed calls use already exist), then untimed calls of `steps` until
    `settle_ms` of GPU time has passed (clocks settle), then `repeats` timed
    calls, each bracketed by HIP events on the library's stream. Returns the
    per-step times of the repeats (ms) and their median, min and max."""
    def timed():
        ctx.sync()
        ctx.timer_start()
        fn(steps)
        ms = ctx.timer_stop()
        ctx.sync()
        return ms

    if warmup > 0:
        fn(warmup)
    fn(steps)
    ctx.sync()
    settled = 0.0
    while settled < settle_ms:
        settled += timed()
    per_step = sorted(timed() / steps for _ in range(repeats))
    med = per_step[len(per_step) // 2] if len(per_step) % 2 else \
        0.5 * (per_step[len(per_step) // 2 - 1] + per_step[len(per_step) // 2])
    return {"ms_per_step": med, "min_ms_per_step": per_step[0], "max_ms_per_step": per_step[-1],
            "repeats_ms_per_step": [round(v, 7) for v in per_step], "settle_ms": round(settled, 1)}


def bench_large(args, G, np) -> dict:
    """FP64 and mixed force evaluation on FCC 4*64^3 = 1,048,576 particles
    at density 1.0 (periodic, full list)."""
    dt, k = 0.01, 64
    q, box = G.lattice_periodic(k, 1.0, 2)
    ctx = G.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], box, G.LIST_FULL)
    info = ctx.list_info()
    e, n = info.n_entries, info.particle_number
    dp = ctx.zeros(q.shape)
    out = {}
    for name, fn in (("f64", lambda s: ctx.force_aos_f64(dq, dp, dt, G.CL2, box, s)),
                     ("mixed_f64acc", lambda s: ctx.force_aos_mixed(dq, dp, dt, G.CL2, box, s,
                                                                    G.ACCUM_F64)),
                     ("mixed_kahan", lambda s: ctx.force_aos_mixed(dq, dp, dt, G.CL2, box, s,
                                                                   G.ACCUM_KAHAN_F32))):
        t = time_leg(ctx, fn, args.steps, args.warmup, args.repeats, args.settle_ms)
        out[name] = {"ms_per_step": t["ms_per_step"], "gpair_s": e / 2 / t["ms_per_step"] / 1e6,
                     "timing": t}
    errs = {}
    dF = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dF, 1.0, G.CL2, box, 1)
    F = dF.download()[:, :3]
    for acc in (G.ACCUM_F64, G.ACCUM_KAHAN_F32):
        dM = ctx.zeros(q.shape)
        ctx.force_aos_mixed(dq, dM, 1.0, G.CL2, box, 1, acc)
        M = dM.download()[:, :3]
        sc = float(np.abs(F).max())
        errs[acc] = (float(np.abs(M - F).max() / sc), float(np.sqrt(((M - F) ** 2).mean()) / sc))
    ctx.close()
    ms = out["f64"]["ms_per_step"]
    return {
        "metric": METRIC, "value": round(out["f64"]["gpair_s"], 3), "unit": "Gpair-interactions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 6),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"LJ force, FCC 4*{k}^3={n} particles, density 1.0, periodic box "
                               f"{box:.4f}, full Verlet list built on the device, dt {dt}, AoS",
                   "particles": n, "list_entries": e, "unique_pairs": e // 2,
                   "lanes_per_particle": info.lanes_per_particle, "unroll": info.unroll,
                   "tile_particles": info.tile_particles},
        "roofline": roofline(4 * e + 104 * n, ms / 1e3),
        "mixed": {"f64acc": {"value": round(out["mixed_f64acc"]["gpair_s"], 3),
                             "speedup_vs_f64": round(ms / out["mixed_f64acc"]["ms_per_step"], 4),
                             "max_rel_force_err_vs_f64": errs[G.ACCUM_F64][0],
                             "rms_rel_force_err_vs_f64": errs[G.ACCUM_F64][1]},
                  "kahan": {"value": round(out["mixed_kahan"]["gpair_s"], 3),
                            "speedup_vs_f64": round(ms / out["mixed_kahan"]["ms_per_step"], 4),
                            "max_rel_force_err_vs_f64": errs[G.ACCUM_KAHAN_F32][0],
                            "rms_rel_force_err_vs_f64": errs[G.ACCUM_KAHAN_F32][1]}},
        "cpu_baseline": None,
    }


def bench_replicas(args, G, np) -> dict:
    """256 periodic replica boxes (FCC 4*k^3, k uniform in 6..20, density
    uniform in [0.5, 1.0]) in one tiled launch per step, configuration from
    the autotuner, 100-step loops replayed as HIP graphs."""
    dt = 0.01
    q, off, boxes = G.replica_workload(256, seed=4)
    ctx = G.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    ctx.set_replicas(off, boxes)
    dq = ctx.array(q)
    ctx.build_list(dq, q.shape[0], 0.0, G.LIST_FULL)
    best, table = ctx.autotune(dq, 0.0, reps=20)
    info = ctx.list_info()
    e, n = info.n_entries, info.particle_number
    dp = ctx.zeros(q.shape)
    res = {}
    for name, fn in (("f64", lambda s: ctx.force_aos_f64(dq, dp, dt, G.CL2, 0.0, s)),
                     ("mixed", lambda s: ctx.force_aos_mixed(dq, dp, dt, G.CL2, 0.0, s,
                                                             G.ACCUM_F64))):
        res[name] = time_leg(ctx, fn, args.steps, args.warmup, args.repeats,
                             args.settle_ms)["ms_per_step"]
    ctx.close()
    ms = res["f64"]
    return {
        "metric": METRIC, "value": round(e / 2 / (ms / 1e3) / 1e9, 3),
        "unit": "Gpair-interactions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 6), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"256 periodic LJ replicas, FCC 4*k^3 with k in 6..20, density in "
                               f"[0.5, 1], {n} particles in total, one tiled launch per step, "
                               f"dt {dt}, AoS", "particles": n, "replicas": 256,
                   "list_entries": e, "unique_pairs": e // 2, "autotuned": best},
        "roofline": roofline(4 * e + 104 * n, ms / 1e3),
        "mixed": {"value": round(e / 2 / (res["mixed"] / 1e3) / 1e9, 3),
                  "speedup_vs_f64": round(ms / res["mixed"], 4)},
        "autotune_table": table,
        "cpu_baseline": None,
    }


def bench_md(args, G, np) -> dict:
    """Velocity Verlet on FCC 4*128^3 = 8,388,608 particles at density 0.8,
    Maxwell-Boltzmann momenta at T = 1 (seed 7), dt 0.005, mixed-precision
    forces, the list rebuilt on the device every --md-every steps (skin
    --md-skin), or, with --md-every 0 (default, skin 0.8), whenever a
    particle has moved half the skin since the last build, so no step runs
    on a stale list.
    Throughput is particle-steps per second. Beside it: pairs inside the
    cutoff per second (the pair count of the start and end states, averaged,
    times the steps, over the run's GPU time) and list entries per second
    (which grow with the skin)."""
    q, box = G.lattice_periodic(128, 0.8, 2)
    p = G.maxwell_boltzmann(q.shape[0], 1.0, 7)
    ctx = G.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    dq, dp = ctx.array(q), ctx.array(p)
    n = q.shape[0]

    def pairs_in_cutoff():
        ctx.build_list(dq, n, box, G.LIST_FULL)
        return ctx.energy_pairs(dq, dp, box)[2]

    pairs0 = pairs_in_cutoff()
    steps, skin, every = args.steps, args.md_skin, args.md_every
    st, samples = ctx.md_run(dq, dp, box, steps, 0.005, skin=skin, rebuild_every=every,
                             precision=G.PREC_MIXED, sample_every=max(steps // 10, 1))
    pairs1 = pairs_in_cutoff()
    ctx.close()
    total_s = st.ms_total / 1e3
    policy = (f"every {every} steps" if every > 0 else
              "when a particle has moved skin/2")
    return {
        "metric": "particle-timesteps/s, LJ velocity-Verlet MD", "value":
            round(n * steps / total_s / 1e9, 4),
        "unit": "Gparticle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(st.ms_total / max(steps, 1), 6), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64-mixed", "data": "synthetic",
        "config": {"workload": f"LJ velocity Verlet, FCC 4*128^3={n} particles, density 0.8, "
                               f"periodic box {box:.4f}, T=1 Maxwell-Boltzmann (seed 7), dt 0.005, "
                               f"device list rebuilt {policy} (skin {skin}), mixed forces, FP64 "
                               f"integration", "particles": n, "rebuild_every": every,
                   "skin": skin},
        "pairs_in_cutoff_per_s": round(0.5 * (pairs0 + pairs1) * steps / total_s / 1e9, 3),
        "pairs_in_cutoff": [pairs0, pairs1],
        "list_entries_per_s": round(st.pair_entries / total_s / 1e9, 3),
        "md": {"rebuilds": st.rebuilds, "ms_rebuild": st.ms_rebuild, "ms_total": st.ms_total,
               "ms_per_rebuild": st.ms_rebuild / max(st.rebuilds, 1),
               "stale_lists": st.stale_lists, "max_displacement": st.max_displacement,
               "energy_start": st.potential0 + st.kinetic0,
               "energy_end": st.potential1 + st.kinetic1,
               "max_abs_energy_drift": st.max_abs_drift,
               "relative_energy_drift": st.max_abs_drift / abs(st.potential0 + st.kinetic0),
               "samples": samples.tolist()},
        "cpu_baseline": None,
    }


def bench_reference_mode(args, G, np) -> dict:
    """The reference's own problem (cpu_ref/force_aos.cpp:43-51 init(): open
    box L = 50, density 0.5, U(0, 0.1) jitter, 62,500 particles; dt 0.001),
    FP64 over its full list, timed like the headline: 100-step calls, median
    of the repeats. The step the reference's measure() loop times."""
    dt = 0.001
    q = G.lattice_reference(0.5, 50.0)
    full = G.makepair(q, 0.0, G.LIST_FULL)
    ctx = G.Context(0)
    ctx.set_list(full)
    dq, dp = ctx.array(q), ctx.zeros(q.shape)
    t = time_leg(ctx, lambda s: ctx.force_aos_f64(dq, dp, dt, G.CL2, 0.0, s), 100, 100,
                 args.repeats, 50.0)
    ctx.close()
    pairs = full.n_entries // 2
    return {"workload": f"reference problem: open box L=50, density 0.5, N={full.n}, full list "
                        f"({full.n_entries} entries, {pairs} unique pairs), dt {dt}, AoS double4",
            "ms_per_step": round(t["ms_per_step"], 7),
            "value": round(pairs / t["ms_per_step"] / 1e6, 3), "unit": "Gpair-interactions/s",
            "steps_per_call": 100, "repeats_ms_per_step": t["repeats_ms_per_step"]}


def roofline(bytes_per_launch: int, seconds: float, traffic=None) -> dict:
    achieved = bytes_per_launch / seconds / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(seconds * 1e6, 3)}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps per repeat (default 1000; 1000 for the md workload)")
    ap.add_argument("--warmup", type=int, default=100, help="untimed warm-up steps")
    ap.add_argument("--repeats", type=int, default=7,
                    help="timed calls of --steps steps; the median is reported (at least 5)")
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed calls of --steps steps before timing, until this much GPU "
                         "time has passed (clock settling)")
    ap.add_argument("--cells", type=int, default=31)
    ap.add_argument("--density", type=float, default=1.0)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per particle (0: library default)")
    ap.add_argument("--block", type=int, default=0, help="workgroup size (0: library default)")
    ap.add_argument("--unroll", type=int, default=0, help="entries per lane per batch (0: default)")
    ap.add_argument("--tile", type=int, default=-1, help="particles per LDS tile (-1: default)")
    ap.add_argument("--chains", type=int, default=0,
                    help="concurrent launch chains over the tiles (0: library default)")
    ap.add_argument("--tile-order", choices=["spatial", "index"], default="spatial",
                    help="order of the particles in LDS tiles")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="seconds of CPU work per CPU baseline (three baselines)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "f64", "mixed"], default="all")
    ap.add_argument("--workload", choices=["force", "large", "replicas", "md"], default="force",
                    help="force: the headline N=119,164 force evaluation (default); large: "
                         "N=1,048,576; replicas: 256 batched boxes; md: 8.4M-particle velocity "
                         "Verlet with device list rebuilds")
    ap.add_argument("--md-skin", type=float, default=0.8,
                    help="md workload: list skin (search length = cutoff + skin)")
    ap.add_argument("--md-every", type=int, default=0,
                    help="md workload: rebuild the list every this many steps (0: whenever a "
                         "particle has moved skin/2 since the last build)")
    ap.add_argument("--autotune", action="store_true",
                    help="force workload: run the autotuner (20-step calls) before timing, "
                         "time its pick and report its table")
    ap.add_argument("--no-reference-mode", action="store_true",
                    help="force workload: skip the timing of the reference's own open-box "
                         "problem")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every step individually instead of replaying captured graphs "
                         "(for per-dispatch profiler counters)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after timing, run three device-to-device copies of this many MiB "
                         "(a known-size kernel that calibrates rocprofv3's byte counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 1000
    args.repeats = max(args.repeats, 5)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the force path is measured on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import lj_gpu_amd as G

    if args.workload != "force":
        fn = {"large": bench_large, "replicas": bench_replicas, "md": bench_md}[args.workload]
        print(json.dumps(fn(args, G, np)), flush=True)
        return 0

    dt = 0.01
    t0 = time.time()
    q, box = G.lattice_periodic(args.cells, args.density, 2)
    full = G.makepair(q, box, G.LIST_FULL)
    half = G.makepair(q, box, G.LIST_HALF)
    n, e = full.n, full.n_entries
    pairs = e // 2
    log(f"setup: N={n} box={box:.4f} full entries={e} pairs={pairs} ({time.time() - t0:.1f}s)")

    ctx = G.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    ctx.set_kernel_config(args.lanes, args.block, args.unroll)
    if args.tile >= 0:
        ctx.set_tiling(args.tile)
    ctx.set_tile_order(G.TILE_ORDER_INDEX if args.tile_order == "index" else G.TILE_ORDER_SPATIAL)
    if args.chains:
        ctx.set_chains(args.chains)
    ctx.set_list(full)
    dq = ctx.array(q)
    dp = ctx.zeros(q.shape)

    tune = None
    if args.autotune:
        t1 = time.time()
        best, table = ctx.autotune(dq, box, reps=20)
        tune = {"best": best, "table": table, "seconds": round(time.time() - t1, 2)}
        log(f"autotune: {best} ({tune['seconds']} s)")

    f64 = lambda s: ctx.force_aos_f64(dq, dp, dt, G.CL2, box, s)
    mixed = lambda s: ctx.force_aos_mixed(dq, dp, dt, G.CL2, box, s, G.ACCUM_F64)

    t_f64 = t_mixed = None
    if args.only in ("all", "f64"):
        t_f64 = time_leg(ctx, f64, args.steps, args.warmup, args.repeats, args.settle_ms)
        log(f"fp64: {args.steps} steps, median {t_f64['ms_per_step'] * 1e3:.3f} us/step "
            f"(min {t_f64['min_ms_per_step'] * 1e3:.3f}, max {t_f64['max_ms_per_step'] * 1e3:.3f})")
    if args.only in ("all", "mixed"):
        t_mixed = time_leg(ctx, mixed, args.steps, args.warmup, args.repeats, args.settle_ms)
        log(f"mixed: {args.steps} steps, median {t_mixed['ms_per_step'] * 1e3:.3f} us/step "
            f"(min {t_mixed['min_ms_per_step'] * 1e3:.3f}, "
            f"max {t_mixed['max_ms_per_step'] * 1e3:.3f})")
    # The tiling in use (spatially ordered at the first evaluation).
    info = ctx.list_info()
    tile_order = ("spatial" if ctx.get_tile_order()[0] == G.TILE_ORDER_SPATIAL else "index")

    # Force error of the mixed kernel against the FP64 kernel (one step, dt = 1).
    dF = ctx.zeros(q.shape)
    dM = ctx.zeros(q.shape)
    ctx.force_aos_f64(dq, dF, 1.0, G.CL2, box, 1)
    ctx.force_aos_mixed(dq, dM, 1.0, G.CL2, box, 1, G.ACCUM_F64)
    ctx.sync()
    F, M = dF.download()[:, :3], dM.download()[:, :3]
    scale = float(np.abs(F).max())
    err_max = float(np.abs(M - F).max() / scale)
    err_rms = float(np.sqrt(((M - F) ** 2).mean()) / scale)

    bytes_per_launch = 4 * e + 104 * n  # list + (nop, pointer) + q + p read/write, AoS double4
    main_t = t_f64 if t_f64 is not None else t_mixed
    per_launch_s = main_t["ms_per_step"] / 1e3
    value = pairs / per_launch_s / 1e9
    achieved = bytes_per_launch / per_launch_s / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_t["ms_per_step"], 7),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if t_f64 is not None else "f32/f64-mixed",
        "data": "synthetic",
        "config": {
            "workload": f"LJ force, FCC 4*{args.cells}^3={n} particles, density {args.density}, "
                        f"periodic box {box:.4f}, full Verlet list (cutoff 3.0, search 3.3), "
                        f"dt {dt}, AoS double4",
            "particles": n, "density": args.density, "box": round(box, 6),
            "list_entries": e, "unique_pairs": pairs,
            "lanes_per_particle": info.lanes_per_particle,
            "unroll": info.unroll,
            "tile_particles": info.tile_particles,
            "tile_max_stage": info.tile_max_stage,
            "tile_order": tile_order,
        },
        "timing": {
            "method": f"HIP events on the library stream around each of {args.repeats} timed "
                      f"calls of {args.steps} steps; median reported. Untimed before them: "
                      f"{args.warmup} warm-up steps, one call of {args.steps} steps, then "
                      f"calls of {args.steps} steps for {main_t['settle_ms']} ms",
            "median_ms_per_step": round(main_t["ms_per_step"], 7),
            "min_ms_per_step": round(main_t["min_ms_per_step"], 7),
            "max_ms_per_step": round(main_t["max_ms_per_step"], 7),
            "repeats_ms_per_step": main_t["repeats_ms_per_step"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(info, tile_order),
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": round(per_launch_s * 1e6, 3),
            "flop_rate_tflops": round(FLOP_PER_ENTRY * e / per_launch_s / 1e12, 3),
        },
        "mixed": None if t_mixed is None else {
            "value": round(pairs / (t_mixed["ms_per_step"] / 1e3) / 1e9, 3),
            "ms_per_step": round(t_mixed["ms_per_step"], 7),
            "min_ms_per_step": round(t_mixed["min_ms_per_step"], 7),
            "max_ms_per_step": round(t_mixed["max_ms_per_step"], 7),
            "speedup_vs_f64": None if t_f64 is None else
            round(t_f64["ms_per_step"] / t_mixed["ms_per_step"], 4),
            "accum": "f64",
            "max_rel_force_err_vs_f64": err_max,
            "rms_rel_force_err_vs_f64": err_rms,
        },
    }
    if tune is not None:
        result["autotune"] = tune
    # The reference's own problem, in the default runs only (the profiling
    # command lines pass --only, --chains or --no-graphs and keep their traces
    # to the headline kernel).
    if not args.no_reference_mode and args.only == "all" and not args.no_graphs and \
            args.chains == 0:
        result["reference_mode"] = bench_reference_mode(args, G, np)
    if args.copy_mib > 0:
        nbytes = args.copy_mib << 20
        a = ctx.zeros(nbytes // 8)
        b = ctx.zeros(nbytes // 8)
        for _ in range(3):
            G.lib().ljgpu_memcpy_dtod(ctx.handle, b.ptr, a.ptr, nbytes)
        ctx.sync()
        a.free()
        b.free()
    if not args.no_cpu:
        base = cpu_baselines(q, box, half, full, dt, args.cpu_seconds)
        result["cpu_baseline"] = base[0]
        result["cpu_baselines_other"] = base[1:]
    else:
        result["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
