#!/usr/bin/env python3
"""Throughput benchmark of the HIP PBF backend (one MI355X).

Default workload: the 1,048,576-particle block (128x128x64 lattice, d=0.05,
+-0.1d seeded jitter) in [0,12]x[0,8]x[0,4], h=0.1, rho0=1000, m=0.125,
eps=100, 4 constraint iterations, dt=1/120, s_corr + XSPH + vorticity,
restitution 0.05 / friction 0.1; W warm-up steps, then K timed steps with
the state resident in HBM.  Prints one JSON line.

    python bench.py [--steps 120] [--warmup 30] [--no-cpu] [--graph]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fluidsimulator_amd import hip, scenes  # noqa: E402

METRIC = "1M-particle steps/s (4 iters) + % of HBM peak; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle and launch (compulsory FP32 SoA
# traffic, neighbor data assumed reused from cache; neighbor-index streams
# not counted).
KERNEL_BYTES = {
    "predict_bbox": 40.0,      # pos 12 + vel 12 -> pred 12 + key 4
    "lambda": 16.0,            # pred 12 -> lambda 4
    "delta": 28.0,             # pred 12 + lambda 4 -> pred' 12 (last one also commits)
    "xsph": 40.0,              # pos 12 + vel 12 + rho 4 -> vel 12
    "vorticity_omega": 40.0,   # pos 12 + vel 12 -> omega 12 + |omega| 4
    "vorticity_apply": 52.0,   # pos 12 + omega 12 + |omega| 4 + vel 12 -> vel 12
}
STEP_BYTES_ALL_FLAGS = 480.0   # whole substep, all flags, 4 iterations
# Algorithmic FP32 operations per neighbor pair (add, mul, div, sqrt and
# compare-select each count one; loop invariants hoisted): lambda 33, delta
# with s_corr 30, XSPH 25, omega 31, eta and confinement 23.
FLOPS_PER_PAIR = {"lambda": 33.0, "delta": 30.0, "xsph": 25.0, "vorticity_omega": 31.0,
                  "vorticity_apply": 23.0}
FP32_PEAK_TFLOPS = 157.3
KERNELS = ["predict_bbox", "grid_setup", "cell_keys", "sort", "cell_count", "reorder", "tiles", "runs",
           "neighbors", "lambda", "delta", "xsph", "vorticity_omega", "vorticity_apply",
           "tile_neighbors", "tile_lambda", "tile_delta", "tile_xsph", "tile_vorticity_omega",
           "tile_vorticity_apply"]


def sync_all():
    """Device-wide synchronize (the library stream is synchronized by
    HipSolver.sync(); torch is only touched if the caller imported it)."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()


TRAFFIC_GLOB = "traffic_million_block_*.json"


def measured_traffic(kernel: str, run: dict):
    """HBM bytes per launch and L2 hit rate of `kernel` from a committed
    rocprofv3 counter summary (profiles/traffic_million_block_*.json) whose
    workload, window (warm-up, timed substeps) and launch configuration are
    this run's.  Returns (bytes or None, hit rate or None, note)."""
    import glob

    wanted = {"dims": list(run["dims"]), "warmup": run["warmup"], "steps": run["steps"],
              "config": run["config"]}
    seen = []
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", TRAFFIC_GLOB))):
        with open(path) as f:
            data = json.load(f)
        rec = data.get("run", {})
        got = {"dims": rec.get("dims"), "warmup": rec.get("warmup"), "steps": rec.get("steps"),
               "config": rec.get("config")}
        seen.append(os.path.basename(path))
        if got != wanted:
            continue
        k = data.get("kernels", {}).get(kernel)
        if k is None:
            return None, None, f"{os.path.basename(path)} has no {kernel} record"
        return (float(k["hbm_bytes_per_launch"]), k.get("l2_hit_rate"),
                f"profiles/{os.path.basename(path)}")
    return None, None, ("no committed counter summary for this window and launch "
                        f"configuration {wanted} (have: {', '.join(seen) or 'none'})")


def cpu_threads() -> int:
    """Host threads of the CPU baseline: OMP_NUM_THREADS when set (the share
    of the host that belongs to this GPU), else every core up to 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(state, params, threads, warm=1, steps=2, one_thread=True):
    """The reference's own CPU step (oracle/_ref, built from core.cpp), or the
    oracle where that library is absent, on the state the GPU timed region
    starts from: `warm` untimed steps, then `steps` timed steps on `threads`
    OpenMP threads (one timed step when a step takes over 15 s), and one
    timed step on one thread from the same state."""
    import oracle

    which = "ref" if oracle.have_ref() else "oracle"
    oracle.set_threads(threads, which)
    t0 = time.perf_counter()
    r = oracle.step(params, state.pos, state.vel, nsteps=warm, which=which)
    t_warm = (time.perf_counter() - t0) / max(warm, 1)
    if t_warm > 15.0:
        steps = 1
    t0 = time.perf_counter()
    oracle.step(params, r["pos"], r["vel"], nsteps=steps, which=which)
    dt = time.perf_counter() - t0
    kind = "reference" if which == "ref" else "port"
    name = "reference core.cpp" if which == "ref" else "oracle"
    out = {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"{name} step() on the {state.n}-particle state the GPU timed region starts "
                   f"from: {warm} untimed + {steps} timed steps on {threads} OpenMP threads "
                   f"({steps * state.n / dt:.3g} particle-steps/s)"),
        "particle_steps_per_s": steps * state.n / dt,
    }
    if one_thread:
        oracle.set_threads(1, which)
        t0 = time.perf_counter()
        oracle.step(params, r["pos"], r["vel"], nsteps=1, which=which)
        d1 = time.perf_counter() - t0
        oracle.set_threads(threads, which)
        out["one_thread"] = {"value": 1.0 / d1, "unit": "steps/s", "cores": 1,
                             "sample": "1 timed step on 1 thread from the same state",
                             "particle_steps_per_s": state.n / d1}
    return out


def copy_calibration(mbytes: int) -> None:
    """One device-to-device copy of `mbytes` MiB (the runtime's blit kernel):
    the reference point that scales FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip_rt = C.CDLL("libamdhip64.so")
    n = int(mbytes) << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip_rt.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip_rt.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip_rt.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip_rt.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # device to device
    assert hip_rt.hipDeviceSynchronize() == 0
    hip_rt.hipFree(a)
    hip_rt.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}))


# rocprofv3 kernel names of the roofline kernels, and their launches per
# substep (lambda and delta: one per constraint iteration)
ROCPROF_NAME = {"predict_bbox": "k_predict_bbox", "lambda": "k_lambda", "delta": "k_delta",
                "xsph": "k_xsph", "vorticity_omega": "k_vort_omega",
                "vorticity_apply": "k_vort_apply"}


def window_bounds(pos: list, take: int) -> tuple:
    """Bounds of the timed window on the ascending dispatch positions (start
    times or dispatch ids) of a per-lane pass kernel: halfway to its launch
    before the window and halfway past its last.  The pass's tile kernel is
    launched next to each per-lane launch, so exactly its launches of the
    window fall inside; those of another context (the early-regime record)
    or of the warm-up fall outside."""
    w = pos[-take:]
    if not w:
        return float("-inf"), float("-inf")
    lo = (pos[-take - 1] + w[0]) / 2 if len(pos) > take else float("-inf")
    hi = w[-1] + (w[-1] - w[-2]) / 2 if len(w) > 1 else float("inf")
    return lo, hi


def summarize_traffic(outdir: str, json_path: str | None = None, steps: int | None = None,
                      iterations: int = 4) -> str:
    """HBM bytes per launch and L2 hit rate of the roofline kernels from
    `rocprofv3 --pmc` runs under `outdir`: pmc_fetch (FETCH_SIZE) and
    pmc_write (WRITE_SIZE) of one bench command, pmc_l2 (TCC_HIT_sum,
    TCC_MISS_sum) of the same command when present, and pmc_cfetch /
    pmc_cwrite (the same counters for `--copy-calibration 256`).  Each
    traffic counter (KiB per dispatch) is scaled by what it reads for the copy
    of known size; every figure is averaged over the kernel's dispatches of
    the timed window (the last `steps` substeps in dispatch order).  The bench
    line the fetch run printed (pmc_fetch/bench.json) supplies the workload,
    window and launch configuration, which the JSON records so that
    `measured_traffic` only uses it for a run of the same kind.  Writes the
    JSON (default profiles/traffic_million_block_warmup<W>_steps<K>.json)
    and returns a text table."""
    import collections
    import glob
    import sqlite3

    def per_dispatch(sub, counter, required=True):
        dbs = sorted(glob.glob(os.path.join(outdir, sub, "**", "*_results.db"), recursive=True))
        if not dbs:
            if required:
                raise SystemExit(f"no rocpd database under {os.path.join(outdir, sub)}")
            return None
        rows = sqlite3.connect(dbs[0]).cursor().execute(
            "select dispatch_id, name, sum(counter_value) from pmc_events where counter_name=? "
            "group by dispatch_id order by dispatch_id", (counter,)).fetchall()
        out = collections.defaultdict(list)
        for did, name, v in rows:
            out[name].append((did, v))
        return out

    run = {}
    line_path = os.path.join(outdir, "pmc_fetch", "bench.json")
    if os.path.exists(line_path):
        with open(line_path) as f:
            line = json.loads(f.read().strip().splitlines()[-1])
        cfg = line.get("config", {})
        run = {"dims": cfg.get("dims"), "warmup": line.get("warmup"), "steps": line.get("steps"),
               "config": {k: cfg.get(k) for k in RUN_KEYS}}
        if steps is None:
            steps = int(line["steps"])
    if steps is None:
        raise SystemExit("give --steps or a pmc_fetch/bench.json bench line")
    dims = run.get("dims") or [128, 128, 64]
    n = int(np.prod(dims))

    copy_kib = 256 * 1024

    def copy_reading(d):
        vals = [v for k in d if "copyBuffer" in k for _, v in d[k]]
        return sum(vals) / len(vals)

    f_read = copy_reading(per_dispatch("pmc_cfetch", "FETCH_SIZE"))
    w_read = copy_reading(per_dispatch("pmc_cwrite", "WRITE_SIZE"))
    f_scale, w_scale = copy_kib / f_read, copy_kib / w_read
    kf = per_dispatch("pmc_fetch", "FETCH_SIZE")
    kw = per_dispatch("pmc_write", "WRITE_SIZE")
    kh = per_dispatch("pmc_l2", "TCC_HIT_sum", required=False)
    km = per_dispatch("pmc_l2", "TCC_MISS_sum", required=False)

    def window(d, key, per_step):
        """Per-launch values of pass `key` over the timed window: the per-lane
        kernel and, when the window's launches have one, its LDS-tile kernel,
        summed launch by launch."""
        if d is None:
            return []
        take = per_step * steps
        lane = sorted(x for k in d if ROCPROF_NAME[key] in k for x in d[k])
        out = [v for _, v in lane[-take:]]
        lo, hi = window_bounds([i for i, _ in lane], take)
        tname = "k_tile_" + ROCPROF_NAME[key][2:]
        tile = sorted(x for k in d if tname in k for x in d[k] if lo < x[0] < hi)
        if tile:
            if len(tile) != len(out):
                raise SystemExit(f"{key}: {len(tile)} tile launches against {len(out)} "
                                 "per-lane launches in the timed window")
            out = [a + v for a, (_, v) in zip(out, tile)]
        return out

    res = {"workload": f"million_block {dims[0]}x{dims[1]}x{dims[2]}, the timed window of "
                       "the bench command",
           "run": run,
           "copy_calibration": {"bytes": copy_kib * 1024, "fetch_kib_read": f_read,
                                "write_kib_read": w_read, "fetch_scale": f_scale,
                                "write_scale": w_scale},
           "kernels": {}}
    lines = [f"Calibration: FETCH_SIZE {f_read:.0f} KiB -> scale {f_scale:.4f}, "
             f"WRITE_SIZE {w_read:.0f} KiB -> scale {w_scale:.4f}", "",
             "kernel              fetch MB   write MB     HBM MB  algorithmic MB   ratio  L2 hit"]
    for key, per_particle in KERNEL_BYTES.items():
        per_step = iterations if key in ("lambda", "delta") else 1
        fv, wv = window(kf, key, per_step), window(kw, key, per_step)
        if not fv or not wv:
            continue
        fetch = sum(fv) / len(fv) * 1024 * f_scale
        write = sum(wv) / len(wv) * 1024 * w_scale
        hv, mv = window(kh, key, per_step), window(km, key, per_step)
        hit = (sum(hv) / (sum(hv) + sum(mv))) if hv and mv and (sum(hv) + sum(mv)) > 0 else None
        alg = per_particle * n
        res["kernels"][key] = {"fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
                               "hbm_bytes_per_launch": fetch + write,
                               "algorithmic_bytes_per_launch": alg,
                               "traffic_over_algorithmic": (fetch + write) / alg,
                               "l2_hit_rate": hit, "launches": len(fv)}
        lines.append(f"{key:<18}{fetch / 1e6:11.1f}{write / 1e6:11.1f}{(fetch + write) / 1e6:11.1f}"
                     f"{alg / 1e6:16.1f}{(fetch + write) / alg:8.2f}"
                     f"{'' if hit is None else f'{hit:8.3f}'}")
    if json_path is None:
        json_path = os.path.join(ROOT, "profiles", f"traffic_million_block_warmup"
                                 f"{run.get('warmup')}_steps{steps}.json")
    with open(json_path, "w") as f:
        json.dump(res, f, indent=1)
    return "\n".join(lines)


def summarize_rocprof(outdir: str, steps: int, launches_per_step: dict,
                      early: bool = True) -> str:
    """Text summary of a `rocprofv3 --kernel-trace --stats` run of this
    script (rocpd database output): per-kernel totals over every dispatch of
    the command, and per-kernel averages over the last `steps` substeps
    (the timed window) from the dispatch trace."""
    import glob
    import sqlite3

    dbs = sorted(glob.glob(os.path.join(outdir, "**", "*_results.db"), recursive=True))
    if not dbs:
        return f"no rocpd database under {outdir}"
    cur = sqlite3.connect(dbs[0]).cursor()
    rows = cur.execute("select name, count(*), avg(duration), sum(duration) from kernels "
                       "group by name order by sum(duration) desc").fetchall()
    total = sum(r[3] for r in rows) or 1
    lines = ["all dispatches of the command (rocprofv3 --stats view):",
             f"{'kernel':28s} {'calls':>7s} {'avg_us':>12s} {'total_ms':>12s} {'pct':>6s}"]
    for name, n, avg, tot in rows:
        m = re.search(r"\b(k_[A-Za-z0-9_]+)(<[^>(]*>)?", name)
        short = (m.group(1) + (m.group(2) or "") if m else name.split("(")[0])[:28]
        lines.append(f"{short:28s} {n:7d} {avg / 1e3:12.1f} {tot / 1e6:12.1f} "
                     f"{100.0 * tot / total:6.2f}")
    lines += ["", f"timed window: last {steps} substeps of each kernel's dispatches",
              "(dispatches that returned at once, under 1% of the kernel's median, are",
              "left out: after a capacity overflow the substeps queued behind the failed",
              "one do no work, and the failed one runs again):"]
    # a tile kernel counts only inside its per-lane pass kernel's window (the
    # tile lists and classification inside predict_bbox's, one per substep);
    # per-lane keys come first
    bounds = {}
    for key, per_step in launches_per_step.items():
        rows = cur.execute("select start, duration from kernels where name like ? order by start",
                           (f"%k_{key}%",)).fetchall()
        take = steps * per_step
        if not rows or not take:
            continue
        med = sorted(r[1] for r in rows)[len(rows) // 2]
        live = [r for r in rows if r[1] >= 0.01 * med]
        if key.startswith("tile_"):
            lane = "predict_bbox" if key in ("tile_classify", "tile_neighbors") else key[5:]
            lo, hi = bounds.get(lane, (float("-inf"), float("inf")))
            w = [r[1] for r in live if lo < r[0] < hi]
        else:
            bounds[key] = window_bounds([r[0] for r in live], take)
            w = [r[1] for r in live[-take:]]
        if not w:
            continue
        lines.append(f"  {key:18s} launches {len(w):6d}  avg_us {sum(w) / len(w) / 1e3:12.1f}"
                     f"  (skipped {len(rows) - len(live)})")
    # the line's roofline_early: substeps 2-3 of the first context (the
    # dispatches between its second and fourth predict_bbox)
    pb = [r[0] for r in cur.execute("select start from kernels where name like "
                                    "'%k_predict_bbox%' order by start")]
    if early and len(pb) >= 4:
        lines += ["", "early-regime record (the line's roofline_early): substeps 2-3 of the",
                  "first context, from t = 0:"]
        for key in launches_per_step:
            w = [r[0] for r in cur.execute(
                "select duration from kernels where name like ? and start >= ? and start < ?",
                (f"%k_{key}%", pb[1], pb[3]))]
            if w:
                lines.append(f"  {key:18s} launches {len(w):6d}  avg_us "
                             f"{sum(w) / len(w) / 1e3:12.1f}")
    return "\n".join(lines)


def bench_ensemble(args) -> dict:
    """Segmented batch against a per-scene loop: `args.scenes` independent dam
    breaks of uneven size (lattice dims uniform in 8..32, |g| in [8,12],
    per-scene v0 in [-1,1]^3, all flags) stepped `args.steps` substeps
    (a) as one batched context replaying a hipGraph per substep,
    (b) as the same context without the graph,
    (c) one context per scene, each stepped through all substeps in turn, once
        eagerly and once replaying its own graph (the speed-up is quoted
        against the faster of the two).
    Final states of (a), (b) and (c) must be bitwise equal."""
    ens = scenes.ensemble(n_scenes=args.scenes)
    counts = [st.n for st in ens.states]
    n = int(sum(counts))
    pos = np.concatenate([st.pos for st in ens.states], axis=1)
    vel = np.concatenate([st.vel for st in ens.states], axis=1)

    # the ensemble's default is tiles off: its scenes turn to sparse spray and
    # wall piles within a few substeps, where the per-lane passes are faster
    # (240 substeps: 5.50 ms per substep batched, against 6.07 with tiles)
    tiles = "off" if args.tiles == "default" else args.tiles
    tg = {"auto": 0, "on": 1, "off": 2}[tiles]

    def run_batched(graph, prof=False):
        s = hip.HipSolver(ens.params, counts, block_size=args.block,
                          xcd_remap=bool(args.xcd_remap), tile_gather=tg)
        s.upload(pos, vel, 0.0)
        s.step(args.warmup, graph=graph)
        s.sync()
        if prof:
            s.profile(True)
        t0 = time.perf_counter()
        s.step(args.steps, graph=graph)
        s.sync()
        dt = time.perf_counter() - t0
        kern = read_kernels(s) if prof else None
        st = s.stats()
        out = s.download()
        s.close()
        return dt, out, kern, st

    if args.batched_only:
        # the batched context alone, eagerly, with per-kernel events: what a
        # kernel trace of the ensemble records (no graph replays, no
        # per-scene contexts)
        t_eager, _, kern, st = run_batched(False, prof=not args.no_profile)
        return {"metric": "ensemble substeps/s (all scenes per substep), batched eager only",
                "value": args.steps / t_eager, "unit": "steps/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": 1000.0 * t_eager / args.steps, "higher_is_better": True,
                "dtype": "f32", "data": "synthetic (seeded jittered lattices)",
                "config": {"workload": f"ensemble {args.scenes} scenes, dims U{{8..32}}^3, "
                                       "all flags, 4 iters, box 3^3",
                           "particles": n, "scenes": args.scenes},
                "tiles": tiles, "tiled_particles_last_step": st["tiled_particles"],
                "kernels_eager": kern}
    t_graph, out_graph, _, st_graph = run_batched(True)
    t_eager, out_eager, _, _ = run_batched(False)
    _, _, kern, _ = run_batched(False, prof=not args.no_profile)

    # one context per scene (created before the timed region, like the batch),
    # looped eagerly and then with a captured graph per scene
    solos = [hip.HipSolver(p, [st.n], cell_capacity=1 << 17,
                           neighbor_capacity=64 * st.n + 4096, block_size=args.block,
                           xcd_remap=bool(args.xcd_remap), tile_gather=tg)
             for p, st in zip(ens.params, ens.states)]
    off = np.concatenate([[0], np.cumsum(counts)])
    same = True
    t_loops = {}
    for graph in (False, True):
        for s, st in zip(solos, ens.states):
            s.upload(st.pos, st.vel, 0.0)
            s.step(args.warmup, graph=graph)
            s.sync()
        t0 = time.perf_counter()
        for s in solos:
            s.step(args.steps, graph=graph)
        t_loops[graph] = time.perf_counter() - t0
        for i, s in enumerate(solos):
            o = s.download()
            sl = slice(off[i], off[i + 1])
            same &= bool(np.array_equal(o.pos.view(np.uint32),
                                        out_graph.pos[:, sl].view(np.uint32)))
            same &= bool(np.array_equal(o.vel.view(np.uint32),
                                        out_graph.vel[:, sl].view(np.uint32)))
    for s in solos:
        s.close()
    t_loop = min(t_loops.values())
    same &= bool(np.array_equal(out_eager.pos.view(np.uint32), out_graph.pos.view(np.uint32)))
    same &= bool(np.array_equal(out_eager.vel.view(np.uint32), out_graph.vel.view(np.uint32)))

    sps = args.steps / t_graph
    return {
        "metric": "ensemble substeps/s (all scenes per substep)",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_graph / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattices)",
        "config": {"workload": f"ensemble {args.scenes} scenes, dims U{{8..32}}^3, all flags, "
                               "4 iters, box 3^3", "particles": n, "scenes": args.scenes},
        "particle_steps_per_s": sps * n,
        "batched_graph_ms_per_step": 1000.0 * t_graph / args.steps,
        "batched_eager_ms_per_step": 1000.0 * t_eager / args.steps,
        "per_scene_loop_ms_per_step": 1000.0 * t_loops[False] / args.steps,
        "per_scene_graph_loop_ms_per_step": 1000.0 * t_loops[True] / args.steps,
        "speedup_vs_per_scene_loop": t_loop / t_graph,
        "graph_vs_eager": t_eager / t_graph,
        "bitwise_equal_to_per_scene": same,
        "tiles": tiles,
        "tiled_particles_last_step": st_graph["tiled_particles"],
        "kernels_eager": kern,
    }


def adjoint_run(args, steps, coefficients, every) -> dict:
    """Differentiable rollout: the 65,536-particle block (64x32x32, d=0.05,
    +-0.1d jitter) in [0,4]x[0,3]x[0,2], s_corr + XSPH, no vorticity,
    per-particle v0 ~ N(0, 0.5^2), `steps` substeps.  Loss = mean squared
    distance of the final positions to a target cloud (a forward run from a
    second seeded velocity field).  Times a plain forward rollout, the
    recorded forward (checkpoint every `every` substeps) and the reverse pass
    (which recomputes all but the last segment), each the best of two runs."""
    sc = scenes.rollout_scene(steps=steps, coefficients=coefficients)
    tg = scenes.rollout_scene(seed=6, steps=steps, coefficients=coefficients)
    p, n = sc.params, sc.state.n
    s = hip.HipSolver(p, [n], block_size=args.block, xcd_remap=bool(args.xcd_remap),
                      tile_gather={"default": 0, "auto": 0, "on": 1, "off": 2}[args.tiles])
    s.upload(tg.state.pos, tg.state.vel)
    s.step(steps)
    target = s.download().pos.astype(np.float64)

    def timed(fn):
        s.sync()
        t0 = time.perf_counter()
        r = fn()
        s.sync()
        return time.perf_counter() - t0, r

    def forward():
        s.upload(sc.state.pos, sc.state.vel)
        return timed(lambda: s.step(steps))[0]

    forward()  # warm-up (capacity growth, code load)
    t_fwd = min(forward() for _ in range(2))
    s.upload(sc.state.pos, sc.state.vel)
    s.forward_record(steps, every)  # warm-up of the recording buffers
    s.backward(np.zeros((3, n), np.float32))
    # each phase is timed like the forward: the best of two runs
    t_rec = t_bwd = float("inf")
    for _ in range(2):
        s.upload(sc.state.pos, sc.state.vel)
        t, _ = timed(lambda: s.forward_record(steps, every))
        t_rec = min(t_rec, t)
        fin = s.download()
        diff = fin.pos.astype(np.float64) - target
        loss = float(np.mean(np.sum(diff * diff, axis=0)))
        seed = ((2.0 / n) * diff).astype(np.float32)
        t, (dx0, dv0, dg) = timed(lambda: s.backward(seed, None))
        t_bwd = min(t_bwd, t)
    st = s.stats()
    s.close()
    finite = float(np.mean(np.isfinite(dv0)))
    return {
        "coefficients": coefficients,
        "particles": n,
        "checkpoint_every": every,
        "ratio": (t_rec + t_bwd) / t_fwd,
        "forward_s": t_fwd,
        "recorded_forward_s": t_rec,
        "backward_s": t_bwd,
        "backward_vs_forward": t_bwd / t_fwd,
        "recomputed_substeps": st["recomputed"],
        "record_bytes": st["record_bytes"],
        "loss": loss,
        "grad_finite_fraction": finite,
        "grad_v0_max_abs": float(np.max(np.abs(dv0[np.isfinite(dv0)]))) if finite > 0 else None,
        "grad_gravity": dg[0].tolist(),
        "pairs_last_step": st["pairs"],
    }


def bench_adjoint(args) -> dict:
    """Adjoint against forward time on the differentiable-rollout workload
    with the reference's default relaxation and s_corr/XSPH coefficients
    (the gradient of 100 substeps is finite there), and the same timing with
    the dam-break coefficients, whose 100-substep gradient overflows float32
    (reported, not the headline)."""
    steps = args.steps
    # the whole rollout recorded as one segment (nothing recomputed; 2.4 GB at
    # 65,536 particles), and the same with a checkpoint every `args.every`
    main_run = adjoint_run(args, steps, "reference", steps)
    ckpt_run = adjoint_run(args, steps, "reference", args.every)
    stated = adjoint_run(args, steps, "stated", steps)
    return {
        "metric": "adjoint/forward time ratio (gradient of a 100-step rollout vs the rollout)",
        "value": main_run["ratio"],
        "unit": "x",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": 1000.0 * (main_run["recorded_forward_s"] + main_run["backward_s"]) / steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, seeded N(0, 0.5^2) velocities)",
        "config": {"workload": "rollout 64x32x32 s_corr+XSPH, 4 iters, box 4x3x2, "
                               "eps 600, k 5e-5, c 2e-4",
                   "particles": main_run["particles"], "checkpoint_every": steps},
        "run": main_run,
        "checkpointed_run": ckpt_run,
        "stated_coefficients_run": stated,
    }


TUNED_KEYS = ("block_size", "xcd_remap", "wave_mode", "brick_edge", "tile_gather")
# launch configuration fields a counter summary must share with a bench line
RUN_KEYS = ("block", "xcd_remap", "wave_mode", "brick_edge", "tile_gather", "graph")


def tuned_config(args):
    """Launch configuration of the timed region and where it came from: the
    file --tuned names (auto: profiles/tuned_million_block_warmup<W>.json,
    the tuner's choice for this warm-up, committed so that every run of the
    same command launches the same kernels), else the library defaults with
    --block / --xcd-remap / --brick-edge."""
    path = args.tuned
    if path == "auto":
        path = os.path.join(ROOT, "profiles", f"tuned_million_block_warmup{args.warmup}.json")
        if not os.path.exists(path):
            path = "none"
    if path != "none" and args.autotune == 0:
        with open(path) as f:
            rec = json.load(f)
        launch = {k: int(rec["launch"][k]) for k in TUNED_KEYS}
        return launch, os.path.relpath(path, ROOT)
    return ({"block_size": args.block, "xcd_remap": int(args.xcd_remap), "wave_mode": 0,
             "brick_edge": args.brick_edge, "tile_gather": 0}, "defaults")


def read_kernels(s) -> dict:
    """Per-kernel device time of the substeps queued since profiling was
    switched on (HIP events on the context's stream)."""
    out = {}
    for k in KERNELS:
        ms, cnt = s.profile_read(k)
        if cnt:
            out[k] = {"total_ms": ms, "launches": cnt, "avg_us": 1000.0 * ms / cnt}
    return out


def roofline_record(kernels: dict, n: int, mean_neighbors: float, n_tiled: int = 0) -> dict:
    """Roofline of the kernel with the most device time among the gather
    passes: algorithmic bytes per launch (KERNEL_BYTES x the particles it
    serves) over its average launch time, against the HBM peak; the FP32
    fraction from the algorithmic flops per neighbor pair (FLOPS_PER_PAIR x
    mean neighbors x particles).  A tile kernel ("tile_<pass>") serves the
    n_tiled particles of the LDS tiles; with n_tiled > 0 the dominant kernel
    is taken among the tile kernels.  Adds both figures to every kernel's
    entry."""
    def base(k):
        return k[5:] if k.startswith("tile_") else k

    def served(k):
        return n_tiled if k.startswith("tile_") else n

    for k, v in kernels.items():
        b = base(k)
        if b in KERNEL_BYTES and served(k) > 0:
            t = v["avg_us"] * 1e-6
            v["algorithmic_GBs"] = KERNEL_BYTES[b] * served(k) / t / 1e9
            v["hbm_frac"] = v["algorithmic_GBs"] / HBM_PEAK_GBS
            if b in FLOPS_PER_PAIR:
                v["algorithmic_TFLOPs"] = FLOPS_PER_PAIR[b] * mean_neighbors * served(k) / t / 1e12
                v["fp32_frac"] = v["algorithmic_TFLOPs"] / FP32_PEAK_TFLOPS
    pool = [k for k in kernels if "algorithmic_GBs" in kernels[k]
            and k.startswith("tile_") == (n_tiled > 0)]
    dom = max(pool, key=lambda k: kernels[k]["total_ms"])
    v = kernels[dom]
    return {"bound": "hbm", "kernel": dom, "achieved": v["algorithmic_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": v["hbm_frac"], "traffic": None,
            "bytes_per_launch": KERNEL_BYTES[base(dom)] * served(dom),
            "avg_us": v["avg_us"],
            "fp32_TFLOPs": v.get("algorithmic_TFLOPs"), "fp32_peak_TFLOPs": FP32_PEAK_TFLOPS,
            "fp32_frac": v.get("fp32_frac")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed substeps (default: 120 million_block, 240 ensemble, 100 adjoint)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed substeps first (default: 30 million_block, 0 ensemble)")
    ap.add_argument("--graph", action="store_true", help="replay a captured hipGraph per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-cpu-1t", action="store_true",
                    help="skip the one-thread step of the CPU baseline")
    ap.add_argument("--no-profile", action="store_true", help="no per-kernel events")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--xcd-remap", type=int, default=0)
    ap.add_argument("--brick-edge", type=int, default=0, help="cell table brick edge (0: 4)")
    ap.add_argument("--dims", type=str, default="128,128,64")
    ap.add_argument("--workload", choices=["million_block", "ensemble", "adjoint"],
                    default="million_block",
                    help="ensemble: batched + hipGraph against a per-scene loop; adjoint: "
                         "gradient of a rollout against the rollout (each its own JSON line)")
    ap.add_argument("--every", type=int, default=10,
                    help="checkpoint interval of the adjoint's checkpointed run")
    ap.add_argument("--autotune", type=int, default=0, metavar="TRIAL_STEPS",
                    help="autotune the launch configuration on the warmed-up state before "
                         "the timed region (0: off; the configuration comes from --tuned)")
    ap.add_argument("--tuned", type=str, default="auto", metavar="JSON",
                    help="launch configuration of the timed region, as written by "
                         "--write-tuned (auto: profiles/tuned_million_block_warmup<W>.json "
                         "when present, else the library defaults; none: the defaults)")
    ap.add_argument("--write-tuned", type=str, default="", metavar="JSON",
                    help="with --autotune: write the tuner's choice for this warm-up here")
    ap.add_argument("--scenes", type=int, default=256, help="ensemble size")
    ap.add_argument("--batched-only", action="store_true",
                    help="ensemble: time the batched context eagerly and nothing else")
    ap.add_argument("--tiles", choices=["default", "auto", "on", "off"], default="default",
                    help="ensemble and adjoint workloads: the LDS-tiled passes (auto: the "
                         "library default, on from 32,768 particles with brick edge 4; "
                         "default: off for the ensemble, auto for the adjoint)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a device copy of this size (counter calibration)")
    ap.add_argument("--summarize-rocprof", type=str, default="", metavar="DIR",
                    help="print a summary of a rocprofv3 output directory of this script")
    ap.add_argument("--summarize-traffic", type=str, default="", metavar="DIR",
                    help="HBM traffic and L2 hit rate per launch from the pmc_fetch / pmc_write "
                         "/ pmc_l2 / pmc_cfetch / pmc_cwrite rocprofv3 runs under DIR (writes "
                         "--traffic-json)")
    ap.add_argument("--traffic-json", type=str, default="",
                    help="output of --summarize-traffic (default: "
                         "profiles/traffic_million_block_warmup<W>_steps<K>.json)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"million_block": 120, "ensemble": 240, "adjoint": 100}[args.workload]
    if args.warmup is None:
        args.warmup = 30 if args.workload == "million_block" else 0
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.summarize_traffic:
        print(summarize_traffic(args.summarize_traffic, args.traffic_json or None))
        return
    if args.summarize_rocprof:
        print(summarize_rocprof(args.summarize_rocprof, args.steps,
                                {"predict_bbox": 1, "lambda": 4, "delta": 4, "xsph": 1,
                                 "vort_omega": 1,
                                 "vort_apply": 1, "neighbors": 1, "reorder": 1,
                                 "tile_lambda": 4, "tile_delta": 4, "tile_xsph": 1,
                                 "tile_vort_omega": 1, "tile_vort_apply": 1,
                                 "tile_neighbors": 1, "tile_classify": 1},
                                early=args.workload == "million_block"))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "layout is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args)))
        return
    if args.workload == "adjoint":
        print(json.dumps(bench_adjoint(args)))
        return

    dims = tuple(int(x) for x in args.dims.split(","))
    sc = scenes.million_block(dims=dims)
    n = sc.state.n
    prof = not args.no_profile and not args.graph
    early = None
    if prof and args.warmup >= 3:
        # substeps 2-3 from the start, before particles pile up (about 28
        # neighbors each): the regime the algorithmic byte model describes,
        # in a context of its own with the default launch configuration
        se = hip.HipSolver(sc.params, [n])
        se.upload(sc.state.pos, sc.state.vel, 0.0)
        se.step(1)
        se.profile(True)
        se.step(2)
        early_k = read_kernels(se)
        se.profile(False)
        st_e = se.stats()
        se.close()
        tiled = st_e["tiled_particles"] if "tile_lambda" in early_k else 0
        early = roofline_record(early_k, n, mean_neighbors=st_e["pairs"] / n, n_tiled=tiled)
        early["substeps"] = "2-3"
        early["mean_neighbors"] = st_e["pairs"] / n
        early["tiled_particles"] = st_e["tiled_particles"]
        early["kernels"] = early_k
    launch, launch_src = tuned_config(args)
    s = hip.HipSolver(sc.params, [n], block_size=launch["block_size"],
                      xcd_remap=bool(launch["xcd_remap"]), wave_mode=launch["wave_mode"],
                      brick_edge=launch["brick_edge"], tile_gather=launch["tile_gather"])
    s.upload(sc.state.pos, sc.state.vel, 0.0)
    s.step(args.warmup, graph=args.graph)
    start_state = s.download() if not args.no_cpu else None
    # the tuner's trials start from this state and leave it in place (they
    # also size the tables for it); the timed region starts here too
    tuned = s.autotune(args.autotune) if args.autotune > 0 else None
    if tuned:
        launch = {k: tuned[k] for k in TUNED_KEYS}
        launch_src = f"autotune({args.autotune}) on the state after {args.warmup} substeps"
        if args.write_tuned:
            with open(args.write_tuned, "w") as f:
                json.dump({"workload": f"million_block {dims[0]}x{dims[1]}x{dims[2]}",
                           "warmup": args.warmup, "launch": launch, "autotune": tuned,
                           "command": " ".join(["python", "bench.py", *sys.argv[1:]])},
                          f, indent=1)
    s.sync()
    if prof:
        s.profile(True)
    grows_before = s.stats()["grow_events"]
    sync_all()
    t0 = time.perf_counter()
    s.step(args.steps, graph=args.graph)  # one sync at the end; grows tables if needed
    sync_all()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    kernels = read_kernels(s) if prof else {}
    if prof:
        s.profile(False)
    st = s.stats()
    sizes = {"neighbor_pairs_last_step": st["pairs"], "list_words_last_step": st["words"],
             "grid_cells_last_step": st["cells"], "tiled_particles_last_step": st["tiled_particles"],
             "capacity_growths_in_window": st["grow_events"] - grows_before}
    s.close()

    config = {"workload": f"million_block {dims[0]}x{dims[1]}x{dims[2]} all flags, 4 iters",
              "dims": list(dims), "particles": n, "graph": bool(args.graph),
              "block": launch["block_size"], "xcd_remap": bool(launch["xcd_remap"]),
              "wave_mode": launch["wave_mode"], "brick_edge": launch["brick_edge"],
              "tile_gather": launch["tile_gather"], "launch_from": launch_src}
    roofline = None
    if kernels:
        roofline = roofline_record(kernels, n, mean_neighbors=st["pairs"] / max(n, 1))
        run = {"dims": list(dims), "warmup": args.warmup, "steps": args.steps,
               "config": {k: config[k] for k in RUN_KEYS}}
        traffic, l2, note = measured_traffic(roofline["kernel"], run)
        roofline["traffic"] = traffic
        roofline["l2_hit_rate"] = l2
        roofline["traffic_source"] = note

    cpu = None
    if start_state is not None:
        cpu = cpu_baseline(start_state, sc.params, cpu_threads(), one_thread=not args.no_cpu_1t)

    step_gbs = STEP_BYTES_ALL_FLAGS * n * steps_per_s / 1e9
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 / steps_per_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": config,
        "autotune": tuned,
        "particle_steps_per_s": steps_per_s * n,
        "step_hbm_fraction": step_gbs / HBM_PEAK_GBS,
        "roofline": roofline,
        "roofline_early": early,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "sizes": sizes,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
