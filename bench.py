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
KERNELS = ["predict_bbox", "grid_setup", "cell_count", "scan", "sort", "reorder", "runs",
           "neighbors", "lambda", "delta", "xsph", "vorticity_omega", "vorticity_apply"]


def sync_all():
    """Device-wide synchronize (the library stream is synchronized by
    HipSolver.sync(); torch is only touched if the caller imported it)."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()


def measured_traffic(kernel: str, dims) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter's reading for a device copy of known size), if present."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_million_block.json")
    if tuple(dims) != (128, 128, 64) or not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    k = data.get("kernels", {}).get(kernel)
    return None if k is None else float(k["hbm_bytes_per_launch"])


def cpu_baseline(state, params, threads, steps=1, warm=0):
    """The reference's own CPU step (oracle/_ref, built from core.cpp) on the
    state the GPU timed region starts from."""
    import oracle

    which = "ref" if oracle.have_ref() else "oracle"
    oracle.set_threads(threads, which)
    r = oracle.step(params, state.pos, state.vel, nsteps=warm, which=which)
    t0 = time.perf_counter()
    oracle.step(params, r["pos"], r["vel"], nsteps=steps, which=which)
    dt = time.perf_counter() - t0
    return {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": threads,
        "kind": "reference" if which == "ref" else "port",
        "sample": (f"{'reference core.cpp' if which == 'ref' else 'oracle'} step() on the "
                   f"{state.n}-particle state the GPU timed region starts from: {warm} untimed "
                   f"+ {steps} timed steps, {threads} OpenMP threads "
                   f"({steps * state.n / dt:.3g} particle-steps/s)"),
    }


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


def summarize_traffic(outdir: str, steps: int, json_path: str) -> str:
    """HBM bytes per launch of lambda and delta from four `rocprofv3 --pmc`
    runs under `outdir` (pmc_fetch: FETCH_SIZE and pmc_write: WRITE_SIZE of
    this script's headline workload; pmc_cfetch / pmc_cwrite: the same
    counters for `--copy-calibration 256`).  Each kernel counter (KiB per
    dispatch) is scaled by what it reads for the copy of known size and
    averaged over the last 4 * `steps` dispatches in dispatch order.  Writes
    the JSON that `measured_traffic` reads and returns a text table."""
    import collections
    import glob
    import sqlite3

    def per_dispatch(sub, counter):
        dbs = sorted(glob.glob(os.path.join(outdir, sub, "**", "*_results.db"), recursive=True))
        if not dbs:
            raise SystemExit(f"no rocpd database under {os.path.join(outdir, sub)}")
        rows = sqlite3.connect(dbs[0]).cursor().execute(
            "select dispatch_id, name, sum(counter_value) from pmc_events where counter_name=? "
            "group by dispatch_id order by dispatch_id", (counter,)).fetchall()
        out = collections.defaultdict(list)
        for did, name, v in rows:
            out[name].append((did, v))
        return out

    copy_kib = 256 * 1024

    def copy_reading(d):
        vals = [v for k in d if "copyBuffer" in k for _, v in d[k]]
        return sum(vals) / len(vals)

    f_read = copy_reading(per_dispatch("pmc_cfetch", "FETCH_SIZE"))
    w_read = copy_reading(per_dispatch("pmc_cwrite", "WRITE_SIZE"))
    f_scale, w_scale = copy_kib / f_read, copy_kib / w_read
    kf = per_dispatch("pmc_fetch", "FETCH_SIZE")
    kw = per_dispatch("pmc_write", "WRITE_SIZE")
    n = 1 << 20
    alg = {"lambda": 16.0 * n, "delta": 28.0 * n}
    res = {"workload": "million_block 128x128x64 all flags, 4 iters; substeps 31-33 "
                       "(the first substeps of the timed window)",
           "copy_calibration": {"bytes": copy_kib * 1024, "fetch_kib_read": f_read,
                                "write_kib_read": w_read, "fetch_scale": f_scale,
                                "write_scale": w_scale},
           "kernels": {}}
    lines = [f"Calibration: FETCH_SIZE {f_read:.0f} KiB -> scale {f_scale:.4f}, "
             f"WRITE_SIZE {w_read:.0f} KiB -> scale {w_scale:.4f}", "",
             "kernel       fetch MB   write MB     HBM MB  algorithmic MB   ratio"]
    for key in ("lambda", "delta"):
        fv = [v for _, v in sorted(x for k in kf if f"k_{key}" in k for x in kf[k])][-4 * steps:]
        wv = [v for _, v in sorted(x for k in kw if f"k_{key}" in k for x in kw[k])][-4 * steps:]
        fetch = sum(fv) / len(fv) * 1024 * f_scale
        write = sum(wv) / len(wv) * 1024 * w_scale
        res["kernels"][key] = {"fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
                               "hbm_bytes_per_launch": fetch + write,
                               "algorithmic_bytes_per_launch": alg[key],
                               "traffic_over_algorithmic": (fetch + write) / alg[key],
                               "launches": len(fv)}
        lines.append(f"{key:<10}{fetch / 1e6:11.1f}{write / 1e6:11.1f}{(fetch + write) / 1e6:11.1f}"
                     f"{alg[key] / 1e6:16.1f}{(fetch + write) / alg[key]:8.1f}")
    with open(json_path, "w") as f:
        json.dump(res, f, indent=1)
    return "\n".join(lines)


def summarize_rocprof(outdir: str, steps: int, launches_per_step: dict) -> str:
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
    for key, per_step in launches_per_step.items():
        d = [r[0] for r in cur.execute(
            "select duration from kernels where name like ? order by start", (f"%k_{key}%",))]
        take = steps * per_step
        if d and take:
            med = sorted(d)[len(d) // 2]
            live = [x for x in d if x >= 0.01 * med]
            w = live[-take:]
            lines.append(f"  {key:18s} launches {len(w):6d}  avg_us {sum(w) / len(w) / 1e3:12.1f}"
                         f"  (skipped {len(d) - len(live)})")
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

    def run_batched(graph):
        s = hip.HipSolver(ens.params, counts, block_size=args.block,
                          xcd_remap=bool(args.xcd_remap))
        s.upload(pos, vel, 0.0)
        s.step(args.warmup, graph=graph)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.steps, graph=graph)
        s.sync()
        dt = time.perf_counter() - t0
        out = s.download()
        s.close()
        return dt, out

    t_graph, out_graph = run_batched(True)
    t_eager, out_eager = run_batched(False)

    # one context per scene (created before the timed region, like the batch),
    # looped eagerly and then with a captured graph per scene
    solos = [hip.HipSolver(p, [st.n], cell_capacity=1 << 17,
                           neighbor_capacity=64 * st.n + 4096, block_size=args.block,
                           xcd_remap=bool(args.xcd_remap))
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
    }


def adjoint_run(args, steps, coefficients, every) -> dict:
    """Differentiable rollout: the 65,536-particle block (64x32x32, d=0.05,
    +-0.1d jitter) in [0,4]x[0,3]x[0,2], s_corr + XSPH, no vorticity,
    per-particle v0 ~ N(0, 0.5^2), `steps` substeps.  Loss = mean squared
    distance of the final positions to a target cloud (a forward run from a
    second seeded velocity field).  Times a plain forward rollout, the
    recorded forward (checkpoint every `every` substeps) and the reverse pass
    (which recomputes all but the last segment)."""
    sc = scenes.rollout_scene(steps=steps, coefficients=coefficients)
    tg = scenes.rollout_scene(seed=6, steps=steps, coefficients=coefficients)
    p, n = sc.params, sc.state.n
    s = hip.HipSolver(p, [n], block_size=args.block, xcd_remap=bool(args.xcd_remap))
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
    s.upload(sc.state.pos, sc.state.vel)
    t_rec, _ = timed(lambda: s.forward_record(steps, every))
    fin = s.download()
    diff = fin.pos.astype(np.float64) - target
    loss = float(np.mean(np.sum(diff * diff, axis=0)))
    seed = ((2.0 / n) * diff).astype(np.float32)
    t_bwd, (dx0, dv0, dg) = timed(lambda: s.backward(seed, None))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed substeps (default: 120 million_block, 240 ensemble, 100 adjoint)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed substeps first (default: 30 million_block, 0 ensemble)")
    ap.add_argument("--graph", action="store_true", help="replay a captured hipGraph per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="no per-kernel events")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--xcd-remap", type=int, default=1)
    ap.add_argument("--brick-edge", type=int, default=0, help="cell table brick edge (0: 4)")
    ap.add_argument("--dims", type=str, default="128,128,64")
    ap.add_argument("--workload", choices=["million_block", "ensemble", "adjoint"],
                    default="million_block",
                    help="ensemble: batched + hipGraph against a per-scene loop; adjoint: "
                         "gradient of a rollout against the rollout (each its own JSON line)")
    ap.add_argument("--every", type=int, default=10,
                    help="checkpoint interval of the adjoint's checkpointed run")
    ap.add_argument("--autotune", type=int, default=2, metavar="TRIAL_STEPS",
                    help="autotune the launch configuration on the warmed-up state before "
                         "the timed region (0: off)")
    ap.add_argument("--scenes", type=int, default=256, help="ensemble size")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a device copy of this size (counter calibration)")
    ap.add_argument("--summarize-rocprof", type=str, default="", metavar="DIR",
                    help="print a summary of a rocprofv3 output directory of this script")
    ap.add_argument("--summarize-traffic", type=str, default="", metavar="DIR",
                    help="HBM traffic per launch from the pmc_fetch / pmc_write / pmc_cfetch / "
                         "pmc_cwrite rocprofv3 runs under DIR (writes --traffic-json)")
    ap.add_argument("--traffic-json", type=str,
                    default=os.path.join(ROOT, "profiles", "hbm_traffic_million_block.json"))
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"million_block": 120, "ensemble": 240, "adjoint": 100}[args.workload]
    if args.warmup is None:
        args.warmup = 30 if args.workload == "million_block" else 0
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.summarize_traffic:
        print(summarize_traffic(args.summarize_traffic, args.steps, args.traffic_json))
        return
    if args.summarize_rocprof:
        print(summarize_rocprof(args.summarize_rocprof, args.steps,
                                {"lambda": 4, "delta": 4, "xsph": 1, "vort_omega": 1,
                                 "vort_apply": 1, "neighbors": 1, "reorder": 1}))
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

    dims =tuple(int(x) for x in args.dims.split(","))
    sc = scenes.million_block(dims=dims)
    n = sc.state.n
    s = hip.HipSolver(sc.params, [n], block_size=args.block, xcd_remap=bool(args.xcd_remap),
                      brick_edge=args.brick_edge)
    s.upload(sc.state.pos, sc.state.vel, 0.0)
    s.step(args.warmup, graph=args.graph)
    start_state = s.download() if not args.no_cpu else None
    # the tuner's trials start from this state and leave it in place (they
    # also size the tables for it); the timed region starts here too
    tuned = s.autotune(args.autotune) if args.autotune > 0 else None
    s.sync()
    prof = not args.no_profile and not args.graph
    if prof:
        s.profile(True)
    sync_all()
    t0 = time.perf_counter()
    grows_before = s.stats()["grow_events"]
    s.step(args.steps, graph=args.graph)  # one sync at the end; grows tables if needed
    sync_all()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    st = s.stats()
    sizes = {"neighbor_pairs_last_step": st["pairs"], "list_words_last_step": st["words"],
             "grid_cells_last_step": st["cells"],
             "capacity_growths_in_window": st["grow_events"] - grows_before}
    kernels = {}
    if prof:
        for k in KERNELS:
            ms, cnt = s.profile_read(k)
            if cnt:
                kernels[k] = {"total_ms": ms, "launches": cnt, "avg_us": 1000.0 * ms / cnt}
        s.profile(False)
    s.close()

    roofline = None
    if kernels:
        ranked = sorted((k for k in kernels if k in KERNEL_BYTES),
                        key=lambda k: -kernels[k]["total_ms"])
        dom = ranked[0]
        avg_s = kernels[dom]["avg_us"] * 1e-6
        achieved = KERNEL_BYTES[dom] * n / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic(dom, dims),
                    "bytes_per_launch": KERNEL_BYTES[dom] * n}
        for k, v in kernels.items():
            if k in KERNEL_BYTES:
                v["algorithmic_GBs"] = KERNEL_BYTES[k] * n / (v["avg_us"] * 1e-6) / 1e9

    cpu = None
    if start_state is not None:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(start_state, sc.params, threads)

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
        "config": {"workload": f"million_block {dims[0]}x{dims[1]}x{dims[2]} all flags, 4 iters",
                   "particles": n, "graph": bool(args.graph),
                   "block": tuned["block_size"] if tuned else args.block,
                   "xcd_remap": bool(tuned["xcd_remap"] if tuned else args.xcd_remap),
                   "wave_mode": tuned["wave_mode"] if tuned else 2,
                   "brick_edge": tuned["brick_edge"] if tuned else (args.brick_edge or 4)},
        "autotune": tuned,
        "particle_steps_per_s": steps_per_s * n,
        "step_hbm_fraction": step_gbs / HBM_PEAK_GBS,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "sizes": sizes,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
