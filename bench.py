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
KERNELS = ["predict_bbox", "grid_setup", "cell_count", "scan", "scatter", "reorder",
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
        short = name.split("::")[-1].split("(")[0].split("<")[0][:28]
        lines.append(f"{short:28s} {n:7d} {avg / 1e3:12.1f} {tot / 1e6:12.1f} "
                     f"{100.0 * tot / total:6.2f}")
    lines += ["", f"timed window: last {steps} substeps of each kernel's dispatches:"]
    for key, per_step in launches_per_step.items():
        d = [r[0] for r in cur.execute(
            "select duration from kernels where name like ? order by start", (f"%k_{key}%",))]
        take = steps * per_step
        if d and take:
            w = d[-take:]
            lines.append(f"  {key:18s} launches {len(w):6d}  avg_us {sum(w) / len(w) / 1e3:12.1f}")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--graph", action="store_true", help="replay a captured hipGraph per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="no per-kernel events")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--xcd-remap", type=int, default=1)
    ap.add_argument("--dims", type=str, default="128,128,64")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a device copy of this size (counter calibration)")
    ap.add_argument("--summarize-rocprof", type=str, default="", metavar="DIR",
                    help="print a summary of a rocprofv3 output directory of this script")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
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

    dims = tuple(int(x) for x in args.dims.split(","))
    sc = scenes.million_block(dims=dims)
    n = sc.state.n
    s = hip.HipSolver(sc.params, [n], block_size=args.block, xcd_remap=bool(args.xcd_remap))
    s.upload(sc.state.pos, sc.state.vel, 0.0)
    s.step(args.warmup, graph=args.graph)
    start_state = s.download() if not args.no_cpu else None
    s.step(1, graph=args.graph)  # settles capacity growth after the download
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
                   "particles": n, "graph": bool(args.graph), "block": args.block,
                   "xcd_remap": bool(args.xcd_remap)},
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
