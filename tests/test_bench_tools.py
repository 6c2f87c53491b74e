"""The measurement tools of bench.py on synthetic rocprofv3 (rocpd) databases:
the HBM-traffic summary scales each counter by its reading for a copy of known
size and averages the last dispatches in dispatch order, across template
variants of one kernel; the kernel-trace window leaves out dispatches that
returned at once after a capacity overflow."""
import json
import os
import sqlite3

import bench


def _pmc_db(path, rows):
    """rows: (dispatch_id, kernel name, counter, value)."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    con = sqlite3.connect(path)
    con.execute("create table pmc_events (dispatch_id int, name text, counter_name text, "
                "counter_value real)")
    con.executemany("insert into pmc_events values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def _kernel_db(path, rows):
    """rows: (kernel name, start, duration)."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, start int, duration int)")
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.commit()
    con.close()


def test_traffic_summary_scales_and_orders_dispatches(tmp_path):
    copy_kib = 256 * 1024
    # the copy reads half its bytes through FETCH_SIZE: scale 2
    _pmc_db(str(tmp_path / "pmc_cfetch" / "c_results.db"),
            [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", copy_kib / 2)] * 1)
    _pmc_db(str(tmp_path / "pmc_cwrite" / "c_results.db"),
            [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", copy_kib)])
    # delta: iterations 0-2 under one template name, the last under another;
    # the early dispatches (warm-up) read 1000 KiB, the last three substeps 10
    fetch, write = [], []
    did = 0
    for step in range(6):
        for it in range(4):
            did += 1
            name = "k_delta<true, true, false>" if it == 3 else "k_delta<true, false, false>"
            v = 1000.0 if step < 3 else 10.0
            fetch += [(did, name, "FETCH_SIZE", v)]
            write += [(did, name, "WRITE_SIZE", v / 2)]
            did += 1
            fetch += [(did, "k_lambda", "FETCH_SIZE", v)]
            write += [(did, "k_lambda", "WRITE_SIZE", v / 2)]
    _pmc_db(str(tmp_path / "pmc_fetch" / "f_results.db"), fetch)
    _pmc_db(str(tmp_path / "pmc_write" / "w_results.db"), write)
    # L2 hits and misses of the same dispatches: 3 hits per miss in the window
    l2 = []
    for did, name, _, v in fetch:
        l2 += [(did, name, "TCC_HIT_sum", 3.0 * v), (did, name, "TCC_MISS_sum", v)]
    _pmc_db(str(tmp_path / "pmc_l2" / "l_results.db"), l2)
    out = str(tmp_path / "traffic.json")
    text = bench.summarize_traffic(str(tmp_path), out, steps=3)
    d = json.load(open(out))
    assert abs(d["copy_calibration"]["fetch_scale"] - 2.0) < 1e-12
    for k in ("lambda", "delta"):
        r = d["kernels"][k]
        assert r["launches"] == 12
        # only the last three substeps (10 KiB each) count, both template variants
        assert abs(r["fetch_bytes_per_launch"] - 10.0 * 1024 * 2.0) < 1e-6
        assert abs(r["write_bytes_per_launch"] - 5.0 * 1024) < 1e-6
        assert abs(r["l2_hit_rate"] - 0.75) < 1e-12
    assert "lambda" in text and "delta" in text


def test_traffic_is_reported_only_for_the_recorded_window_and_config(tmp_path, monkeypatch):
    prof = tmp_path / "profiles"
    prof.mkdir()
    cfg = {"block": 256, "xcd_remap": False, "wave_mode": 2, "brick_edge": 1, "graph": False}
    rec = {"run": {"dims": [128, 128, 64], "warmup": 5, "steps": 20, "config": cfg},
           "kernels": {"lambda": {"hbm_bytes_per_launch": 123.0, "l2_hit_rate": 0.9}}}
    (prof / "traffic_million_block_warmup5_steps20.json").write_text(json.dumps(rec))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    run = {"dims": (128, 128, 64), "warmup": 5, "steps": 20, "config": dict(cfg)}
    t, hit, src = bench.measured_traffic("lambda", run)
    assert t == 123.0 and hit == 0.9 and src.endswith("steps20.json")
    # another brick edge, another window: no traffic, and the reason says why
    for other in (dict(run, config=dict(cfg, brick_edge=4)), dict(run, warmup=30, steps=120)):
        t, hit, why = bench.measured_traffic("lambda", other)
        assert t is None and hit is None and "no committed counter summary" in why


def test_rocprof_window_skips_instant_dispatches(tmp_path):
    rows, t = [], 0
    # 5 real substeps of lambda at 2000 us, then a capacity overflow: 2
    # substeps queued behind it return at once (2 us), then 3 real again
    for us in [2000.0] * 5 + [2.0] * 2 + [2000.0] * 3:
        for _ in range(4):
            t += 1
            rows.append(("k_lambda", t, int(us * 1000)))
    _kernel_db(str(tmp_path / "kt" / "h_results.db"), rows)
    text = bench.summarize_rocprof(str(tmp_path / "kt"), 6, {"lambda": 4})
    line = [l for l in text.splitlines() if l.strip().startswith("lambda")][0]
    assert "launches     24" in line
    assert "avg_us       2000.0" in line
    assert "(skipped 8)" in line


def test_more_than_one_gpu_is_refused_with_a_message():
    """bench.py measures one GPU; asking for more exits with status 2 and
    says why, before touching any device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(os.path.dirname(bench.__file__), "bench.py"),
                        "--gpus", "2", "--steps", "1", "--warmup", "0", "--no-cpu"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_tile_launches_count_only_inside_the_window(tmp_path):
    """An early-regime context runs 3 substeps with tile kernels before the
    timed context; the timed window (per-lane kernels only, or per-lane and
    tile kernels launched side by side) must not take the early tile launches
    in, in the traffic summary or the kernel-trace window."""
    copy_kib = 256 * 1024
    _pmc_db(str(tmp_path / "pmc_cfetch" / "c_results.db"),
            [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", copy_kib)])
    _pmc_db(str(tmp_path / "pmc_cwrite" / "c_results.db"),
            [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", copy_kib)])

    def record(tiles_in_window):
        fetch, kt, did = [], [], 0
        # early context: 3 substeps, tile kernel (500 KiB) launched before
        # the per-lane one (1 KiB)
        for _ in range(3 * 4):
            did += 1
            fetch.append((did, "k_tile_lambda", "FETCH_SIZE", 500.0))
            kt.append(("k_tile_lambda", did * 10, 50_000))
            did += 1
            fetch.append((did, "k_lambda", "FETCH_SIZE", 1.0))
            kt.append(("k_lambda", did * 10, 10_000))
            did += 1
            fetch.append((did, "k_delta", "FETCH_SIZE", 1.0))
            kt.append(("k_delta", did * 10, 10_000))
        # timed context: 2 warm-up and 4 timed substeps of 100 KiB per-lane
        # launches, with 7 KiB tile launches beside them when tiled (each
        # lambda launch followed by a delta launch, as in a substep)
        for step in range(6):
            for _ in range(4):
                if tiles_in_window:
                    did += 1
                    fetch.append((did, "k_tile_lambda", "FETCH_SIZE", 7.0))
                    kt.append(("k_tile_lambda", did * 10, 70_000))
                did += 1
                fetch.append((did, "k_lambda", "FETCH_SIZE", 100.0))
                kt.append(("k_lambda", did * 10, 1_000_000))
                did += 1
                fetch.append((did, "k_delta", "FETCH_SIZE", 100.0))
                kt.append(("k_delta", did * 10, 1_000_000))
        return fetch, kt

    for tiles_in_window, expect_kib, tile_line in ((False, 100.0, None), (True, 107.0, 16)):
        sub = tmp_path / f"t{int(tiles_in_window)}"
        fetch, kt = record(tiles_in_window)
        for d in ("pmc_cfetch", "pmc_cwrite"):
            os.makedirs(sub / d, exist_ok=True)
            os.link(tmp_path / d / "c_results.db", sub / d / "c_results.db")
        _pmc_db(str(sub / "pmc_fetch" / "f_results.db"), fetch)
        _pmc_db(str(sub / "pmc_write" / "w_results.db"),
                [(i, n, "WRITE_SIZE", 0.0) for i, n, _, _ in fetch])
        out = str(sub / "traffic.json")
        bench.summarize_traffic(str(sub), out, steps=4)
        r = json.load(open(out))["kernels"]["lambda"]
        assert r["launches"] == 16
        assert abs(r["fetch_bytes_per_launch"] - expect_kib * 1024) < 1e-6
        _kernel_db(str(sub / "kt" / "k_results.db"), kt)
        text = bench.summarize_rocprof(str(sub / "kt"), 4, {"lambda": 4, "tile_lambda": 4})
        lines = [l for l in text.splitlines() if l.strip().startswith("tile_lambda")]
        if tile_line is None:
            assert not lines
        else:
            assert f"launches     {tile_line}" in lines[0] and "avg_us         70.0" in lines[0]
