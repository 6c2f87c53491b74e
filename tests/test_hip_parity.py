"""HIP backend vs the CPU oracle and the reference's golden vectors, through
the C ABI.  Bar: cell keys, sorted order, neighbor counts and lists
bit-exact; positions and velocities within 1e-5*h after one step (the
kernels follow the reference's operation order, so they are compared
bitwise as well)."""
import numpy as np
import pytest

import oracle
from fluidsimulator_amd import hip, scenes
from tests.helpers import golden_names, load_golden, neighbor_lists, same_bits

pytestmark = pytest.mark.gpu

H = 0.1
POS_TOL = 1e-5 * H


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()  # raises when the library is missing: no fallback
    ok, count, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)


def run_hip(p, pos, vel, steps, record_diag=True, graph=False, **kw):
    with hip.HipSolver(p, [pos.shape[1]], record_diag=record_diag, **kw) as s:
        s.upload(pos, vel, 0.0)
        s.step(steps, graph=graph)
        st = s.download()
        out = {"pos": st.pos, "vel": st.vel, "time": np.float32(st.time)}
        if steps > 0:
            out.update(s.diag())
            out["nbr_ids"] = s.neighbors()
            out["keys"], out["desc"] = s.grid()
    return out


def brick_keys(cells, desc, edge=4):
    """Sort key of each cell: scene offset + brick index * edge^3 + Z-order
    of the cell inside its brick (bits of x, y, z interleaved; the same
    definition as the kernels)."""
    ox, oy, oz, nbx, nby, nbz, off, _ = [int(v) for v in desc]
    rx, ry, rz = cells[:, 0] - ox, cells[:, 1] - oy, cells[:, 2] - oz
    bits = edge.bit_length() - 1
    b = ((rz >> bits) * nby + (ry >> bits)) * nbx + (rx >> bits)
    m = np.zeros_like(rx)
    for k in range(bits):
        m |= (((rx >> k) & 1) << (3 * k)) | (((ry >> k) & 1) << (3 * k + 1)) | \
             (((rz >> k) & 1) << (3 * k + 2))
    return off + b * edge ** 3 + m


def expected_desc(cells, edge=4):
    lo = cells.min(axis=0)
    hi = cells.max(axis=0)
    size = hi - lo + 3
    nb = (size + edge - 1) // edge
    return np.array([lo[0] - 1, lo[1] - 1, lo[2] - 1, nb[0], nb[1], nb[2], 0,
                     int(nb.prod()) * edge ** 3], np.int64)


@pytest.mark.parametrize("name", [n for n in golden_names() if not n.startswith(("bruteforce", "powf_"))])
def test_golden_bitwise(name):
    p, exp = load_golden(name)
    r = run_hip(p, exp["pos0"], exp["vel0"], int(exp["steps"]))
    assert same_bits(r["pos"], exp["pos"])
    assert same_bits(r["vel"], exp["vel"])
    assert same_bits(r["time"], exp["time"])
    assert np.array_equal(r["nbr_count"], exp["nbr_count"])
    assert np.array_equal(r["nbr_ids"], exp["nbr_ids"])
    assert np.array_equal(r["cell"], exp["cell"])
    assert same_bits(r["rho"], exp["rho"])
    assert same_bits(r["lambda"], exp["lambda"])


def test_bruteforce_path_is_refused():
    p, exp = load_golden("bruteforce_120_1step")
    with pytest.raises(hip.HipError) as ei:
        hip.HipSolver(p, [exp["pos0"].shape[1]])
    assert ei.value.code == hip.EUNSUPPORTED


@pytest.mark.parametrize("flags", [False, True])
def test_dam_break_32k_grid_bit_exact(flags):
    sc = scenes.dam_break(seed=7, flags=flags)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1, diag=True)
    r = run_hip(sc.params, sc.state.pos, sc.state.vel, 1)
    # cell keys (integer cells of pred0)
    assert np.array_equal(r["cell"], o["cell"])
    # grid descriptor and sorted order: (scene, brick Z-order key, id)
    desc = expected_desc(o["cell"])
    assert np.array_equal(r["desc"][0], desc)
    keys = brick_keys(o["cell"], desc)
    order = np.lexsort((np.arange(sc.state.n), keys))
    assert np.array_equal(r["sorted_ids"], order)
    assert np.array_equal(r["keys"], keys[order])
    # neighbor counts and lists, in the reference's order
    assert np.array_equal(r["nbr_count"], o["nbr_count"])
    assert np.array_equal(r["nbr_ids"], o["nbr_ids"])
    # one step: within 1e-5 h, and in fact bitwise
    assert np.abs(r["pos"] - o["pos"]).max() <= POS_TOL
    assert same_bits(r["pos"], o["pos"])
    assert same_bits(r["vel"], o["vel"])
    assert same_bits(r["rho"], o["rho"])
    assert same_bits(r["lambda"], o["lambda"])


def test_double_dam_break_all_flags_262k_one_step():
    sc = scenes.double_dam_break()
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1, diag=True)
    r = run_hip(sc.params, sc.state.pos, sc.state.vel, 1)
    assert np.array_equal(r["cell"], o["cell"])
    assert np.array_equal(r["nbr_count"], o["nbr_count"])
    assert np.abs(r["pos"] - o["pos"]).max() <= POS_TOL
    assert np.abs(r["vel"] - o["vel"]).max() <= POS_TOL / float(sc.params.dt)
    assert same_bits(r["pos"], o["pos"])
    assert same_bits(r["vel"], o["vel"])


def test_multi_step_rollout_bitwise_and_density_stats():
    # 60 steps of the all-flags dam break: the kernels repeat the
    # reference's arithmetic exactly, so the rollout stays bitwise equal;
    # the density-error statistic is compared as well.
    sc = scenes.dam_break(seed=8, dims=(24, 24, 24), flags=True)
    steps = 60
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=steps, diag=True)
    r = run_hip(sc.params, sc.state.pos, sc.state.vel, steps)
    err_o = np.mean(np.abs(o["rho"] / 1000.0 - 1.0))
    err_r = np.mean(np.abs(r["rho"] / 1000.0 - 1.0))
    assert abs(err_r - err_o) <= 0.01 * err_o
    assert same_bits(r["pos"], o["pos"])
    assert same_bits(r["vel"], o["vel"])


def test_graph_replay_equals_eager():
    sc = scenes.dam_break(seed=9, dims=(20, 20, 20), flags=True)
    a = run_hip(sc.params, sc.state.pos, sc.state.vel, 5, graph=False)
    b = run_hip(sc.params, sc.state.pos, sc.state.vel, 5, graph=True)
    assert same_bits(a["pos"], b["pos"]) and same_bits(a["vel"], b["vel"])


@pytest.mark.parametrize("block,remap", [(64, 0), (128, 1), (512, 1), (1024, 0)])
def test_launch_shapes_do_not_change_results(block, remap):
    sc = scenes.dam_break(seed=10, dims=(20, 16, 12), flags=True)
    a = run_hip(sc.params, sc.state.pos, sc.state.vel, 3)
    b = run_hip(sc.params, sc.state.pos, sc.state.vel, 3, block_size=block, xcd_remap=remap)
    assert same_bits(a["pos"], b["pos"]) and same_bits(a["vel"], b["vel"])


@pytest.mark.parametrize("edge", [1, 2, 8])
def test_brick_edge_orders_cells_but_keeps_results(edge):
    sc = scenes.dam_break(seed=7, dims=(20, 16, 12), flags=True)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1, diag=True)
    r = run_hip(sc.params, sc.state.pos, sc.state.vel, 1, brick_edge=edge)
    desc = expected_desc(o["cell"], edge)
    assert np.array_equal(r["desc"][0], desc)
    keys = brick_keys(o["cell"], desc, edge)
    order = np.lexsort((np.arange(sc.state.n), keys))
    assert np.array_equal(r["sorted_ids"], order)
    assert np.array_equal(r["keys"], keys[order])
    assert same_bits(r["pos"], o["pos"]) and same_bits(r["vel"], o["vel"])
    a = run_hip(sc.params, sc.state.pos, sc.state.vel, 5)
    b = run_hip(sc.params, sc.state.pos, sc.state.vel, 5, brick_edge=edge)
    assert same_bits(a["pos"], b["pos"]) and same_bits(a["vel"], b["vel"])


def test_autotune_restores_state_and_keeps_results():
    sc = scenes.dam_break(seed=13, dims=(24, 20, 16), flags=True)
    n = sc.state.n
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel, 0.0)
        s.step(5)
        before = s.download()
        rep = s.autotune(trial_steps=2, knobs=hip.TUNE_WAVE_MODE | hip.TUNE_BLOCK |
                         hip.TUNE_BRICK | hip.TUNE_XCD)
        after = s.download()
        assert same_bits(before.pos, after.pos) and same_bits(before.vel, after.vel)
        assert after.time == before.time
        assert rep["block_size"] in (64, 128, 256, 512, 1024)
        assert rep["wave_mode"] in (1, 2, 3) and rep["brick_edge"] in (1, 2, 4, 8)
        assert rep["candidates_rejected"] == 0 and rep["candidates_timed"] >= 12
        assert 0 < rep["best_ms_per_step"] <= rep["start_ms_per_step"]
        s.step(7)
        tuned = s.download()
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel, 0.0)
        s.step(12)
        plain = s.download()
    assert same_bits(tuned.pos, plain.pos) and same_bits(tuned.vel, plain.vel)


def test_capacity_auto_growth_keeps_results():
    sc = scenes.dam_break(seed=11, dims=(16, 16, 16), flags=True)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=4)
    r = run_hip(sc.params, sc.state.pos, sc.state.vel, 4, cell_capacity=64,
                neighbor_capacity=64)
    assert same_bits(r["pos"], o["pos"]) and same_bits(r["vel"], o["vel"])


def test_profile_counts_each_substep_once_across_capacity_growth():
    """Per-kernel event totals count the launches of completed substeps
    only: a substep that overflowed a capacity is run again after the
    growth, and the substeps queued behind it did no work, so neither
    counts."""
    sc = scenes.dam_break(seed=11, dims=(16, 16, 16), flags=True)
    iters = sc.params.solver_iterations
    with hip.HipSolver(sc.params, [sc.state.n], cell_capacity=64, neighbor_capacity=64) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.profile(True)
        s.step(5)
        grown = s.stats()["grow_events"]
        counts = {k: s.profile_read(k)[1] for k in ("neighbors", "lambda", "delta", "xsph")}
        s.profile(False)
    assert grown >= 1
    assert counts == {"neighbors": 5, "lambda": 5 * iters, "delta": 5 * iters, "xsph": 5}


def test_non_finite_position_reports_range_error_and_keeps_state():
    sc = scenes.dam_break(seed=12, dims=(8, 8, 8))
    pos = sc.state.pos.copy()
    pos[1, 17] = np.nan
    with hip.HipSolver(sc.params, [sc.state.n]) as s:
        s.upload(pos, sc.state.vel, 0.0)
        with pytest.raises(hip.HipError) as ei:
            s.step(1)
        assert ei.value.code == hip.ERANGE
        st = s.download()
    assert same_bits(st.vel, sc.state.vel)
    assert np.isnan(st.pos[1, 17])


def test_empty_and_single_particle():
    p = scenes.base_params((1, 1, 1), True)
    with hip.HipSolver(p, [0]) as s:
        s.upload(np.zeros((3, 0), np.float32), np.zeros((3, 0), np.float32), 0.0)
        s.step(3)
        assert s.download().time == pytest.approx(3 * float(np.float32(p.dt)), rel=1e-6)
    x = np.array([[0.5], [0.5], [0.5]], np.float32)
    v = np.array([[0.1], [0.2], [0.3]], np.float32)
    o = oracle.step(p, x, v, nsteps=2)
    r = run_hip(p, x, v, 2, record_diag=False)
    assert same_bits(r["pos"], o["pos"]) and same_bits(r["vel"], o["vel"])


def test_hip_step_stateless_call_shape():
    # hip_step(params, state) has the call shape of cuda_step(params, state)
    sc = scenes.dam_break(seed=13, dims=(10, 10, 10), flags=True)
    st = sc.state.copy()
    hip.hip_step(sc.params, st)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1)
    assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"])
    assert st.time == pytest.approx(float(o["time"]))
