"""Dense-cell regime: with these parameters the dam break piles particles
onto walls and corners (hundreds to tens of thousands per cell, mostly
exactly coincident).  Wavefronts that share a cell take the shared-tile and
coincident-tile paths there; every wave mode must stay bitwise equal to the
CPU oracle."""
import numpy as np
import pytest

import oracle
from fluidsimulator_amd import hip, scenes
from fluidsimulator_amd.params import Params
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu

WAVE_MODES = {1: "per-lane lists", 2: "shared tiles + coincident tiles", 3: "shared tiles"}


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)


@pytest.fixture(scope="module")
def pileup_reference():
    # 131,072 particles, all flags; by step 40 cells hold several hundred
    # coincident particles.
    sc = scenes.million_block(dims=(64, 64, 32))
    checkpoints = (10, 25, 40)
    out, pos, vel, prev = {}, sc.state.pos, sc.state.vel, 0
    for k in checkpoints:
        r = oracle.step(sc.params, pos, vel, nsteps=k - prev)
        pos, vel, prev = r["pos"], r["vel"], k
        out[k] = (pos, vel)
    return sc, out


@pytest.mark.parametrize("mode", sorted(WAVE_MODES))
def test_pileup_rollout_bitwise(pileup_reference, mode):
    sc, ref = pileup_reference
    with hip.HipSolver(sc.params, [sc.state.n], wave_mode=mode) as s:
        s.upload(sc.state.pos, sc.state.vel)
        prev = 0
        for k, (pos, vel) in ref.items():
            s.step(k - prev)
            prev = k
            st = s.download()
            assert same_bits(st.pos, pos), (WAVE_MODES[mode], k)
            assert same_bits(st.vel, vel), (WAVE_MODES[mode], k)


def _clusters(seed=3):
    # coincident clusters in neighboring cells (some cells hold two
    # clusters), with scattered particles around them
    rng = np.random.default_rng(seed)
    pts = []
    centers = [(0.52, 0.51, 0.53), (0.58, 0.51, 0.53), (0.45, 0.43, 0.61), (0.55, 0.66, 0.47),
               (0.52, 0.51, 0.44), (0.61, 0.59, 0.52)]
    for k, c in enumerate(centers):
        pts.append(np.repeat(np.array(c, np.float32)[:, None], 90 + 20 * k, axis=1))
    pts.append((0.53 + 0.06 * rng.standard_normal((3, 300))).astype(np.float32))
    pos = np.concatenate(pts, axis=1)
    perm = rng.permutation(pos.shape[1])  # interleave ids across clusters
    return np.ascontiguousarray(pos[:, perm])


@pytest.mark.parametrize("mode", sorted(WAVE_MODES))
@pytest.mark.parametrize("flags", [False, True])
def test_coincident_clusters_bitwise(mode, flags):
    p = scenes.base_params((1.0, 1.0, 1.0), flags)
    pos = _clusters()
    vel = np.zeros_like(pos)
    vel[1] = -0.5
    o = oracle.step(p, pos, vel, nsteps=3, diag=True)
    with hip.HipSolver(p, [pos.shape[1]], wave_mode=mode, record_diag=True) as s:
        s.upload(pos, vel)
        s.step(3)
        st = s.download()
        d = s.diag()
        nb = s.neighbors()
    assert np.array_equal(d["nbr_count"], o["nbr_count"])
    assert np.array_equal(nb, o["nbr_ids"])
    assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"])


def test_all_particles_at_one_point():
    p = Params(dt=float(np.float32(1 / 120)), density=1000.0, particle_mass=0.125, h=0.1,
               epsilon=100.0, solver_iterations=4, external_forces=(0.0, -9.8, 0.0),
               enable_xsph=True, visc_c=0.01, enable_vorticity=True, vort_epsilon=5e-4)
    p.set_box(1.0, 1.0, 1.0)
    pos = np.full((3, 700), 0.5, np.float32)
    vel = np.zeros_like(pos)
    o = oracle.step(p, pos, vel, nsteps=2)
    for mode in sorted(WAVE_MODES):
        with hip.HipSolver(p, [pos.shape[1]], wave_mode=mode) as s:
            s.upload(pos, vel)
            s.step(2)
            st = s.download()
        assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"]), WAVE_MODES[mode]


@pytest.mark.parametrize("flags", [False, True])
def test_runs_of_identical_particles_exactly(flags):
    """The default gather mode computes each run of identical consecutive
    slots (same position, velocity and cell, bit for bit) once and copies the
    run's results to its other slots; walks take long runs as one step.  Long
    runs of identical particles in one cell, broken by a few others at nearby
    points (so that lists differ between runs), must still give the
    reference's results bit for bit, with and without runs (diagnostics turn
    them off)."""
    p = scenes.base_params((1.0, 1.0, 1.0), flags)
    rng = np.random.default_rng(11)
    big = np.repeat(np.array([[0.512], [0.507], [0.503]], np.float32), 1200, axis=1)
    others = np.array([[0.515, 0.509, 0.501], [0.52, 0.5, 0.51], [0.508, 0.514, 0.506]],
                      np.float32).T
    others = np.repeat(others, 3, axis=1)
    far = (0.5 + 0.05 * rng.standard_normal((3, 400))).astype(np.float32)
    pos = np.concatenate([big, others, far], axis=1)
    order = rng.permutation(pos.shape[1])
    pos = np.ascontiguousarray(pos[:, order])
    vel = np.zeros_like(pos)
    vel[1] = -0.5
    o = oracle.step(p, pos, vel, nsteps=3)
    for diag in (False, True):  # diagnostics turn runs off
        with hip.HipSolver(p, [pos.shape[1]], record_diag=diag) as s:
            s.upload(pos, vel)
            s.step(3)
            st = s.download()
        assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"]), diag


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("flags", [False, True])
def test_coincident_clusters_bitwise_with_skips(flags, mode):
    """Without diagnostics, particles of dense cells store no neighbor list
    (their passes retest the distance predicate) and runs of identical slots
    are computed once; results stay bitwise and the pair statistics still
    count every particle's neighbors."""
    p = scenes.base_params((1.0, 1.0, 1.0), flags)
    pos = _clusters()
    vel = np.zeros_like(pos)
    vel[1] = -0.5
    o = oracle.step(p, pos, vel, nsteps=3, diag=True)
    with hip.HipSolver(p, [pos.shape[1]], wave_mode=mode) as s:
        s.upload(pos, vel)
        s.step(3)
        st = s.download()
        pairs = s.stats()["pairs"]
    assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"])
    assert pairs == int(o["nbr_count"].astype(np.int64).sum())


@pytest.mark.parametrize("flags", [False, True])
def test_lane_map_fallback_bitwise(flags):
    """Dense cells of one run each need a whole wave of lanes apiece; with the
    lane capacity at n that layout does not fit, and the passes fall back to
    one packed lane per run.  Both layouts give the reference's results bit
    for bit."""
    p = scenes.base_params((1.0, 1.0, 1.0), flags)
    rng = np.random.default_rng(21)
    pts = []
    for k in range(24):  # 24 cells of 40 identical particles each
        c = (0.15 + 0.11 * (k % 6), 0.25 + 0.13 * (k // 6), 0.5)
        pts.append(np.repeat(np.array(c, np.float32)[:, None], 40, axis=1))
    pts.append((0.5 + 0.15 * rng.standard_normal((3, 200))).astype(np.float32).clip(0.05, 0.95))
    pos = np.concatenate(pts, axis=1)
    pos = np.ascontiguousarray(pos[:, rng.permutation(pos.shape[1])])
    vel = np.zeros_like(pos)
    vel[1] = -0.5
    o = oracle.step(p, pos, vel, nsteps=3)
    n = pos.shape[1]
    for cap in (0, n):
        with hip.HipSolver(p, [n], lane_capacity=cap) as s:
            s.upload(pos, vel)
            s.step(3)
            st = s.download()
        assert same_bits(st.pos, o["pos"]) and same_bits(st.vel, o["vel"]), cap
