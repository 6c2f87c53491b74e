"""The LDS-tiled passes of sparse bricks (pbf_tiles.hip) against the CPU
oracle and against the per-lane passes, bitwise.

A brick of 4x4x4 cells runs through LDS when its 6^3-cell halo holds at most
2,048 particles, one of its cells is good and no particle's list outgrows
the tile list (256 entries below 2^18 particles, 512 from there on, where a
brick also needs 32 particles).  A good cell has fewer than 512 candidates
(particles in its 27 cells) and, from 2^18 particles on, no coincident pair;
the brick's other cells, and the bricks that are no tile, take the per-lane
passes in the same substep.  Every sum visits the neighbors in the
reference's order (core.cpp:205-247), so the mixed substep must equal the
oracle bit for bit."""
import numpy as np
import pytest

import oracle
from fluidsimulator_amd import hip, scenes
from tests.helpers import golden_names, load_golden, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)


def run(p, pos, vel, steps, counts=None, **kw):
    counts = counts or [pos.shape[1]]
    with hip.HipSolver(p, counts, **kw) as s:
        s.upload(pos, vel, 0.0)
        s.step(steps)
        out = s.download()
        d = s.diag()
        st = s.stats()
    return out, d, st


@pytest.mark.parametrize("name", [n for n in golden_names() if not n.startswith(("bruteforce", "powf_"))])
def test_goldens_with_tiles(name):
    p, exp = load_golden(name)
    out, d, st = run(p, exp["pos0"], exp["vel0"], int(exp["steps"]), tile_gather=1)
    assert same_bits(out.pos, exp["pos"])
    assert same_bits(out.vel, exp["vel"])
    assert same_bits(d["rho"], exp["rho"])
    assert same_bits(d["lambda"], exp["lambda"])
    assert np.array_equal(d["nbr_count"], exp["nbr_count"])


@pytest.mark.parametrize("flags", [False, True])
def test_dam_break_tiles_equal_oracle_and_per_lane(flags):
    sc = scenes.dam_break(seed=7, dims=(32, 32, 32), flags=flags)
    steps = 8
    on, d_on, st_on = run(sc.params, sc.state.pos, sc.state.vel, steps, tile_gather=1)
    off, d_off, st_off = run(sc.params, sc.state.pos, sc.state.vel, steps, tile_gather=2)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=steps, diag="light")
    print(f"tiles in the last substep: {st_on['tiles']} (off: {st_off['tiles']})", flush=True)
    assert st_on["tiles"] > 0 and st_off["tiles"] == 0
    # the decoded neighbor lists of tile particles, in the reference's order
    with hip.HipSolver(sc.params, [sc.state.n], tile_gather=1) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.step(1)
        assert s.stats()["tiles"] > 0
        lists = s.neighbors()
    o1 = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1, diag=True)
    assert np.array_equal(lists, o1["nbr_ids"])
    for r, d in ((on, d_on), (off, d_off)):
        assert same_bits(r.pos, o["pos"])
        assert same_bits(r.vel, o["vel"])
        assert same_bits(d["rho"], o["rho"])
        assert np.array_equal(d["nbr_count"], o["nbr_count"])


def test_tiles_next_to_piled_up_cells():
    """A rollout deep enough that walls and corners pile up: dense cells take
    the per-lane passes, sparse bricks the tiles, in the same substeps."""
    sc = scenes.dam_break(seed=8, dims=(24, 24, 24), box=(1.6, 1.6, 1.6), flags=True)
    steps = 40
    on, d_on, st_on = run(sc.params, sc.state.pos, sc.state.vel, steps, tile_gather=1)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=steps, diag="light")
    print(f"pairs {st_on['pairs']}, tiles {st_on['tiles']}", flush=True)
    assert st_on["tiles"] > 0
    assert same_bits(on.pos, o["pos"])
    assert same_bits(on.vel, o["vel"])
    assert same_bits(d_on["rho"], o["rho"])


@pytest.mark.parametrize("flags", [False, True])
def test_tiles_with_long_lists_go_back_to_the_lanes(flags):
    """A cluster of 7^3 particles at spacing 0.01 (h = 0.1): its inner
    particles have more neighbors than the 256 a tile list of a small batch
    holds, so its brick goes back to the per-lane passes in the same substep,
    while the lattice block at spacing 0.05 beside it stays tiles."""
    sc = scenes.dam_break(seed=9, dims=(20, 20, 20), flags=flags)
    a = scenes.lattice_block(7, 7, 7, d=0.01, origin=(0.52, 0.52, 0.52), seed=9)
    b = scenes.lattice_block(16, 16, 16, d=0.05, origin=(2.0, 0.2, 0.2), seed=9, stream=1)
    pos = np.concatenate([a, b], axis=1)
    vel = np.zeros_like(pos)
    steps = 3
    out, d, st = run(sc.params, pos, vel, steps, tile_gather=1)
    o = oracle.step(sc.params, pos, vel, nsteps=steps, diag="light")
    _, d1, st1 = run(sc.params, pos, vel, 1, tile_gather=1)
    o1 = oracle.step(sc.params, pos, vel, nsteps=1, diag="light")
    print(f"substep 1: tiles {st1['tiles']} of their particles {st1['tiled_particles']}, "
          f"max neighbors {int(o1['nbr_count'].max())}", flush=True)
    assert int(o1["nbr_count"][: a.shape[1]].max()) > 256
    assert st1["tiled_particles"] == b.shape[1]
    assert np.array_equal(d1["nbr_count"], o1["nbr_count"])
    assert same_bits(out.pos, o["pos"])
    assert same_bits(out.vel, o["vel"])
    assert same_bits(d["rho"], o["rho"])
    assert np.array_equal(d["nbr_count"], o["nbr_count"])


@pytest.mark.parametrize("graph", [False, True])
def test_batched_scenes_with_tiles(graph):
    ens = scenes.ensemble(seed=23, n_scenes=10, lo=6, hi=20)
    counts = [s.n for s in ens.states]
    pos = np.concatenate([s.pos for s in ens.states], axis=1)
    vel = np.concatenate([s.vel for s in ens.states], axis=1)
    steps = 5
    with hip.HipSolver(ens.params, counts, tile_gather=1) as b:
        b.upload(pos, vel)
        b.step(steps, graph=graph)
        out = b.download()
        assert b.stats()["tiles"] > 0
    off = np.concatenate([[0], np.cumsum(counts)])
    for i, (p, st) in enumerate(zip(ens.params, ens.states)):
        o = oracle.step(p, st.pos, st.vel, nsteps=steps)
        sl = slice(off[i], off[i + 1])
        assert same_bits(out.pos[:, sl], o["pos"]), i
        assert same_bits(out.vel[:, sl], o["vel"]), i


def test_brick_edges_without_tiles_agree():
    """Tiles need 4-cell bricks; other brick edges take the per-lane passes.
    The default tile mode tiles from 32,768 particles, so this 3,840-particle
    scene has none; asked for, it has them."""
    sc = scenes.dam_break(seed=12, dims=(20, 16, 12), flags=True)
    ref, _, st4 = run(sc.params, sc.state.pos, sc.state.vel, 4, brick_edge=4, tile_gather=1)
    assert st4["tiles"] > 0
    _, _, st0 = run(sc.params, sc.state.pos, sc.state.vel, 1, brick_edge=4)
    assert st0["tiles"] == 0
    for edge in (1, 2, 8):
        r, _, st = run(sc.params, sc.state.pos, sc.state.vel, 4, brick_edge=edge, tile_gather=1)
        assert st["tiles"] == 0
        assert same_bits(r.pos, ref.pos) and same_bits(r.vel, ref.vel)


def _large_batch_scene(pair=False, cluster=False):
    """64^3 particles at spacing 0.0625 (h = 0.1: about 4 per cell, 2^18 in
    all) floating in a 6 m box.  `pair` moves particle a+1 onto particle a
    (one coincident pair in consecutive slots); `cluster` adds 450 particles
    in a 0.01 m cube inside the block, so the cluster's cell and its 26
    neighbors have more than 512 candidates while the halos stay under
    2,048."""
    nx = 64
    pos = scenes.lattice_block(nx, nx, nx, d=0.0625, origin=(0.5, 0.5, 0.5), seed=31)
    if pair:
        a = (32 * nx + 32) * nx + 32
        pos[:, a + 1] = pos[:, a]
    if cluster:
        c0 = pos[:, (20 * nx + 20) * nx + 44].astype(np.float64)
        u = scenes.uniform01(32, 3 * 450).reshape(3, 450)
        pos = np.concatenate([pos, (c0[:, None] + 0.01 * (u - 0.5)).astype(np.float32)], axis=1)
    return scenes.base_params((6.0, 6.0, 6.0), True), pos


def test_large_batch_tiles_leave_only_dense_and_coincident_cells_to_the_lanes():
    """From 2^18 particles on, a cell with a coincident pair or with 512
    candidates or more goes to the per-lane passes and the rest of its brick
    stays a tile: a pair costs the tiles its one cell (about 4 particles,
    where its brick holds about 260), the cluster its 27 cells.  Tiles on,
    tiles off and the oracle agree bit for bit."""
    untiled = {}
    for name, kw in (("plain", {}), ("pair", {"pair": True}), ("cluster", {"cluster": True})):
        p, pos = _large_batch_scene(**kw)
        vel = np.zeros_like(pos)
        on, d_on, st_on = run(p, pos, vel, 1, tile_gather=1)
        off, d_off, st_off = run(p, pos, vel, 1, tile_gather=2)
        o = oracle.step(p, pos, vel, nsteps=1, diag="light")
        untiled[name] = pos.shape[1] - st_on["tiled_particles"]
        print(f"{name}: n {pos.shape[1]}, tiles {st_on['tiles']}, "
              f"untiled {untiled[name]}", flush=True)
        assert st_on["tiles"] > 0 and st_off["tiles"] == 0
        for r, d in ((on, d_on), (off, d_off)):
            assert same_bits(r.pos, o["pos"]), name
            assert same_bits(r.vel, o["vel"]), name
            assert same_bits(d["rho"], o["rho"]), name
            assert np.array_equal(d["nbr_count"], o["nbr_count"]), name
    assert 2 <= untiled["pair"] - untiled["plain"] <= 40
    assert 450 < untiled["cluster"] - untiled["plain"] <= 450 + 64 * 12


def test_large_batch_mixed_bricks_over_substeps():
    """The pair and the cluster together over 3 substeps (the cluster blows
    apart): tiles on equal the oracle bit for bit."""
    p, pos = _large_batch_scene(pair=True, cluster=True)
    vel = np.zeros_like(pos)
    steps = 3
    on, d_on, st_on = run(p, pos, vel, steps, tile_gather=1)
    o = oracle.step(p, pos, vel, nsteps=steps, diag="light")
    print(f"tiles {st_on['tiles']}, tiled {st_on['tiled_particles']} of {pos.shape[1]}", flush=True)
    assert st_on["tiles"] > 0
    assert same_bits(on.pos, o["pos"])
    assert same_bits(on.vel, o["vel"])
    assert same_bits(d_on["rho"], o["rho"])
    assert np.array_equal(d_on["nbr_count"], o["nbr_count"])
