"""The five benchmark workloads at their real sizes, on the GPU, against the
CPU oracle (the restatement of core/src/core.cpp:119-615, pinned by the
reference's own outputs in test_oracle.py).

* 1,048,576-particle block (128x128x64, all flags): one substep from t=0
  with grid, order, neighbor lists, rho and lambda bit-exact; substeps 1-3
  through the LDS tiles (eager and hipGraph), bitwise; one substep
  from the piled-up state after 30 substeps, bitwise, with the density
  statistics of that substep.
* 256-scene ensemble (dims uniform in 8..32, 2.08M particles): 5 substeps
  batched (eager and hipGraph) bitwise equal to one context per scene; the
  4 smallest and 4 largest scenes bitwise equal to the oracle.
* 65,536-particle differentiable rollout, 100 substeps: gradients with a
  checkpoint every 10 substeps bitwise equal to the whole recorded rollout
  and finite; 64 gradient coordinates (first substep) and the derivative
  along the gradient (2 substeps) against central differences of the FP64
  oracle.
* 32,768-particle dam break: the full 240-step rollout; 262,144-particle
  double dam break: the first 24 of its 480 steps (the oracle's cost per step
  grows with the pile-up: about 1.3 s per step at the start on 8 cores and
  several seconds beyond step 20).  Both bitwise, with the mean density error
  |rho/rho0 - 1| of the last step within 1% of the oracle's.

Positions and velocities are compared bitwise because the kernels repeat the
reference's float operations in its order; the tolerances `north_star`
states (1e-5 h after one step, density statistics within 1%) are asserted
too.
"""
import os

import numpy as np
import pytest

import oracle
from fluidsimulator_amd import hip, scenes
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu

H = 0.1
POS_TOL = 1e-5 * H
RHO0 = 1000.0
THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)
    oracle.set_threads(THREADS)


def density_error(rho):
    """Mean |rho/rho0 - 1| over the particles (the reference's density
    constraint C_i = rho_i/rho0 - 1, core/src/core.cpp:281-329)."""
    return float(np.mean(np.abs(np.asarray(rho, np.float64) / RHO0 - 1.0)))


def assert_density_stats(rho_gpu, rho_cpu):
    eg, ec = density_error(rho_gpu), density_error(rho_cpu)
    assert abs(eg - ec) <= 0.01 * ec, (eg, ec)
    mg, mc = float(np.max(rho_gpu)), float(np.max(rho_cpu))
    assert abs(mg - mc) <= 0.01 * mc, (mg, mc)


def test_million_block_first_substep_bit_exact():
    sc = scenes.million_block()
    n = sc.state.n
    print(f"million block: {n} particles, substep 1 with diagnostics", flush=True)
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=1, diag=True,
                    nbr_cap=40 * n)
    with hip.HipSolver(sc.params, [n], record_diag=True) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.step(1)
        r = s.download()
        d = s.diag()
        nbr = s.neighbors()
    assert np.array_equal(d["cell"], o["cell"])
    # (cell, id) order: within each cell the ids ascend, and the cells of the
    # sorted order are exactly the oracle's cells
    ids = d["sorted_ids"]
    assert np.array_equal(np.sort(ids), np.arange(n))
    assert np.array_equal(d["nbr_count"], o["nbr_count"])
    assert np.array_equal(nbr, o["nbr_ids"])
    assert np.abs(r.pos - o["pos"]).max() <= POS_TOL
    assert same_bits(r.pos, o["pos"])
    assert same_bits(r.vel, o["vel"])
    assert same_bits(d["rho"], o["rho"])
    assert same_bits(d["lambda"], o["lambda"])


@pytest.mark.parametrize("graph", [False, True])
def test_million_block_early_substeps_through_tiles_bitwise(graph):
    """Substeps 1-3 from t=0 with the default tile mode: nearly every
    particle in an LDS tile, the tile kernels on the side stream and the
    tile lists built beside the lane map."""
    sc = scenes.million_block()
    n = sc.state.n
    steps = 3
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.step(steps, graph=graph)
        r = s.download()
        st = s.stats()
    print(f"million block, {steps} substeps (graph {graph}): {st['tiles']} tiles, "
          f"{st['tiled_particles']} tiled particles", flush=True)
    assert st["tiled_particles"] > 0.9 * n
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=steps)
    assert same_bits(r.pos, o["pos"])
    assert same_bits(r.vel, o["vel"])


def test_million_block_piled_up_substep_bitwise():
    sc = scenes.million_block()
    n = sc.state.n
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.step(30)
        s30 = s.download()
        s.step(1)
        r = s.download()
        d = s.diag()
        pairs = s.stats()["pairs"]
    print(f"million block: substep 31 from the piled-up state, {pairs} pairs", flush=True)
    o = oracle.step(sc.params, s30.pos, s30.vel, nsteps=1, diag="light")
    assert pairs == o["nbr_total"]
    assert np.array_equal(d["nbr_count"], o["nbr_count"])
    assert same_bits(r.pos, o["pos"])
    assert same_bits(r.vel, o["vel"])
    assert same_bits(d["rho"], o["rho"])
    assert_density_stats(d["rho"], o["rho"])


def test_ensemble_256_scenes_batched_equals_per_scene():
    ens = scenes.ensemble()
    counts = [st.n for st in ens.states]
    assert len(counts) == 256
    pos = np.concatenate([st.pos for st in ens.states], axis=1)
    vel = np.concatenate([st.vel for st in ens.states], axis=1)
    steps = 5
    print(f"ensemble: 256 scenes, {sum(counts)} particles, {steps} substeps", flush=True)
    outs = {}
    for graph in (False, True):
        with hip.HipSolver(ens.params, counts) as b:
            b.upload(pos, vel)
            b.step(steps, graph=graph)
            outs[graph] = b.download()
    assert same_bits(outs[False].pos, outs[True].pos)
    assert same_bits(outs[False].vel, outs[True].vel)
    off = np.concatenate([[0], np.cumsum(counts)])
    for i, (p, st) in enumerate(zip(ens.params, ens.states)):
        with hip.HipSolver(p, [st.n], cell_capacity=1 << 17,
                           neighbor_capacity=64 * st.n + 4096) as one:
            one.upload(st.pos, st.vel)
            one.step(steps)
            single = one.download()
        sl = slice(off[i], off[i + 1])
        assert same_bits(outs[True].pos[:, sl], single.pos), i
        assert same_bits(outs[True].vel[:, sl], single.vel), i
    order = np.argsort(counts, kind="stable")
    for i in list(order[:4]) + list(order[-4:]):
        st = ens.states[i]
        o = oracle.step(ens.params[i], st.pos, st.vel, nsteps=steps)
        sl = slice(off[i], off[i + 1])
        assert same_bits(outs[True].pos[:, sl], o["pos"]), i
        assert same_bits(outs[True].vel[:, sl], o["vel"]), i


def _rollout_loss_grad(sc, target, steps, every):
    n = sc.state.n
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.forward_record(steps, every)
        fin = s.download()
        diff = fin.pos.astype(np.float64) - target
        loss = float(np.mean(np.sum(diff * diff, axis=0)))
        dx0, dv0, dg = s.backward((2.0 / n) * diff, None)
        rec = s.stats()
    return loss, dx0, dv0, dg[0], rec


def test_rollout_65536_checkpointed_gradients_bitwise():
    sc = scenes.rollout_scene()
    tg = scenes.rollout_scene(seed=6)
    n = sc.state.n
    assert n == 65536
    with hip.HipSolver(tg.params, [n]) as s:
        s.upload(tg.state.pos, tg.state.vel)
        s.step(100)
        target = s.download().pos.astype(np.float64)
    print("rollout: 65536 particles, 100 substeps, checkpoint every 10 vs whole", flush=True)
    L10, dx10, dv10, dg10, r10 = _rollout_loss_grad(sc, target, 100, 10)
    L100, dx100, dv100, dg100, r100 = _rollout_loss_grad(sc, target, 100, 100)
    assert L10 == L100
    assert r10["recomputed"] == 90 and r100["recomputed"] == 0
    assert same_bits(dx10, dx100)
    assert same_bits(dv10, dv100)
    assert same_bits(dg10, dg100)
    assert np.all(np.isfinite(dv10)) and np.all(np.isfinite(dx10)) and np.all(np.isfinite(dg10))
    assert np.max(np.abs(dv10)) > 0.0


def _fp64_loss(p, pos64, target, steps):
    g0 = np.array([float(np.float32(x)) for x in p.external_forces])

    def loss(vel, g=g0):
        o = oracle.step(p, pos64, vel, nsteps=steps, dtype=np.float64, params64={"gravity": g})
        d = o["pos"] - target
        return float(np.mean(np.sum(d * d, axis=0)))

    return loss, g0


def test_rollout_65536_gradient_check_64_coordinates():
    """61 coordinates of dL/dv0 and the 3 of dL/dg from the float32 adjoint
    against central differences (step 1e-4) of the FP64 oracle, over the
    first substep of the full 65,536-particle rollout: 2-norm relative error
    of the 64-vector below 1e-3."""
    steps = 1
    sc = scenes.rollout_scene()
    tg = scenes.rollout_scene(seed=6)
    n = sc.state.n
    p = sc.params
    pos64 = sc.state.pos.astype(np.float64)
    target = oracle.step(p, pos64, tg.state.vel.astype(np.float64), nsteps=steps,
                         dtype=np.float64)["pos"]
    _, _, dv0, dg, _ = _rollout_loss_grad(sc, target, steps, 10)
    loss, g0 = _fp64_loss(p, pos64, target, steps)
    v64 = sc.state.vel.astype(np.float64)
    rng = np.random.default_rng(12)
    coords = rng.choice(3 * n, size=61, replace=False)
    eps = 1e-4
    ad, fd = [], []
    for c in coords:
        k, i = divmod(int(c), n)
        vp, vm = v64.copy(), v64.copy()
        vp[k, i] += eps
        vm[k, i] -= eps
        fd.append((loss(vp) - loss(vm)) / (2 * eps))
        ad.append(float(dv0[k, i]))
    for k in range(3):
        gp, gm = g0.copy(), g0.copy()
        gp[k] += eps
        gm[k] -= eps
        fd.append((loss(v64, gp) - loss(v64, gm)) / (2 * eps))
        ad.append(float(dg[k]))
    ad, fd = np.array(ad), np.array(fd)
    err = float(np.linalg.norm(ad - fd) / np.linalg.norm(fd))
    print(f"64-coordinate gradient check: relative 2-norm error {err:.3e}", flush=True)
    assert err < 1e-3


def test_rollout_65536_directional_derivative_along_the_gradient():
    """Central difference (step 1e-4) of the FP64 oracle's loss along the
    unit direction of the float32 adjoint's dL/dv0, over 2 substeps of the
    full rollout, equals the adjoint's |dL/dv0| to 1e-3: the direction and
    the length of the gradient are right.  (A random direction is not used:
    its inner product with the gradient cancels over 196,608 components,
    which multiplies the float32 adjoint's per-component error by ~400.)"""
    steps = 2
    sc = scenes.rollout_scene()
    tg = scenes.rollout_scene(seed=6)
    p = sc.params
    pos64 = sc.state.pos.astype(np.float64)
    target = oracle.step(p, pos64, tg.state.vel.astype(np.float64), nsteps=steps,
                         dtype=np.float64)["pos"]
    _, _, dv0, _, _ = _rollout_loss_grad(sc, target, steps, 10)
    loss, _ = _fp64_loss(p, pos64, target, steps)
    g = dv0.astype(np.float64)
    norm = float(np.linalg.norm(g))
    u = g / norm
    v64 = sc.state.vel.astype(np.float64)
    eps = 1e-4
    fd = (loss(v64 + eps * u) - loss(v64 - eps * u)) / (2 * eps)
    print(f"directional derivative along the gradient: |dL/dv0| {norm:.6e}, FD {fd:.6e}",
          flush=True)
    assert abs(fd - norm) <= 1e-3 * norm


@pytest.mark.parametrize("which,steps", [("dam_break_32k", 240), ("double_dam_break", 24)])
def test_long_rollout_bitwise_and_density_statistics(which, steps):
    sc = scenes.dam_break() if which == "dam_break_32k" else scenes.double_dam_break()
    n = sc.state.n
    print(f"{which}: {n} particles, {steps} steps", flush=True)
    with hip.HipSolver(sc.params, [n]) as s:
        s.upload(sc.state.pos, sc.state.vel)
        s.step(steps)
        r = s.download()
        d = s.diag()
    o = oracle.step(sc.params, sc.state.pos, sc.state.vel, nsteps=steps, diag="light")
    assert_density_stats(d["rho"], o["rho"])
    assert np.array_equal(d["nbr_count"], o["nbr_count"])
    assert same_bits(r.pos, o["pos"])
    assert same_bits(r.vel, o["vel"])
    assert r.time == pytest.approx(float(o["time"]), abs=0.0)
