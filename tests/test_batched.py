"""Segmented batch: many independent scenes of uneven size in one launch per
phase must equal stepping each scene alone (HIP and CPU oracle), bitwise."""
import numpy as np
import pytest

import oracle
from fluidsimulator_amd import hip, scenes
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)


def _concat(states):
    pos = np.concatenate([s.pos for s in states], axis=1)
    vel = np.concatenate([s.vel for s in states], axis=1)
    return pos, vel


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_per_scene_and_oracle(graph):
    ens = scenes.ensemble(seed=21, n_scenes=12, lo=4, hi=12)
    counts = [s.n for s in ens.states]
    pos, vel = _concat(ens.states)
    steps = 6
    with hip.HipSolver(ens.params, counts) as b:
        b.upload(pos, vel)
        b.step(steps, graph=graph)
        out = b.download()
    off = np.concatenate([[0], np.cumsum(counts)])
    for s, (p, st) in enumerate(zip(ens.params, ens.states)):
        with hip.HipSolver(p, [st.n]) as one:
            one.upload(st.pos, st.vel)
            one.step(steps)
            single = one.download()
        o = oracle.step(p, st.pos, st.vel, nsteps=steps)
        sl = slice(off[s], off[s + 1])
        assert same_bits(out.pos[:, sl], single.pos), s
        assert same_bits(out.vel[:, sl], single.vel), s
        assert same_bits(out.pos[:, sl], o["pos"]), s
        assert same_bits(out.vel[:, sl], o["vel"]), s


def test_batch_with_empty_scene():
    ens = scenes.ensemble(seed=22, n_scenes=3, lo=4, hi=6)
    empty = scenes.State.empty(0) if hasattr(scenes, "State") else None
    params = [ens.params[0], ens.params[1], ens.params[2]]
    counts = [ens.states[0].n, 0, ens.states[2].n]
    pos = np.concatenate([ens.states[0].pos, ens.states[2].pos], axis=1)
    vel = np.concatenate([ens.states[0].vel, ens.states[2].vel], axis=1)
    del empty
    with hip.HipSolver(params, counts) as b:
        b.upload(pos, vel)
        b.step(3)
        out = b.download()
    o0 = oracle.step(params[0], ens.states[0].pos, ens.states[0].vel, nsteps=3)
    o2 = oracle.step(params[2], ens.states[2].pos, ens.states[2].vel, nsteps=3)
    n0 = counts[0]
    assert same_bits(out.pos[:, :n0], o0["pos"])
    assert same_bits(out.pos[:, n0:], o2["pos"])
