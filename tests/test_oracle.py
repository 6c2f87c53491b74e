"""Pins the CPU oracle (oracle/pbf_oracle.cpp) before anything is checked
against it: golden vectors from the reference's own core.cpp, the live
reference library, closed forms, brute force, and planted mistakes."""
import math

import numpy as np
import pytest

import oracle
from fluidsimulator_amd import scenes
from fluidsimulator_amd.params import Params
from tests.helpers import golden_names, load_golden, neighbor_lists, same_bits


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_golden_bitwise(name):
    p, exp = load_golden(name)
    r = oracle.step(p, exp["pos0"], exp["vel0"], nsteps=int(exp["steps"]), diag=True)
    assert same_bits(r["pos"], exp["pos"])
    assert same_bits(r["vel"], exp["vel"])
    assert same_bits(np.float32(r["time"]), exp["time"])
    assert np.array_equal(r["nbr_count"], exp["nbr_count"])
    assert np.array_equal(r["nbr_ids"], exp["nbr_ids"])
    assert same_bits(r["rho"], exp["rho"])
    assert same_bits(r["lambda"], exp["lambda"])
    if p.use_uniform_grid:
        assert np.array_equal(r["cell"], exp["cell"])
        assert np.array_equal(r["sorted_ids"], exp["sorted_ids"])


@pytest.mark.skipif(not oracle.have_ref(), reason="reference library not built here")
@pytest.mark.parametrize("flags,seed", [(False, 101), (True, 102), (True, 103)])
def test_oracle_matches_reference_library(flags, seed):
    sc = scenes.dam_break(seed=seed, dims=(14, 12, 10), flags=flags)
    vel = scenes.normal(seed, 3 * sc.state.n, sigma=0.7).reshape(3, -1).astype(np.float32)
    a = oracle.step(sc.params, sc.state.pos, vel, nsteps=4, diag=True)
    b = oracle.step(sc.params, sc.state.pos, vel, nsteps=4, diag=True, which="ref")
    for k in ("pos", "vel", "rho", "lambda"):
        assert same_bits(a[k], b[k]), k
    for k in ("nbr_count", "nbr_ids", "cell", "sorted_ids"):
        assert np.array_equal(a[k], b[k]), k


def test_kernel_coefficients_closed_form():
    h = 0.1
    w0 = oracle.poly6_f32(0.0, h)
    assert w0 == pytest.approx(315.0 / (64.0 * math.pi * h ** 3), rel=1e-6)
    assert w0 == pytest.approx(1566.68, rel=1e-5)
    r2 = 0.05 ** 2
    assert oracle.poly6_f32(r2, h) == pytest.approx(
        315.0 / (64.0 * math.pi * h ** 9) * (h * h - r2) ** 3, rel=1e-5)
    assert oracle.poly6_f32(h * h * 1.01, h) == 0.0
    # spiky gradient factor -45/(pi h^6) (h - r)^2, multiplied by (pi - pj)
    # directly (not divided by r)
    assert oracle.spiky_f32(0.0, h) == pytest.approx(-45.0 / (math.pi * h ** 4), rel=1e-5)
    assert oracle.spiky_f32(0.05, h) == pytest.approx(-45.0 / (math.pi * h ** 6) * 0.05 ** 2,
                                                      rel=1e-5)
    assert oracle.spiky_f32(0.1001, h) == 0.0


def _rest_density_closed_form(m=0.125, d=0.05, h=0.1):
    def w(r):
        return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r < h else 0.0
    return m * (w(0) + 6 * w(d) + 12 * w(math.sqrt(2) * d) + 8 * w(math.sqrt(3) * d))


def test_rest_lattice_density_closed_form():
    # Unjittered lattice, no gravity, no planes: the first lambda pass sees
    # the lattice itself; interior particles have the 26 lattice neighbors.
    p = Params(dt=1.0 / 120, density=1000.0, particle_mass=0.125, h=0.1, epsilon=100.0,
               solver_iterations=1, external_forces=(0.0, 0.0, 0.0))
    pos = scenes.lattice_block(9, 9, 9, d=0.05, origin=(1.0, 1.0, 1.0), jitter=0.0)
    r = oracle.step(p, pos, np.zeros_like(pos), nsteps=1, diag=True)
    interior = []
    for z in range(2, 7):
        for y in range(2, 7):
            for x in range(2, 7):
                interior.append((z * 9 + y) * 9 + x)
    rho = r["rho"][interior]
    ref = _rest_density_closed_form()
    assert ref == pytest.approx(1009.8, abs=0.1)
    np.testing.assert_allclose(rho, ref, rtol=1e-5)  # fp32 summation; one W term is >1e-3
    # lambda = -C / (sum |grad C|^2 + eps) with C = rho/rho0 - 1 > 0
    assert np.all(r["lambda"][interior] < 0)


def test_isolated_particle_free_fall_and_lambda():
    p = Params(dt=float(np.float32(1 / 120)), density=1000.0, particle_mass=0.125, h=0.1,
               epsilon=100.0, external_forces=(0.25, -9.8, 0.5))
    x = np.array([[0.3], [0.7], [-0.2]], np.float32)
    v = np.array([[1.0], [-2.0], [0.5]], np.float32)
    r = oracle.step(p, x, v, nsteps=1, diag=True)
    dt = np.float32(p.dt)
    g = np.array(p.external_forces, np.float32)[:, None]
    v1 = v + g * dt
    x1 = x + v1 * dt
    assert same_bits(r["pos"], x1)
    assert same_bits(r["vel"], (x1 - x) / dt)
    w0 = np.float32(oracle.poly6_f32(0.0, 0.1))
    rho = w0 * np.float32(0.125)
    lam = -(rho * (np.float32(1) / np.float32(1000)) - np.float32(1)) / np.float32(100)
    assert same_bits(r["rho"], np.array([rho], np.float32))
    assert same_bits(r["lambda"], np.array([lam], np.float32))


def test_plane_projection_and_restitution_single_particle():
    e, f = 0.5, 0.25
    p = Params(dt=float(np.float32(1 / 120)), density=1000.0, particle_mass=0.125, h=0.1,
               epsilon=100.0, external_forces=(0.0, -9.8, 0.0), plane_restitution=e,
               plane_friction=f)
    p.add_plane_normalized(0, 1, 0, 0)  # floor y = 0
    x = np.array([[0.5], [0.01], [0.5]], np.float32)
    v = np.array([[2.0], [-3.0], [0.0]], np.float32)
    r = oracle.step(p, x, v, nsteps=1, diag=True)
    # pushed exactly onto the plane
    assert r["pos"][1, 0] == 0.0
    dt = np.float32(p.dt)
    vel = (r["pos"][:, 0] - x[:, 0]) / dt
    vn = vel[1]
    assert vn < 0
    # reflected normal component, damped tangential one
    np.testing.assert_allclose(r["vel"][1, 0], -e * vn, rtol=1e-6)
    np.testing.assert_allclose(r["vel"][0, 0], vel[0] * (1 - f), rtol=1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_grid_neighbors_equal_brute_force_sets(seed):
    sc = scenes.dam_break(seed=seed, dims=(7, 6, 5))
    p = sc.params
    a = oracle.step(p, sc.state.pos, sc.state.vel, nsteps=1, diag=True)
    b = oracle.step(p.copy(use_uniform_grid=False), sc.state.pos, sc.state.vel, nsteps=1,
                    diag=True)
    la = neighbor_lists(a["nbr_count"], a["nbr_ids"])
    lb = neighbor_lists(b["nbr_count"], b["nbr_ids"])
    for i, (x, y) in enumerate(zip(la, lb)):
        assert sorted(x.tolist()) == y.tolist(), i
    # the relation is symmetric (r2_ij == r2_ji bitwise) and the total even
    pairs = {(i, int(j)) for i, lst in enumerate(la) for j in lst}
    assert all((j, i) in pairs for (i, j) in pairs)
    assert len(pairs) % 2 == 0


def test_delta_is_antisymmetric_without_planes():
    # Without planes, sum_i delta_i = 0: the pairwise terms are antisymmetric.
    p = Params(dt=float(np.float32(1 / 120)), density=1000.0, particle_mass=0.125, h=0.1,
               epsilon=100.0, solver_iterations=1, external_forces=(0.0, 0.0, 0.0))
    pos = scenes.lattice_block(8, 8, 8, d=0.05, seed=9)
    r = oracle.step(p, pos, np.zeros_like(pos), nsteps=1, dtype=np.float64)
    disp = r["pos"] - pos.astype(np.float64)
    assert np.abs(disp).max() > 1e-6
    np.testing.assert_allclose(disp.sum(axis=1), 0.0, atol=1e-12)


# Each planted mistake must be caught by at least one of the pins above.
MUTATIONS = {1: "drop self density", 2: "spiky sign", 3: "swapped delta operand",
             4: "Gauss-Seidel instead of Jacobi", 5: "r2 <= h2 neighbor test"}


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(mutation):
    caught = []
    for name in ("all_1000_3step", "plain_1000_3step"):
        p, exp = load_golden(name)
        r = oracle.step(p, exp["pos0"], exp["vel0"], nsteps=int(exp["steps"]), diag=True,
                        mutation=mutation)
        caught.append(not same_bits(r["pos"], exp["pos"]))
    if mutation == 5:
        # exactly representable distance h: in the list only with <=
        p = Params(dt=1.0 / 120, density=1000.0, particle_mass=0.125, h=0.1, epsilon=100.0,
                   external_forces=(0.0, 0.0, 0.0))
        h = float(np.float32(0.1))
        x = np.array([[0.0, h], [0.5, 0.5], [0.5, 0.5]], np.float32)
        a = oracle.step(p, x, np.zeros_like(x), diag=True, mutation=5)
        b = oracle.step(p.copy(use_uniform_grid=False), x, np.zeros_like(x), diag=True)
        caught.append(not np.array_equal(a["nbr_count"], b["nbr_count"]))
    assert any(caught), MUTATIONS[mutation]
