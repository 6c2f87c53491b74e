"""Seeded scene construction for the benchmark workloads.

The reference builds lattices with `init_block` (core/src/init.cpp:72-107,
z-outer / x-inner particle order) and box walls with `init_planes`
(core/src/init.cpp:109-117); it has no random jitter.  The workloads here add
a seeded jitter, so this module defines the random stream itself:

* splitmix64 over a counter (public-domain generator by S. Vigna), one 64-bit
  output per draw;
* uniform u in [0,1) from the top 24 bits of an output;
* jitter = d * 0.1 * (2u - 1) per axis;
* normal draws by Box-Muller in float64, then cast to float32.

Lattice site (ix, iy, iz) of a block sits at origin + (i + 1/2) d; positions
are computed in float64 and rounded once to float32, so every backend gets
identical float32 arrays.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .params import Params, State

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """`count` consecutive splitmix64 outputs for (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x5851F42D4C957F2D) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """float64 uniforms in [0,1) with 24-bit resolution."""
    return (splitmix64(seed, count, stream) >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def normal(seed: int, count: int, stream: int = 0, sigma: float = 1.0) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m, stream)
    u1 = np.maximum(u[0::2], 1.0 / (1 << 25))
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return (sigma * z[:count]).astype(np.float64)


def lattice_block(nx, ny, nz, d, origin=(0.0, 0.0, 0.0), jitter=0.1, seed=0, stream=0):
    """(3, nx*ny*nz) float32 positions, x fastest then y then z."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()]).astype(np.float64)
    pos = np.asarray(origin, np.float64)[:, None] + (idx + 0.5) * d
    n = pos.shape[1]
    if jitter:
        u = uniform01(seed, 3 * n, stream).reshape(3, n)
        pos = pos + d * jitter * (2.0 * u - 1.0)
    return pos.astype(np.float32)


def base_params(box, flags: bool) -> Params:
    """Physics shared by every workload: h 0.1, rho0 1000, m 0.125, eps 100,
    4 iterations, dt 1/120 s, restitution 0.05, friction 0.1; with `flags`,
    s_corr (k 0.1, n 4, dq 0.2h), XSPH (c 0.01), vorticity (eps_v 5e-4)."""
    p = Params(
        dt=float(np.float32(1.0 / 120.0)),
        density=1000.0,
        particle_mass=0.125,
        h=0.1,
        particle_radius=0.025,
        epsilon=100.0,
        solver_iterations=4,
        plane_restitution=0.05,
        plane_friction=0.1,
        external_forces=(0.0, -9.8, 0.0),
    )
    p.set_box(*box)
    if flags:
        p.enable_scorr = True
        p.scorr_k = 0.1
        p.scorr_n = 4
        p.scorr_dq_coeff = 0.2
        p.enable_xsph = True
        p.visc_c = 0.01
        p.enable_vorticity = True
        p.vort_epsilon = 5e-4
    return p


@dataclass
class Scene:
    name: str
    params: Params
    state: State
    steps: int


def dam_break(seed: int = 1, dims=(32, 32, 32), box=(4.0, 3.0, 2.0), flags=False, steps=240) -> Scene:
    """Plain dam break: a jittered lattice block at the origin corner."""
    pos = lattice_block(*dims, d=0.05, seed=seed)
    p = base_params(box, flags)
    return Scene(f"dam_break_{dims[0]}x{dims[1]}x{dims[2]}", p,
                 State(pos, np.zeros_like(pos), 0.0), steps)


def double_dam_break(seed: int = 2, dims=(64, 64, 32), box=(8.0, 4.0, 4.0), steps=480) -> Scene:
    """Two blocks at opposite x ends, all flags."""
    d = 0.05
    a = lattice_block(*dims, d=d, seed=seed, stream=0)
    ox = box[0] - dims[0] * d
    b = lattice_block(*dims, d=d, origin=(ox, 0.0, 0.0), seed=seed, stream=1)
    pos = np.concatenate([a, b], axis=1)
    p = base_params(box, True)
    return Scene("double_dam_break", p, State(pos, np.zeros_like(pos), 0.0), steps)


def million_block(seed: int = 3, dims=(128, 128, 64), box=(12.0, 8.0, 4.0), steps=120) -> Scene:
    """The 1,048,576-particle block, all flags."""
    pos = lattice_block(*dims, d=0.05, seed=seed)
    p = base_params(box, True)
    return Scene("million_block", p, State(pos, np.zeros_like(pos), 0.0), steps)


@dataclass
class Ensemble:
    params: list  # per-scene Params
    states: list  # per-scene State


def ensemble(seed: int = 4, n_scenes: int = 256, lo: int = 8, hi: int = 32, box=(3.0, 3.0, 3.0)) -> Ensemble:
    """Independent dam breaks of uneven size: lattice dims uniform in lo..hi,
    |g| uniform in [8,12] (pointing -y), initial velocity a per-scene constant
    uniform in [-1,1]^3, all flags."""
    u = uniform01(seed, 7 * n_scenes, stream=99).reshape(n_scenes, 7)
    params, states = [], []
    for s in range(n_scenes):
        nx, ny, nz = (lo + np.minimum((u[s, :3] * (hi - lo + 1)).astype(int), hi - lo)).tolist()
        g = 8.0 + 4.0 * u[s, 3]
        v0 = (2.0 * u[s, 4:7] - 1.0).astype(np.float32)
        pos = lattice_block(nx, ny, nz, d=0.05, seed=seed, stream=1000 + s)
        p = base_params(box, True)
        p.external_forces = (0.0, float(np.float32(-g)), 0.0)
        vel = np.repeat(v0[:, None], pos.shape[1], axis=1).astype(np.float32)
        params.append(p)
        states.append(State(pos, vel, 0.0))
    return Ensemble(params, states)


def rollout_scene(seed: int = 5, dims=(64, 32, 32), box=(4.0, 3.0, 2.0), sigma=0.5, steps=100,
                  coefficients: str = "reference"):
    """Differentiable-rollout workload: s_corr and XSPH on, vorticity off,
    per-particle initial velocity N(0, sigma^2).

    coefficients="reference": relaxation eps and the s_corr / XSPH
    coefficients at the reference's Params defaults (eps 600, k 5e-5,
    dq 0.3h, c 2e-4; core/include/fluid/core.h:24-34).  The gradient of a
    100-substep rollout stays bounded (|dL/dv0| ~ 10 at 100 substeps).
    coefficients="stated": eps 100 and the dam-break coefficients (k 0.1,
    dq 0.2h, c 0.01).  Each constraint iteration is then strongly expansive
    and the gradient grows ~10^2 per substep, beyond float32 after about 18
    substeps."""
    pos = lattice_block(*dims, d=0.05, jitter=0.0 if seed is None else 0.1, seed=seed or 0)
    p = base_params(box, True)
    p.enable_vorticity = False
    if coefficients == "reference":
        p.epsilon = 600.0
        p.scorr_k = 5e-5
        p.scorr_dq_coeff = 0.3
        p.visc_c = 2e-4
    elif coefficients != "stated":
        raise ValueError(coefficients)
    n = pos.shape[1]
    vel = normal(seed, 3 * n, stream=7, sigma=sigma).reshape(3, n).astype(np.float32)
    return Scene("rollout", p, State(pos, vel, 0.0), steps)
