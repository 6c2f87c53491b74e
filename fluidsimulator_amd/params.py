"""Solver parameters and particle state on the host.

`Params` mirrors the reference solver contract `fluid::Params`
(core/include/fluid/core.h:11-79): the same field names, meanings and
defaults, plus the plane list as (nx, ny, nz, d) rows (`PlaneSoA`,
core.h:44-78).  `State` mirrors `fluid::State` (core.h:121-132): SoA positions
and velocities in float32 and the float `time`.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Params:
    dt: float = 1.0 / 60.0
    backend: str = "cpu"  # "cpu" | "cuda" | "hip"; set by the caller, not read by the step
    density: float = 6000.0
    particle_mass: float = 0.0
    h: float = 0.0
    particle_radius: float = 0.01
    epsilon: float = 600.0
    solver_iterations: int = 4
    neighbor_reserve_factor: float = 1.5
    use_uniform_grid: bool = True
    enable_scorr: bool = False
    enable_xsph: bool = False
    enable_vorticity: bool = False
    scorr_k: float = 0.00005
    scorr_n: int = 4
    scorr_dq_coeff: float = 0.3
    visc_c: float = 0.0002
    plane_restitution: float = 0.0
    plane_friction: float = 0.0
    vort_epsilon: float = 0.5
    vort_norm_eps: float = 1e-6
    external_forces: tuple = (0.0, -9.8, 0.0)
    planes: list = field(default_factory=list)  # [(nx, ny, nz, d), ...]
    # Extension beyond fluid::Params: one (restitution, friction) pair per
    # plane, or None for plane_restitution / plane_friction on every plane.
    plane_props: list | None = None

    def f32(self, name: str) -> float:
        """The field as the float32 value the C++ struct would hold."""
        return float(np.float32(getattr(self, name)))

    def copy(self, **kw) -> "Params":
        p = dataclasses.replace(self, **kw)
        if "planes" not in kw:
            p.planes = list(self.planes)
        if "plane_props" not in kw and self.plane_props is not None:
            p.plane_props = list(self.plane_props)
        return p

    def add_plane_normalized(self, nx, ny, nz, d):
        """PlaneSoA::add_normalized (core.h:69-77) in float32 arithmetic."""
        nx, ny, nz, d = (np.float32(v) for v in (nx, ny, nz, d))
        len_sq = np.float32(np.float32(nx * nx) + np.float32(ny * ny)) + np.float32(nz * nz)
        if len_sq > 0:
            inv = np.float32(np.float32(1.0) / np.float32(np.sqrt(len_sq)))
            self.planes.append((float(nx * inv), float(ny * inv), float(nz * inv), float(d)))
        else:
            self.planes.append((float(nx), float(ny), float(nz), float(d)))

    def set_box(self, bx, by, bz):
        """Six axis-aligned walls of [0,bx]x[0,by]x[0,bz] (init_planes,
        core/src/init.cpp:109-117)."""
        self.planes = []
        self.add_plane_normalized(1, 0, 0, 0)
        self.add_plane_normalized(-1, 0, 0, -bx)
        self.add_plane_normalized(0, 1, 0, 0)
        self.add_plane_normalized(0, -1, 0, -by)
        self.add_plane_normalized(0, 0, 1, 0)
        self.add_plane_normalized(0, 0, -1, -bz)


@dataclass
class State:
    pos: np.ndarray  # (3, n) float32, rows x, y, z
    vel: np.ndarray  # (3, n) float32
    time: float = 0.0

    @property
    def n(self) -> int:
        return int(self.pos.shape[1])

    def copy(self) -> "State":
        return State(self.pos.copy(), self.vel.copy(), self.time)

    @staticmethod
    def empty(n: int) -> "State":
        return State(np.zeros((3, n), np.float32), np.zeros((3, n), np.float32), 0.0)
