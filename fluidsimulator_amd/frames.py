"""State frames (.fsf) of the parity harness, in numpy.

The format (written by the fluidsim_hip app, the parity harness's reference
runner and this module): a 64-byte little-endian header — magic "FLUIDFR1",
uint32 version 1, uint32 flags (bit 0: densities present), uint64 n, int64
step, float32 time, float32 rest density, 24 zero bytes — then positions and
velocities as SoA float32 blocks of 3n (x block, y block, z block, original
particle order) and, when flagged, n float32 densities.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

MAGIC = b"FLUIDFR1"
_HEAD = struct.Struct("<8sIIQqff24x")


@dataclass
class Frame:
    pos: np.ndarray            # (3, n) float32
    vel: np.ndarray            # (3, n) float32
    step: int = 0
    time: float = 0.0
    rho0: float = 0.0
    rho: np.ndarray | None = None  # (n,) float32

    @property
    def n(self) -> int:
        return int(self.pos.shape[1])


def write(path: str, f: Frame) -> None:
    n = f.n
    flags = 0 if f.rho is None else 1
    with open(path, "wb") as fp:
        fp.write(_HEAD.pack(MAGIC, 1, flags, n, int(f.step), float(np.float32(f.time)),
                            float(np.float32(f.rho0))))
        fp.write(np.ascontiguousarray(f.pos, "<f4").reshape(3, n).tobytes())
        fp.write(np.ascontiguousarray(f.vel, "<f4").reshape(3, n).tobytes())
        if f.rho is not None:
            fp.write(np.ascontiguousarray(f.rho, "<f4").reshape(n).tobytes())


def read(path: str) -> Frame:
    with open(path, "rb") as fp:
        head = fp.read(_HEAD.size)
        magic, version, flags, n, step, t, rho0 = _HEAD.unpack(head)
        if magic != MAGIC or version != 1:
            raise ValueError(f"{path} is not a version-1 state frame")
        pos = np.frombuffer(fp.read(12 * n), "<f4").reshape(3, n).copy()
        vel = np.frombuffer(fp.read(12 * n), "<f4").reshape(3, n).copy()
        rho = np.frombuffer(fp.read(4 * n), "<f4").copy() if flags & 1 else None
    return Frame(pos, vel, step, t, rho0, rho)


def read_dir(path: str) -> list:
    """Every .fsf frame of a directory, in file-name order."""
    return [read(os.path.join(path, f)) for f in sorted(os.listdir(path)) if f.endswith(".fsf")]


def density_error(f: Frame) -> float:
    """Mean |rho / rho0 - 1| of a frame with densities."""
    rho0 = f.rho0 if f.rho0 > 0 else 1.0
    return float(np.mean(np.abs(f.rho.astype(np.float64) / rho0 - 1.0)))
