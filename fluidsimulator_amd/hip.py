"""Python side of the `hip` backend: ctypes over the C ABI in include/fluid_hip.h.

Mirrors the reference backend interface (cuda/include/fluid/cuda.h:7-9):
`hip_version()`, `hip_device_available()` and `hip_step(params, state)` with
the meaning of `cuda_version`, `cuda_device_available` and
`cuda_step(const Params&, State&)`.  `HipSolver` is the device-resident form
that the rest of the project uses: upload once, step many times, download.

There is no fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import Params, State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfluid_hip.so")
MAX_PLANES = 8

OK, EINVAL, ENODEV, EHIP, ECAPACITY, ERANGE, EUNSUPPORTED, ESTATE = 0, -1, -2, -3, -4, -5, -6, -7
TUNE_WAVE_MODE, TUNE_BLOCK, TUNE_BRICK, TUNE_XCD, TUNE_TILES = 1, 2, 4, 8, 16


class FluidHipParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("density", C.c_float), ("particle_mass", C.c_float),
        ("h", C.c_float), ("particle_radius", C.c_float), ("epsilon", C.c_float),
        ("solver_iterations", C.c_int32), ("use_uniform_grid", C.c_int32),
        ("enable_scorr", C.c_int32), ("enable_xsph", C.c_int32), ("enable_vorticity", C.c_int32),
        ("scorr_k", C.c_float), ("scorr_n", C.c_int32), ("scorr_dq_coeff", C.c_float),
        ("visc_c", C.c_float), ("plane_restitution", C.c_float), ("plane_friction", C.c_float),
        ("vort_epsilon", C.c_float), ("vort_norm_eps", C.c_float),
        ("gravity", C.c_float * 3), ("n_planes", C.c_int32),
        ("planes", (C.c_float * 4) * MAX_PLANES),
        ("plane_props", C.c_int32), ("plane_restitution_of", C.c_float * MAX_PLANES),
        ("plane_friction_of", C.c_float * MAX_PLANES),
    ]


class FluidHipConfig(C.Structure):
    _fields_ = [
        ("cell_capacity", C.c_int64), ("neighbor_capacity", C.c_int64), ("device", C.c_int32),
        ("block_size", C.c_int32), ("xcd_remap", C.c_int32), ("record_diag", C.c_int32),
        ("wave_mode", C.c_int32), ("brick_edge", C.c_int32), ("lane_capacity", C.c_int64),
        ("tile_gather", C.c_int32),
    ]


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fluid_hip error {code}: {msg}")
        self.code = code


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "fluid_hip_version": (C.c_int, []),
    "fluid_hip_device_available": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_char_p)]),
    "fluid_hip_last_error": (C.c_char_p, []),
    "fluid_hip_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int64),
                                   C.POINTER(FluidHipParams), C.POINTER(FluidHipConfig)]),
    "fluid_hip_destroy": (C.c_int, [C.c_void_p]),
    "fluid_hip_set_params": (C.c_int, [C.c_void_p, C.POINTER(FluidHipParams)]),
    "fluid_hip_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fluid_hip_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]),
    "fluid_hip_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "fluid_hip_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fluid_hip_step_async": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fluid_hip_sync": (C.c_int, [C.c_void_p]),
    "fluid_hip_get_diag": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "fluid_hip_get_neighbors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64,
                                          C.POINTER(C.c_int64)]),
    "fluid_hip_get_grid": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fluid_hip_get_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]),
    "fluid_hip_download_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fluid_hip_download_wait": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "fluid_hip_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "fluid_hip_host_free": (C.c_int, [C.c_void_p]),
    "fluid_hip_profile": (C.c_int, [C.c_void_p, C.c_int32]),
    "fluid_hip_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double),
                                         C.POINTER(C.c_int64)]),
}

# Autotuning and the differentiable rollout.
_OPTIONAL_SIGS: dict = {
    "fluid_hip_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32,
                                     C.POINTER(FluidHipConfig), C.POINTER(C.c_double)]),
    "fluid_hip_forward_record": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fluid_hip_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "fluid_hip_step_noalloc": (C.c_int, [C.POINTER(FluidHipParams), C.c_int64, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_float), C.c_int32,
                                         C.c_void_p]),
}

NOALLOC_MAX_PARTICLES = 32768
NOALLOC_SCRATCH_BYTES = 64 << 20


def exported_symbols():
    """Every C symbol the header declares (checked by the ABI tests)."""
    return list(_SIGS) + list(_OPTIONAL_SIGS)


def load_library(path: str | None = None):
    """Load the HIP library (`path`, else $FLUID_HIP_LIBRARY, else the
    in-tree build); raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("FLUID_HIP_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m fluidsimulator_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in {**_SIGS, **_OPTIONAL_SIGS}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != OK:
        msg = load_library().fluid_hip_last_error()
        raise HipError(rc, msg.decode() if msg else "")
    return rc


def hip_version() -> int:
    """Counterpart of cuda_version() (cuda/include/fluid/cuda.h:7)."""
    return int(load_library().fluid_hip_version())


def hip_device_available():
    """Counterpart of cuda_device_available (cuda/include/fluid/cuda.h:8):
    returns (available, device_count, error_string_or_None)."""
    cnt = C.c_int(0)
    err = C.c_char_p()
    ok = load_library().fluid_hip_device_available(C.byref(cnt), C.byref(err))
    return bool(ok), int(cnt.value), (err.value.decode() if err.value else None)


def pack_params(p: Params) -> FluidHipParams:
    """Params -> the flat C struct (every real rounded to float32)."""
    if len(p.planes) > MAX_PLANES:
        raise ValueError(f"at most {MAX_PLANES} planes")
    q = FluidHipParams()
    for name in ("dt", "density", "particle_mass", "h", "particle_radius", "epsilon", "scorr_k",
                 "scorr_dq_coeff", "visc_c", "plane_restitution", "plane_friction",
                 "vort_epsilon", "vort_norm_eps"):
        setattr(q, name, float(np.float32(getattr(p, name))))
    q.solver_iterations = int(p.solver_iterations)
    q.use_uniform_grid = int(bool(p.use_uniform_grid))
    q.enable_scorr = int(bool(p.enable_scorr))
    q.enable_xsph = int(bool(p.enable_xsph))
    q.enable_vorticity = int(bool(p.enable_vorticity))
    q.scorr_n = int(p.scorr_n)
    for k in range(3):
        q.gravity[k] = float(np.float32(p.external_forces[k]))
    q.n_planes = len(p.planes)
    for i, pl in enumerate(p.planes):
        for k in range(4):
            q.planes[i][k] = float(np.float32(pl[k]))
    props = getattr(p, "plane_props", None)
    if props:
        if len(props) != len(p.planes):
            raise ValueError("plane_props needs one (restitution, friction) pair per plane")
        q.plane_props = 1
        for i, (e, f) in enumerate(props):
            q.plane_restitution_of[i] = float(np.float32(e))
            q.plane_friction_of[i] = float(np.float32(f))
    return q


class HipSolver:
    """A device-resident context for one scene or a segmented batch of scenes.

    params: one Params, or a list with one Params per scene.
    counts: particle count per scene (defaults to [state.n] via upload)."""

    def __init__(self, params, counts, *, cell_capacity=0, neighbor_capacity=0, block_size=256,
                 xcd_remap=False, record_diag=False, device=0, wave_mode=0, brick_edge=0,
                 lane_capacity=0, tile_gather=0):
        self.lib = load_library()
        plist = list(params) if isinstance(params, (list, tuple)) else [params]
        counts = [int(c) for c in counts]
        if len(plist) == 1 and len(counts) > 1:
            plist = plist * len(counts)
        if len(plist) != len(counts):
            raise ValueError("one Params per scene")
        self.params = plist
        self.counts = counts
        self.n = int(sum(counts))
        self.n_scenes = len(counts)
        arr = (FluidHipParams * self.n_scenes)(*[pack_params(p) for p in plist])
        cnt = (C.c_int64 * self.n_scenes)(*counts)
        cfg = FluidHipConfig(int(cell_capacity), int(neighbor_capacity), int(device),
                             int(block_size), int(bool(xcd_remap)), int(bool(record_diag)),
                             int(wave_mode), int(brick_edge), int(lane_capacity),
                             int(tile_gather))
        h = C.c_void_p()
        _check(self.lib.fluid_hip_create(C.byref(h), self.n_scenes, cnt, arr, C.byref(cfg)))
        self.h = h
        self.record_diag = record_diag

    def close(self):
        if getattr(self, "h", None):
            self.lib.fluid_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, params):
        plist = list(params) if isinstance(params, (list, tuple)) else [params] * self.n_scenes
        arr = (FluidHipParams * self.n_scenes)(*[pack_params(p) for p in plist])
        _check(self.lib.fluid_hip_set_params(self.h, arr))
        self.params = plist

    def upload(self, pos, vel, time=0.0):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        if pos.shape != (3, self.n) or vel.shape != (3, self.n):
            raise ValueError(f"pos/vel must be (3, {self.n})")
        _check(self.lib.fluid_hip_upload(self.h, pos.ctypes.data, vel.ctypes.data,
                                         C.c_float(time)))

    def upload_state(self, st: State):
        self.upload(st.pos, st.vel, st.time)

    def download(self) -> State:
        pos = np.empty((3, self.n), np.float32)
        vel = np.empty((3, self.n), np.float32)
        t = C.c_float(0)
        _check(self.lib.fluid_hip_download(self.h, pos.ctypes.data, vel.ctypes.data, C.byref(t)))
        return State(pos, vel, float(t.value))

    def step(self, nsteps=1, graph=False):
        _check(self.lib.fluid_hip_step(self.h, int(nsteps), int(bool(graph))))

    def step_async(self, nsteps=1, graph=False):
        _check(self.lib.fluid_hip_step_async(self.h, int(nsteps), int(bool(graph))))

    def sync(self):
        _check(self.lib.fluid_hip_sync(self.h))

    def diag(self):
        n = self.n
        out = {
            "sorted_ids": np.empty(n, np.int32),
            "nbr_count": np.empty(n, np.int32),
            "rho": np.empty(n, np.float32),
            "lambda": np.empty(n, np.float32),
        }
        cell = np.empty(3 * n, np.int32) if self.record_diag else None
        _check(self.lib.fluid_hip_get_diag(
            self.h, cell.ctypes.data if cell is not None else None,
            out["sorted_ids"].ctypes.data, out["nbr_count"].ctypes.data, out["rho"].ctypes.data,
            out["lambda"].ctypes.data))
        if cell is not None:
            out["cell"] = cell.reshape(n, 3)
        return out

    def neighbors(self):
        total = C.c_int64(0)
        _check(self.lib.fluid_hip_get_neighbors(self.h, None, 0, C.byref(total)))
        ids = np.empty(max(1, total.value), np.int32)
        _check(self.lib.fluid_hip_get_neighbors(self.h, ids.ctypes.data, total.value,
                                                C.byref(total)))
        return ids[: total.value]

    def grid(self):
        keys = np.empty(self.n, np.int32) if self.record_diag else None
        desc = np.empty((self.n_scenes, 8), np.int32)
        _check(self.lib.fluid_hip_get_grid(self.h, keys.ctypes.data if keys is not None else None,
                                           desc.ctypes.data))
        return keys, desc

    def stats(self):
        """Sizes of the last step: neighbor pairs, encoded list words, grid
        cells, list capacity (words), cell capacity, capacity growths; device
        bytes of the rollout recording, substeps the last backward pass
        recomputed, and the bricks the last substep ran as LDS tiles and
        the particles those tiles computed (a tile leaves its brick's dense
        or coincident cells to the per-lane passes)."""
        out = (C.c_int64 * 10)()
        _check(self.lib.fluid_hip_get_stats(self.h, out, 10))
        keys = ("pairs", "words", "cells", "list_capacity", "cell_capacity", "grow_events",
                "record_bytes", "recomputed", "tiles", "tiled_particles")
        return dict(zip(keys, [int(x) for x in out]))

    def autotune(self, trial_steps=2, knobs=0):
        """Time the result-preserving launch variants on the current state
        and keep the fastest (the state is restored).  knobs: mask of
        TUNE_* (0: workgroup size, brick edge and XCD order).  Returns the
        chosen knobs and the timings."""
        cfg = FluidHipConfig()
        rep = (C.c_double * 4)()
        _check(self.lib.fluid_hip_autotune(self.h, int(trial_steps), int(knobs), C.byref(cfg),
                                           rep))
        return {"block_size": cfg.block_size, "xcd_remap": cfg.xcd_remap,
                "wave_mode": cfg.wave_mode, "brick_edge": cfg.brick_edge,
                "tile_gather": cfg.tile_gather,
                "start_ms_per_step": rep[0], "best_ms_per_step": rep[1],
                "candidates_timed": int(rep[2]), "candidates_rejected": int(rep[3])}

    def forward_record(self, nsteps, checkpoint_every=10):
        """Run `nsteps` substeps keeping what the reverse pass needs (state
        every `checkpoint_every` substeps, recordings of the last segment)."""
        _check(self.lib.fluid_hip_forward_record(self.h, int(nsteps), int(checkpoint_every)))

    def backward(self, dpos=None, dvel=None):
        """Reverse pass through the recorded rollout.  dpos, dvel: dL/dx_T,
        dL/dv_T as (3, n) arrays in original order (None = zero).  Returns
        (dL/dx_0, dL/dv_0, dL/dg) with dL/dg of shape (n_scenes, 3)."""
        n = self.n
        def arg(a):
            if a is None:
                return None, None
            a = np.ascontiguousarray(a, dtype=np.float32).reshape(3, n)
            return a, a.ctypes.data_as(C.c_void_p)
        kp, pp = arg(dpos)
        kv, pv = arg(dvel)
        dx0 = np.zeros((3, n), np.float32)
        dv0 = np.zeros((3, n), np.float32)
        dg = np.zeros((self.n_scenes, 3), np.float32)
        _check(self.lib.fluid_hip_backward(self.h, pp, pv, dx0.ctypes.data_as(C.c_void_p),
                                           dv0.ctypes.data_as(C.c_void_p),
                                           dg.ctypes.data_as(C.c_void_p)))
        del kp, kv
        return dx0, dv0, dg

    def profile(self, enable=True):
        _check(self.lib.fluid_hip_profile(self.h, int(bool(enable))))

    def profile_read(self, name):
        ms = C.c_double(0)
        cnt = C.c_int64(0)
        _check(self.lib.fluid_hip_profile_read(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return float(ms.value), int(cnt.value)


def step_noalloc(params: Params, n: int, pos_ptr: int, vel_ptr: int, time: float = 0.0,
                 nsteps: int = 1, stream: int | None = None) -> float:
    """fluid_hip_step_noalloc: `nsteps` substeps of one scene on DEVICE SoA
    arrays (3*n floats each, original order) in place, with every buffer in
    the library's fixed per-GPU scratch array (no device allocation).
    Returns the advanced time."""
    t = C.c_float(time)
    _check(load_library().fluid_hip_step_noalloc(
        C.byref(pack_params(params)), int(n), C.c_void_p(pos_ptr), C.c_void_p(vel_ptr),
        C.byref(t), int(nsteps), C.c_void_p(stream) if stream else None))
    return float(t.value)


def hip_step(params: Params, state: State) -> None:
    """One substep with the call shape of cuda_step(const Params&, State&)
    (cuda/include/fluid/cuda.h:9): stateless, host arrays in and out."""
    with HipSolver(params, [state.n]) as s:
        s.upload(state.pos, state.vel, state.time)
        s.step(1)
        out = s.download()
    state.pos[...] = out.pos
    state.vel[...] = out.vel
    state.time = out.time
