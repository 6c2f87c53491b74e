"""TEST INFRASTRUCTURE ONLY — ctypes access to the CPU checkers.

* `liboracle.so` (pbf_oracle.cpp): this project's restatement of the
  reference CPU step (core/src/core.cpp:119-615), float and double.
* `_ref/libfluid_ref.so` (ref_shim.cpp + the reference's own core.cpp and
  init.cpp): the reference step itself, float only; present where the
  reference sources were available at build time.

Only tests/, `__graft_entry__.smoke()` and bench.py's CPU-baseline leg use
this module.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfluid_ref.so")
REF_SRC = "/root/reference"


class OracleParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("density", C.c_double), ("particle_mass", C.c_double),
        ("h", C.c_double), ("epsilon", C.c_double), ("solver_iterations", C.c_int),
        ("use_uniform_grid", C.c_int), ("enable_scorr", C.c_int), ("enable_xsph", C.c_int),
        ("enable_vorticity", C.c_int), ("scorr_n", C.c_int), ("scorr_k", C.c_double),
        ("scorr_dq_coeff", C.c_double), ("visc_c", C.c_double),
        ("plane_restitution", C.c_double), ("plane_friction", C.c_double),
        ("vort_epsilon", C.c_double), ("vort_norm_eps", C.c_double),
        ("gravity", C.c_double * 3), ("n_planes", C.c_int), ("mutation", C.c_int),
        ("planes", C.POINTER(C.c_double)), ("plane_props", C.POINTER(C.c_double)),
    ]


class OracleDiag(C.Structure):
    _fields_ = [
        ("cell", C.POINTER(C.c_int)), ("sorted_ids", C.POINTER(C.c_int)),
        ("nbr_count", C.POINTER(C.c_int)), ("nbr_ids", C.POINTER(C.c_int)),
        ("nbr_cap", C.c_longlong), ("nbr_total", C.c_longlong),
        ("rho", C.c_void_p), ("lambda_", C.c_void_p), ("pred0", C.c_void_p),
    ]


def build(force: bool = False, ref: bool = True) -> None:
    """Compile the checkers with oracle/Makefile (the reference library only
    when its sources are present)."""
    targets = ["all"]
    if ref and os.path.isdir(REF_SRC):
        targets.append("ref")
    cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else []) + targets
    subprocess.run(cmd, check=True)


_libs: dict = {}


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle.build())")
        lib = C.CDLL(path)
        _libs[path] = lib
    return _libs[path]


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def _pack_params(p, mutation=0):
    """Params (fluidsimulator_amd.params.Params) -> OracleParams.  Every real
    is first rounded to float32, as the C++ struct holds it."""
    f = lambda name: float(np.float32(getattr(p, name)))
    planes = np.ascontiguousarray(np.asarray(p.planes, np.float32).reshape(-1, 4).astype(np.float64))
    op = OracleParams()
    op.dt, op.density, op.particle_mass = f("dt"), f("density"), f("particle_mass")
    op.h, op.epsilon = f("h"), f("epsilon")
    op.solver_iterations = int(p.solver_iterations)
    op.use_uniform_grid = int(bool(p.use_uniform_grid))
    op.enable_scorr, op.enable_xsph = int(bool(p.enable_scorr)), int(bool(p.enable_xsph))
    op.enable_vorticity = int(bool(p.enable_vorticity))
    op.scorr_n = int(p.scorr_n)
    op.scorr_k, op.scorr_dq_coeff, op.visc_c = f("scorr_k"), f("scorr_dq_coeff"), f("visc_c")
    op.plane_restitution, op.plane_friction = f("plane_restitution"), f("plane_friction")
    op.vort_epsilon, op.vort_norm_eps = f("vort_epsilon"), f("vort_norm_eps")
    g = [float(np.float32(v)) for v in p.external_forces]
    op.gravity[0], op.gravity[1], op.gravity[2] = g
    op.n_planes = planes.shape[0]
    op.mutation = int(mutation)
    op.planes = planes.ctypes.data_as(C.POINTER(C.c_double)) if planes.size else None
    props = getattr(p, "plane_props", None)
    keep = [planes]
    if props:
        pp = np.ascontiguousarray(np.asarray(props, np.float32).reshape(-1, 2).astype(np.float64))
        if pp.shape[0] != planes.shape[0]:
            raise ValueError("plane_props needs one (restitution, friction) pair per plane")
        op.plane_props = pp.ctypes.data_as(C.POINTER(C.c_double))
        keep.append(pp)
    return op, keep


def step(params, pos, vel, nsteps=1, time=0.0, diag=False, which="oracle",
         dtype=np.float32, mutation=0, threads=None, params64=None, nbr_cap=None):
    """Run `nsteps` CPU steps.  pos, vel: (3, n) arrays.  Returns a dict with
    pos, vel, time and, with diag=True, the last step's grid/neighbor/rho data
    (original particle order); diag="light" returns only the last step's
    neighbor counts, pair total, rho and lambda.

    which: "oracle" (this project's restatement) or "ref" (the reference
    core.cpp).  dtype float64 runs the FP64 oracle; `params64` then supplies
    double-precision overrides (e.g. gravity for finite differences).
    nbr_cap sizes the neighbor-list buffer of the diagnostics (default 64
    per particle; grown automatically when the lists are longer)."""
    n = int(pos.shape[1])
    pos = np.ascontiguousarray(pos, dtype=dtype).copy()
    vel = np.ascontiguousarray(vel, dtype=dtype).copy()
    op, keep = _pack_params(params, mutation)
    if params64:
        for k, v in params64.items():
            if k == "gravity":
                op.gravity[0], op.gravity[1], op.gravity[2] = [float(x) for x in v]
            else:
                setattr(op, k, float(v))
    t = np.array([time], dtype=dtype)
    out = {}
    d = None
    light = diag == "light"
    if light:
        # per-particle sums only (neighbor counts, rho, lambda): no lists, so
        # piled-up states with 1e9 pairs stay cheap
        out["nbr_count"] = np.zeros(n, np.int32)
        out["rho"] = np.zeros(n, dtype)
        out["lambda"] = np.zeros(n, dtype)
        d = OracleDiag(None, None, out["nbr_count"].ctypes.data_as(C.POINTER(C.c_int)), None, 0,
                       0, out["rho"].ctypes.data, out["lambda"].ctypes.data, None)
    elif diag:
        cap = int(nbr_cap) if nbr_cap else max(1, n * 64)
        out["cell"] = np.zeros(3 * n, np.int32)
        out["sorted_ids"] = np.zeros(n, np.int32)
        out["nbr_count"] = np.zeros(n, np.int32)
        out["nbr_ids"] = np.zeros(cap, np.int32)
        out["rho"] = np.zeros(n, dtype)
        out["lambda"] = np.zeros(n, dtype)
        out["pred0"] = np.zeros(3 * n, dtype)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        d = OracleDiag(ip(out["cell"]), ip(out["sorted_ids"]), ip(out["nbr_count"]),
                       ip(out["nbr_ids"]), cap, 0, out["rho"].ctypes.data,
                       out["lambda"].ctypes.data,
                       out["pred0"].ctypes.data if which == "oracle" else None)
    if which == "oracle":
        lib = _load(ORACLE_SO)
        fn = lib.oracle_step_f32 if dtype == np.float32 else lib.oracle_step_f64
        if threads:
            lib.oracle_set_threads(int(threads))
    elif which == "ref":
        if dtype != np.float32:
            raise ValueError("the reference step is float32 only")
        lib = _load(REF_SO)
        fn = lib.ref_step_f32
        if threads:
            lib.ref_set_threads(int(threads))
    else:
        raise ValueError(which)
    fn.restype = C.c_int
    pos_in, vel_in = pos.copy(), vel.copy()
    rc = fn(C.byref(op), C.c_longlong(n), pos.ctypes.data_as(C.c_void_p),
            vel.ctypes.data_as(C.c_void_p), C.c_int(nsteps), t.ctypes.data_as(C.c_void_p),
            C.byref(d) if d is not None else None)
    if rc != 0:
        raise RuntimeError(f"{which} step failed rc={rc}")
    if d is not None and not light and d.nbr_total > d.nbr_cap:
        # dense scenes: rerun with a list buffer of the exact size
        return step(params, pos_in, vel_in, nsteps=nsteps, time=time, diag=True, which=which,
                    dtype=dtype, mutation=mutation, threads=threads, params64=params64,
                    nbr_cap=int(d.nbr_total))
    out["pos"], out["vel"], out["time"] = pos, vel, t[0]
    if light:
        out["nbr_total"] = int(d.nbr_total)
    elif diag:
        out["cell"] = out["cell"].reshape(n, 3)
        out["nbr_ids"] = out["nbr_ids"][: d.nbr_total]
        out["pred0"] = out["pred0"].reshape(3, n)
        if which == "ref":
            out.pop("pred0")
    del keep
    return out


def set_pow_mode(mode: int) -> int:
    """s_corr's power for n outside {2, 3, 4}: 0 std::pow as the reference
    (core.cpp:59-71), 1 the correctly rounded power (a sibling trajectory for
    measuring how far a 1-ulp pow difference carries; not the reference).
    Returns the previous mode."""
    return int(_load(ORACLE_SO).oracle_set_pow_mode(int(mode)))


def set_threads(t: int, which="oracle") -> int:
    lib = _load(ORACLE_SO if which == "oracle" else REF_SO)
    f = lib.oracle_set_threads if which == "oracle" else lib.ref_set_threads
    return int(f(int(t)))


def poly6_f32(r2, h):
    lib = _load(ORACLE_SO)
    lib.oracle_poly6_f32.restype = C.c_double
    return lib.oracle_poly6_f32(C.c_double(r2), C.c_double(h))


def spiky_f32(r, h):
    lib = _load(ORACLE_SO)
    lib.oracle_spiky_f32.restype = C.c_double
    return lib.oracle_spiky_f32(C.c_double(r), C.c_double(h))
