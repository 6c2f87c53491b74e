"""fluid_hip_step_noalloc: the entry point that allocates no device memory
(every buffer in the library's fixed per-GPU scratch array) against the
context path (fluid_hip_create + upload + fluid_hip_step + download), bitwise,
through the C ABI: two calls in a row, the second with a smaller scene, so
the second call runs on a scratch array holding the first call's leftovers."""
import ctypes as C

import numpy as np
import pytest

from fluidsimulator_amd import hip, scenes
from tests.helpers import same_bits


def _context_path(p, pos, vel, steps):
    with hip.HipSolver(p, [pos.shape[1]]) as s:
        s.upload(pos, vel, 0.0)
        s.step(steps)
        return s.download()


def _hip_runtime():
    """The HIP runtime the library itself loaded (same soname, so dlopen
    returns the library already in the process)."""
    hip.load_library()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    rt.hipStreamDestroy.argtypes = [C.c_void_p]
    return rt


H2D, D2H = 1, 2


def _noalloc_path(p, pos, vel, steps, stream=None):
    rt = _hip_runtime()
    n = pos.shape[1]
    host = [np.ascontiguousarray(pos, np.float32).copy(), np.ascontiguousarray(vel, np.float32).copy()]
    dev = [C.c_void_p(), C.c_void_p()]
    nbytes = 12 * n
    try:
        for d, h in zip(dev, host):
            assert rt.hipMalloc(C.byref(d), nbytes) == 0
            assert rt.hipMemcpy(d, h.ctypes.data, nbytes, H2D) == 0
        t = hip.step_noalloc(p, n, dev[0].value, dev[1].value, 0.0, steps, stream)
        for d, h in zip(dev, host):
            assert rt.hipMemcpy(h.ctypes.data, d, nbytes, D2H) == 0
    finally:
        for d in dev:
            if d.value:
                rt.hipFree(d)
    return host[0], host[1], t


@pytest.mark.gpu
def test_noalloc_equals_context_path_on_two_calls_second_smaller():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    rt = _hip_runtime()
    stream = C.c_void_p()
    assert rt.hipStreamCreate(C.byref(stream)) == 0
    try:
        # the second call on a caller stream, with a smaller scene
        for dims, seed, steps, st in (((20, 20, 20), 31, 4, None), ((12, 10, 8), 32, 3, stream)):
            sc = scenes.dam_break(seed=seed, dims=dims, box=(1.6, 1.6, 1.6), flags=True)
            print(f"noalloc vs context: {sc.state.n} particles, {steps} substeps", flush=True)
            ref = _context_path(sc.params, sc.state.pos, sc.state.vel, steps)
            pos, vel, t = _noalloc_path(sc.params, sc.state.pos, sc.state.vel, steps,
                                        st.value if st is not None else None)
            assert same_bits(pos, ref.pos)
            assert same_bits(vel, ref.vel)
            assert np.float32(t) == np.float32(ref.time)
    finally:
        rt.hipStreamDestroy(stream)


def test_noalloc_rejects_scenes_beyond_its_scratch_array():
    """Checked on the host before any device work (runs without a GPU)."""
    p = scenes.base_params((1.0, 1.0, 1.0), False)
    with pytest.raises(hip.HipError) as ei:
        hip.step_noalloc(p, hip.NOALLOC_MAX_PARTICLES + 1, 0, 0, 0.0, 1)
    assert ei.value.code == hip.EINVAL
    assert str(hip.NOALLOC_MAX_PARTICLES) in str(ei.value)
    # zero particles: only the time advances (core.cpp:122-125)
    t = hip.step_noalloc(p, 0, 0, 0, 1.0, 3)
    dt = np.float32(p.dt)
    assert np.float32(t) == np.float32(np.float32(np.float32(np.float32(1.0) + dt) + dt) + dt)
