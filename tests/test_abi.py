"""The C-ABI library loads and exports what include/fluid_hip.h declares;
host-side validation and the no-device path (no kernel launches)."""
import os
import re

import numpy as np
import pytest

from fluidsimulator_amd import hip, scenes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "fluid_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fluid_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = hip.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(hip.exported_symbols())


def test_version_and_device_query():
    assert hip.hip_version() == 2
    ok, count, err = hip.hip_device_available()
    assert ok == (count > 0)
    if not ok:
        assert count == 0


def _params(**kw):
    p = scenes.base_params((1.0, 1.0, 1.0), False)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("kw,code", [
    (dict(use_uniform_grid=False), hip.EUNSUPPORTED),
    (dict(h=0.0), hip.EUNSUPPORTED),
    (dict(dt=0.0), hip.EINVAL),
    (dict(density=-1.0), hip.EINVAL),
    (dict(solver_iterations=-1), hip.EINVAL),
    (dict(epsilon=float("nan")), hip.EINVAL),
])
def test_invalid_params_rejected_before_any_device_work(kw, code):
    with pytest.raises(hip.HipError) as ei:
        hip.HipSolver(_params(**kw), [10])
    assert ei.value.code == code


def test_too_many_planes_rejected():
    p = _params()
    p.planes = p.planes + p.planes
    with pytest.raises(ValueError):
        hip.pack_params(p)


def test_batch_shared_fields_must_agree():
    a = _params()
    b = _params(epsilon=50.0)
    with pytest.raises(hip.HipError) as ei:
        hip.HipSolver([a, b], [5, 5])
    assert ei.value.code == hip.EINVAL


def test_no_device_path():
    ok, _, _ = hip.hip_device_available()
    if ok:
        pytest.skip("a device is present")
    with pytest.raises(hip.HipError) as ei:
        hip.HipSolver(_params(), [10])
    assert ei.value.code == hip.ENODEV
    msg = hip.load_library().fluid_hip_last_error().decode()
    assert "device" in msg


def test_pack_params_rounds_to_float32():
    p = _params(dt=1.0 / 120.0)
    q = hip.pack_params(p)
    assert q.dt == float(np.float32(1.0 / 120.0))
    assert q.n_planes == 6
    assert q.planes[1][0] == -1.0 and q.planes[1][3] == -1.0
