"""The hip backend behind the reference's own C++ Params/State: a binary
built from the reference's core.cpp/init.cpp and include/fluid_hip_backend.hpp
steps one scene with fluid::step and with fluid::hip_step / HipSession and
compares them bitwise (oracle/dropin_test.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_test")


@pytest.mark.skipif(not os.path.exists(BIN), reason="built only where the reference sources are")
def test_hip_step_behind_reference_types_is_bitwise():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "hip_step bitwise=1 session(8 steps) bitwise=1" in r.stdout


SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


@pytest.mark.skipif(not os.path.exists(BIN), reason="built only where the reference sources are")
@pytest.mark.parametrize("scene,steps,flags", [("dam_break_32k.json", 6, 0),
                                               ("sphere_cube_slanted.json", 8, 1)])
def test_scene_json_through_the_reference_loader_is_bitwise(scene, steps, flags):
    """init_scene_from_json (core/src/init.cpp:158-418) feeds HipSession, which
    must match fluid::step bitwise after several substeps."""
    r = subprocess.run([BIN, os.path.join(SCENES, scene), str(steps), str(flags)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"session({steps} steps) bitwise=1" in r.stdout
