"""The gather passes' square root (sqrt_rn, pbf_device.h): a fast form of
the correctly rounded sqrtf for finite x >= 2^-96.  Checked on the device
against sqrtf over every non-negative finite float (tests/native/sqrt_check,
built by fluidsimulator_amd.build)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "sqrt_check")


def test_sqrt_rn_equals_sqrtf_on_every_float():
    assert os.path.exists(EXE), "not built (python -m fluidsimulator_amd.build)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ": 0 differ from sqrtf" in r.stdout
