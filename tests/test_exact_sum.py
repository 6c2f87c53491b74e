"""The exact shortcut for runs of identical float additions used by the
gather passes (fluidsimulator_amd/csrc/exact_sum.h): built for the host with
hipcc and checked against plain additions, bit for bit, on random and
constructed cases (ties, both signs, zero crossings, subnormals, overflow)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc is needed to build the host check")
def test_add_repeat_equals_plain_additions(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "test_add_repeat"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-o", str(exe),
                    os.path.join(HERE, "csrc", "test_add_repeat.cpp")], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
