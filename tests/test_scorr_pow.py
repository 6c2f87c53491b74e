"""s_corr with an exponent outside {2, 3, 4} (core.cpp:59-71): the
reference evaluates std::pow(ratio, float(n)), i.e. the host libm's powf;
the device evaluates its own powf.  Neither is correctly rounded in every
case, so this branch is held to a stated tolerance instead of bit equality:

* each powf is within 1 ulp of the correctly rounded power over the ratio's
  whole range [0, 1.5] (host: a sample below; device: every float, by
  tests/native/powf_check), so the two differ by at most 2 ulp per term;
* the n = 5 goldens (written by the reference's own core.cpp) start from the
  violent first substeps of a lattice block, where a 1-ulp change in a few
  percent of the s_corr terms already moves positions by ~1e-5 after one
  substep.  The yardstick is a sibling trajectory of the oracle whose only
  change is x^5 as a product chain (within 3 ulp): after one substep the
  device's positions and velocities must be no further from the reference
  than the sibling's (or within 1e-5 h), with equal neighbor counts; after
  three substeps, where every such trajectory has diverged, the density
  statistics must agree within 1%.

n = 2 (and 3, 4) are products, bitwise like every other golden."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import load_golden

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")


def correctly_rounded_pow(x, n):
    return np.power(x.astype(np.float64), float(n)).astype(np.float32)


def test_host_powf_within_one_ulp_over_the_ratio_range():
    """numpy's float32 power calls the C library's powf, as std::pow(float,
    float) does; every 16th float of [0, 1.5]."""
    bits = np.arange(0, 0x3FC00001, 16, dtype=np.uint32)
    x = bits.view(np.float32)
    p = np.power(x, np.float32(5.0))
    d = correctly_rounded_pow(x, 5)
    ulp = np.abs(p.view(np.int32).astype(np.int64) - d.view(np.int32).astype(np.int64))
    print(f"host powf(x, 5): max {ulp.max()} ulp, {np.count_nonzero(ulp)} of {x.size} "
          "not correctly rounded")
    assert ulp.max() <= 1


@pytest.mark.gpu
def test_device_powf_within_one_ulp_over_the_ratio_range():
    exe = os.path.join(NATIVE, "powf_check")
    assert os.path.exists(exe), "not built (python -m fluidsimulator_amd.build)"
    r = subprocess.run([exe, "5"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def run_hip(name):
    from fluidsimulator_amd import hip

    p, exp = load_golden(name)
    with hip.HipSolver(p, [exp["pos0"].shape[1]]) as s:
        s.upload(exp["pos0"], exp["vel0"], 0.0)
        s.step(int(exp["steps"]))
        out = s.download()
        d = s.diag()
    return p, exp, out, d


def max_dev(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))


@pytest.mark.gpu
def test_scorr_n5_one_substep_within_the_spread_of_a_few_ulp_pow():
    import oracle

    p, exp, out, d = run_hip("powf_scorr_n5_1000_1step")
    old = oracle.set_pow_mode(2)
    try:
        sib = oracle.step(p, exp["pos0"], exp["vel0"], nsteps=1)
    finally:
        oracle.set_pow_mode(old)
    dp, dv = max_dev(out.pos, exp["pos"]), max_dev(out.vel, exp["vel"])
    sp, sv = max_dev(sib["pos"], exp["pos"]), max_dev(sib["vel"], exp["vel"])
    same = float(np.mean(out.pos.view(np.uint32) == exp["pos"].view(np.uint32)))
    print(f"device: max |dpos| {dp:.3g}, |dvel| {dv:.3g} (bitwise-equal positions {same:.3f}); "
          f"product-chain sibling: {sp:.3g}, {sv:.3g}; 1e-5 h = {1e-5 * p.h:.3g}")
    assert dp <= max(sp, 1e-5 * p.h)
    assert dv <= max(sv, 1e-5 * p.h / p.dt)
    assert np.array_equal(d["nbr_count"], exp["nbr_count"])


@pytest.mark.gpu
def test_scorr_n5_three_substeps_density_statistics():
    p, exp, out, d = run_hip("powf_scorr_n5_1000_3step")
    rho0 = p.density

    def stats(rho):
        e = np.abs(rho.astype(np.float64) / rho0 - 1.0)
        return float(np.mean(e)), float(np.max(e))

    (gm, gx), (rm, rx) = stats(d["rho"]), stats(exp["rho"])
    print(f"density error mean/max: device {gm:.5f}/{gx:.5f}, reference {rm:.5f}/{rx:.5f}")
    assert abs(gm - rm) <= 0.01 * rm
