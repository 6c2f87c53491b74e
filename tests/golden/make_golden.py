"""Writes the golden vectors under tests/golden/ from the reference CPU step.

The reference `fluid::step` (core/src/core.cpp:119-615) is compiled from its
own sources by oracle/Makefile into oracle/_ref/libfluid_ref.so; this script
runs it on small seeded scenes and stores inputs and outputs as .npz
fixtures (no pickles): initial pos/vel, the parameters, and after `steps`
steps the positions, velocities, time, and the last step's cells, sorted
order, neighbor counts and lists, rho and lambda.

    python tests/golden/make_golden.py [NAME ...]   (needs oracle/_ref built)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fluidsimulator_amd import scenes  # noqa: E402
from fluidsimulator_amd.params import Params  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def params_json(p: Params) -> str:
    d = {k: getattr(p, k) for k in p.__dataclass_fields__}
    d["external_forces"] = list(d["external_forces"])
    d["planes"] = [list(x) for x in d["planes"]]
    return json.dumps(d)


def cases():
    """(name, params, pos, vel, steps)"""
    out = []
    base = scenes.base_params((1.0, 1.0, 1.0), False)
    small = scenes.lattice_block(6, 5, 4, d=0.05, origin=(0.1, 0.05, 0.1), seed=11)
    mid = scenes.lattice_block(10, 10, 10, d=0.05, seed=12)
    zeros_s = np.zeros_like(small)
    zeros_m = np.zeros_like(mid)
    out.append(("plain_120_1step", base, small, zeros_s, 1))
    out.append(("plain_1000_3step", base, mid, zeros_m, 3))
    p = base.copy(enable_scorr=True, scorr_k=0.1, scorr_n=4, scorr_dq_coeff=0.2)
    out.append(("scorr_1000_2step", p, mid, zeros_m, 2))
    p = base.copy(enable_scorr=True, scorr_k=0.1, scorr_n=3, scorr_dq_coeff=0.0)
    out.append(("scorr_n3_dqdefault_1000_1step", p, mid, zeros_m, 1))
    p = base.copy(enable_scorr=True, scorr_k=0.1, scorr_n=2, scorr_dq_coeff=0.2)
    out.append(("scorr_n2_1000_2step", p, mid, zeros_m, 2))
    # n outside {2, 3, 4}: the reference's std::pow branch (core.cpp:59-71);
    # the device's powf may differ from the host libm's in the last bit, so
    # these are compared to a stated tolerance, not bitwise ("powf_" prefix)
    p = base.copy(enable_scorr=True, scorr_k=0.1, scorr_n=5, scorr_dq_coeff=0.2)
    out.append(("powf_scorr_n5_1000_1step", p, mid, zeros_m, 1))
    p = base.copy(enable_scorr=True, scorr_k=0.1, scorr_n=5, scorr_dq_coeff=0.2)
    out.append(("powf_scorr_n5_1000_3step", p, mid, zeros_m, 3))
    p = base.copy(enable_xsph=True, visc_c=0.01)
    out.append(("xsph_1000_2step", p, mid, zeros_m, 2))
    p = base.copy(enable_vorticity=True, vort_epsilon=5e-4)
    vel = scenes.normal(21, 3 * mid.shape[1], sigma=0.5).reshape(3, -1).astype(np.float32)
    out.append(("vort_1000_2step", p, mid, vel, 2))
    allf = scenes.base_params((1.0, 1.0, 1.0), True)
    out.append(("all_1000_3step", allf, mid, vel, 3))
    p = allf.copy(plane_restitution=0.0, plane_friction=0.0)
    out.append(("all_norest_1000_2step", p, mid, vel, 2))
    p = allf.copy(solver_iterations=1)
    out.append(("all_iter1_1000_2step", p, mid, vel, 2))
    p = base.copy(solver_iterations=0)
    out.append(("iter0_120_1step", p, small, zeros_s, 1))
    # an isolated particle and a pair, with restitution
    one = np.array([[0.5], [0.02], [0.5]], np.float32)
    out.append(("single_1step", base, one, np.array([[0.3], [-2.0], [0.0]], np.float32), 1))
    pair = np.array([[0.5, 0.55], [0.5, 0.52], [0.5, 0.49]], np.float32)
    out.append(("pair_2step", base, pair, np.zeros_like(pair), 2))
    # a slanted plane (normalized) and no gravity in x
    p = base.copy(external_forces=(1.0, -9.8, 0.5))
    p.planes = list(base.planes)
    p.add_plane_normalized(1.0, 1.0, 0.0, 0.3)
    out.append(("slanted_1000_2step", p, mid, zeros_m, 2))
    # the reference's O(N^2) neighbor path
    p = base.copy(use_uniform_grid=False)
    out.append(("bruteforce_120_1step", p, small, zeros_s, 1))
    return out


def main(only=()):
    """Writes every case, or only the cases named on the command line."""
    if not oracle.have_ref():
        raise SystemExit("oracle/_ref/libfluid_ref.so is not built (make -C oracle ref)")
    for name, p, pos, vel, steps in cases():
        if only and name not in only:
            continue
        r = oracle.step(p, pos, vel, nsteps=steps, diag=True, which="ref")
        np.savez_compressed(
            os.path.join(OUT, f"{name}.npz"),
            params=np.array(params_json(p)), steps=np.int32(steps),
            pos0=pos.astype(np.float32), vel0=vel.astype(np.float32),
            pos=r["pos"], vel=r["vel"], time=np.float32(r["time"]),
            cell=r["cell"], sorted_ids=r["sorted_ids"], nbr_count=r["nbr_count"],
            nbr_ids=r["nbr_ids"], rho=r["rho"], lam=r["lambda"])
        print(name, pos.shape[1], "particles", steps, "steps")


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
