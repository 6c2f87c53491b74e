"""The unrolled float64 PyTorch PBF (tests/torch_pbf.py) that the adjoint is
checked against, pinned on the CPU: its forward against the FP64 CPU
oracle, its autograd gradient against central finite differences of that
oracle."""
import numpy as np
import pytest
import torch

import oracle
from tests import torch_pbf
from tests.test_adjoint import make_scene, fp64_target


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build(ref=False)


@pytest.mark.parametrize("kind,dims,steps,tol", [
    ("smooth", (8, 8, 8), 10, 1e-12),
    ("reference", (8, 6, 5), 5, 1e-10),
    ("stated", (8, 6, 5), 1, 1e-11),
])
def test_forward_matches_fp64_oracle(kind, dims, steps, tol):
    p, pos, vel, _ = make_scene(kind, dims)
    o = oracle.step(p, pos.astype(np.float64), vel.astype(np.float64), nsteps=steps,
                    dtype=np.float64)
    x0 = torch.tensor(pos.astype(np.float64).T.copy())
    v0 = torch.tensor(vel.astype(np.float64).T.copy())
    g = torch.tensor(np.asarray(p.external_forces, np.float32).astype(np.float64))
    xT, vT = torch_pbf.rollout(x0, v0, g, p, steps)
    # summation order differs from the oracle's; the rest is the same map
    assert np.max(np.abs(xT.numpy().T - o["pos"])) < tol
    assert np.max(np.abs(vT.numpy().T - o["vel"])) < tol / p.dt


def test_forward_drop_in_mistake_is_caught():
    """A planted mistake in the oracle (the spiky gradient's sign flipped)
    separates the two, so the comparison above has teeth."""
    p, pos, vel, _ = make_scene("stated", (6, 5, 4))
    bad = oracle.step(p, pos.astype(np.float64), vel.astype(np.float64), nsteps=1,
                      dtype=np.float64, mutation=2)
    x0 = torch.tensor(pos.astype(np.float64).T.copy())
    v0 = torch.tensor(vel.astype(np.float64).T.copy())
    g = torch.tensor(np.asarray(p.external_forces, np.float32).astype(np.float64))
    xT, _ = torch_pbf.rollout(x0, v0, g, p, 1)
    assert np.max(np.abs(xT.numpy().T - bad["pos"])) > 1e-6


@pytest.mark.parametrize("kind,dims,steps", [("reference", (8, 6, 5), 8), ("stated", (8, 6, 5), 1)])
def test_autograd_matches_fd(kind, dims, steps):
    p, pos, vel, vtg = make_scene(kind, dims)
    tgt = fp64_target(p, pos, vtg, steps)
    _, dx, dv, dg = torch_pbf.gradients(p, pos, vel, p.external_forces, steps, tgt)
    pos64, vel64 = pos.astype(np.float64), vel.astype(np.float64)
    g64 = np.asarray(p.external_forces, np.float32).astype(np.float64)

    def loss(v0, g):
        out = oracle.step(p, pos64, v0, nsteps=steps, dtype=np.float64,
                          params64={"gravity": g})["pos"]
        return float(np.mean(np.sum((out - tgt) ** 2, axis=0)))

    rng = np.random.default_rng(2)
    n = pos.shape[1]
    fd, ad = [], []
    for _ in range(16):
        c, i = int(rng.integers(3)), int(rng.integers(n))
        a, b = vel64.copy(), vel64.copy()
        a[c, i] += 1e-5
        b[c, i] -= 1e-5
        fd.append((loss(a, g64) - loss(b, g64)) / 2e-5)
        ad.append(dv[c, i])
    for c in range(3):
        a, b = g64.copy(), g64.copy()
        a[c] += 1e-5
        b[c] -= 1e-5
        fd.append((loss(vel64, a) - loss(vel64, b)) / 2e-5)
        ad.append(dg[c])
    fd, ad = np.array(fd), np.array(ad)
    # within 1e-3 relative (2-norm); a stencil that brushes a seam of the
    # piecewise map can still carry ~5e-4 on its own coordinate
    assert np.linalg.norm(ad - fd) / np.linalg.norm(fd) < 1e-3
