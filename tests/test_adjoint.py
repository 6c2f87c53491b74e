"""Hand-written adjoint of the PBF substep (frozen neighbor topology) through
the C ABI: against torch autograd of an unrolled float64 PBF, against central
finite differences of the FP64 CPU oracle, checkpointed against
non-checkpointed (bitwise), batched against per-scene (bitwise), and wrapped
in a torch.autograd.Function.

Finite differences need a map that is smooth over the FD stencil.  Lists
frozen at p0 but evaluated at p^k, plane pushes and the restitution contact
test make the PBF map piecewise smooth, and a block thrown at the walls with
N(0, 0.5^2) velocities crosses those seams: with the reference's default
coefficients (eps 600, k 5e-5) FD with steps 1e-4 and 1e-5 agree to 1e-6
over 10 substeps but not over 30; with the dam-break coefficients (eps 100,
k 0.1) not beyond one substep.  So FD checks run over 1 ("stated"), 10
("reference") and 20 substeps (a smooth mid-air regime: eps 1000, k 1e-4,
N(0, 0.05^2)); longer horizons are checked against the composition of
torch's per-substep vector-Jacobian products along the GPU's own
trajectory, which differentiates the same piecewise map.
"""
import numpy as np
import pytest
import torch

import oracle
from fluidsimulator_amd import hip, scenes
from tests import torch_pbf
from tests.helpers import same_bits_or_nan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip.load_library()
    ok, _, err = hip.hip_device_available()
    assert ok, f"no HIP device: {err}"
    oracle.build(ref=False)


def make_scene(kind, dims, seed=5, scorr=True, xsph=True, restitution=True):
    p = scenes.base_params((4.0, 3.0, 2.0), True)
    p.enable_vorticity = False
    p.enable_scorr = scorr
    p.enable_xsph = xsph
    if not restitution:
        p.plane_restitution = 0.0
        p.plane_friction = 0.0
    if kind == "smooth":
        p.epsilon = 1000.0
        p.scorr_k = 1e-4
        pos = scenes.lattice_block(*dims, d=0.05, origin=(1.6, 1.5, 0.8), seed=seed)
        sigma = 0.05
    else:  # the differentiable-rollout geometry: block in the box corner
        if kind == "reference":  # the reference's default eps and s_corr / XSPH coefficients
            p.epsilon = 600.0
            p.scorr_k = 5e-5
            p.scorr_dq_coeff = 0.3
            p.visc_c = 2e-4
        else:
            assert kind == "stated"
        pos = scenes.lattice_block(*dims, d=0.05, seed=seed)
        sigma = 0.5
    n = pos.shape[1]
    vel = scenes.normal(seed, 3 * n, stream=7, sigma=sigma).reshape(3, n).astype(np.float32)
    vtg = scenes.normal(seed + 1, 3 * n, stream=7, sigma=sigma).reshape(3, n).astype(np.float32)
    return p, pos, vel, vtg


def fp64_target(p, pos, vtg, steps):
    return oracle.step(p, pos.astype(np.float64), vtg.astype(np.float64), nsteps=steps,
                       dtype=np.float64)["pos"]


def hip_gradients(p, pos, vel, steps, target, every=10, tile_gather=0):
    """L = mean_i |x_T,i - target_i|^2 and its gradients through the C ABI."""
    n = pos.shape[1]
    with hip.HipSolver(p, [n], tile_gather=tile_gather) as s:
        s.upload(pos, vel)
        s.forward_record(steps, every)
        out = s.download()
        diff = out.pos.astype(np.float64) - target
        L = float(np.mean(np.sum(diff * diff, axis=0)))
        dx0, dv0, dg = s.backward((2.0 / n) * diff, None)
    return L, dx0, dv0, dg[0]


def rel_inf(a, b):
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("kind,dims,steps,flags", [
    ("smooth", (8, 8, 8), 10, {}),
    ("smooth", (8, 8, 8), 10, {"xsph": False}),
    ("stated", (8, 6, 5), 1, {}),
    ("stated", (8, 6, 5), 1, {"scorr": False}),
    ("stated", (8, 6, 5), 1, {"restitution": False}),
    ("stated", (8, 6, 5), 1, {"xsph": False, "restitution": False}),
    ("reference", (8, 6, 5), 10, {}),
    ("reference", (8, 6, 5), 10, {"restitution": False}),
])
def test_adjoint_matches_torch_autograd(kind, dims, steps, flags):
    p, pos, vel, vtg = make_scene(kind, dims, **flags)
    tgt = fp64_target(p, pos, vtg, steps)
    L_t, dx_t, dv_t, dg_t = torch_pbf.gradients(p, pos, vel, p.external_forces, steps, tgt)
    L_h, dx_h, dv_h, dg_h = hip_gradients(p, pos, vel, steps, tgt)
    assert abs(L_h - L_t) <= 1e-4 * L_t
    # tolerance of the FP32 adjoint against the FP64 autograd gradient
    assert rel_inf(dx_h, dx_t) < 1e-3, rel_inf(dx_h, dx_t)
    assert rel_inf(dv_h, dv_t) < 1e-3, rel_inf(dv_h, dv_t)
    assert rel_inf(dg_h, dg_t) < 1e-3, (dg_h, dg_t)


def test_adjoint_slanted_plane_and_friction():
    p, pos, vel, vtg = make_scene("stated", (6, 6, 6))
    p.add_plane_normalized(1.0, 1.0, 0.0, 0.12)
    tgt = fp64_target(p, pos, vtg, 1)
    _, dx_t, dv_t, dg_t = torch_pbf.gradients(p, pos, vel, p.external_forces, 1, tgt)
    _, dx_h, dv_h, dg_h = hip_gradients(p, pos, vel, 1, tgt)
    assert rel_inf(dx_h, dx_t) < 1e-3
    assert rel_inf(dv_h, dv_t) < 1e-3
    assert rel_inf(dg_h, dg_t) < 1e-3


def _fd_check(p, pos, vel, steps, tgt, ncoords, seed):
    """Central differences of the FP64 oracle on `ncoords` seeded coordinates
    (seven velocity coordinates for each position coordinate) and the three
    gravity components.  Velocities and gravity, the rollout's parameters,
    use step 1e-4; positions use 1e-6 m, about the displacement a 1e-4 m/s
    velocity step makes in one substep (at 1e-4 m the position derivative is
    visibly nonlinear: FD with 1e-4, 1e-5, 1e-6 m converges to the autograd
    value only at the smallest step)."""
    n = pos.shape[1]
    pos64, vel64 = pos.astype(np.float64), vel.astype(np.float64)
    g64 = np.asarray(p.external_forces, np.float32).astype(np.float64)

    def loss(x0, v0, g):
        out = oracle.step(p, x0, v0, nsteps=steps, dtype=np.float64, threads=8,
                          params64={"gravity": g})["pos"]
        return float(np.mean(np.sum((out - tgt) ** 2, axis=0)))

    _, dx_h, dv_h, dg_h = hip_gradients(p, pos, vel, steps, tgt)
    rng = np.random.default_rng(seed)
    hstep, hpos = 1e-4, 1e-6

    def central(which, c, i, h):
        x0, v0, g = pos64.copy(), vel64.copy(), g64.copy()
        x1, v1, g1 = pos64.copy(), vel64.copy(), g64.copy()
        if which == "x":
            x0[c, i] += h
            x1[c, i] -= h
        elif which == "v":
            v0[c, i] += h
            v1[c, i] -= h
        else:
            g[c] += h
            g1[c] -= h
        return (loss(x0, v0, g) - loss(x1, v1, g1)) / (2 * h)

    coords = []
    for t in range(ncoords):
        c, i = int(rng.integers(3)), int(rng.integers(n))
        coords.append(("x", c, i, hpos, dx_h[c, i]) if t % 8 == 7 else ("v", c, i, hstep, dv_h[c, i]))
    coords += [("g", c, 0, hstep, dg_h[c]) for c in range(3)]
    fd, fd10, ad = [], [], []
    for which, c, i, h, a in coords:
        fd.append(central(which, c, i, h))
        fd10.append(central(which, c, i, h / 10))
        ad.append(a)
    fd, fd10, ad = np.array(fd), np.array(fd10), np.array(ad, np.float64)
    # A coordinate whose stencil crosses a seam of the piecewise map (a pair
    # entering a list, a plane push, a contact) has no finite-difference
    # derivative: there FD with h and h/10 disagree by far more than their
    # truncation error (1e-3 relative here).  Those are left out; at least
    # 80% of the coordinates must remain.
    floor = 1e-2 * np.max(np.abs(fd10))
    smooth = np.abs(fd - fd10) <= 1e-3 * np.maximum(np.abs(fd10), floor)
    f, a = fd[smooth], ad[smooth]
    # relative error of the checked gradient vector, and per coordinate
    # floored at 1% of the largest checked entry (entries near zero carry
    # only rounding)
    rel = float(np.linalg.norm(a - f) / np.linalg.norm(f))
    err = np.abs(a - f) / np.maximum(np.abs(f), floor)
    return rel, err, f, a, float(np.mean(smooth))


# Tolerance: the checked gradient vector within 1e-3 relative (2-norm) of the
# FP64 central differences; each coordinate within 1e-2.  The FP32 forward
# itself limits the per-coordinate agreement: the commit velocity
# (p - x) / dt loses about 3e-4 relative per substep in float32.
def test_adjoint_matches_fd_smooth_20_steps():
    p, pos, vel, vtg = make_scene("smooth", (16, 8, 8))
    tgt = fp64_target(p, pos, vtg, 20)
    rel, err, fd, ad, kept = _fd_check(p, pos, vel, 20, tgt, 64, seed=3)
    print(f"smooth 20 steps: rel {rel:.2e}, worst coordinate {np.max(err):.2e}, kept {kept:.2f}")
    assert kept >= 0.8
    assert rel < 1e-3, rel
    assert np.max(err) < 1e-2, (np.max(err), fd[np.argmax(err)], ad[np.argmax(err)])


def test_adjoint_matches_fd_stated_coefficients_one_step():
    p, pos, vel, vtg = make_scene("stated", (16, 16, 16))
    tgt = fp64_target(p, pos, vtg, 1)
    rel, err, fd, ad, kept = _fd_check(p, pos, vel, 1, tgt, 64, seed=4)
    print(f"stated coefficients 1 step: rel {rel:.2e}, worst coordinate {np.max(err):.2e}, "
          f"kept {kept:.2f}")
    assert kept >= 0.8
    assert rel < 1e-3, rel
    assert np.max(err) < 1e-2, (np.max(err), fd[np.argmax(err)], ad[np.argmax(err)])


def test_adjoint_matches_fd_reference_coefficients_10_steps():
    p, pos, vel, vtg = make_scene("reference", (16, 8, 8))
    tgt = fp64_target(p, pos, vtg, 10)
    rel, err, fd, ad, kept = _fd_check(p, pos, vel, 10, tgt, 64, seed=5)
    print(f"reference coefficients 10 steps: rel {rel:.2e}, worst coordinate {np.max(err):.2e}, "
          f"kept {kept:.2f}")
    assert kept >= 0.8
    assert rel < 1e-3, rel
    assert np.max(err) < 1e-2, (np.max(err), fd[np.argmax(err)], ad[np.argmax(err)])


@pytest.mark.parametrize("kind,steps", [("reference", 40), ("stated", 6)])
def test_adjoint_chain_along_gpu_trajectory(kind, steps):
    """Many substeps: the adjoint equals the composition of torch's
    per-substep vector-Jacobian products taken at the states the GPU forward
    itself produced."""
    p, pos, vel, vtg = make_scene(kind, (8, 6, 5))
    n = pos.shape[1]
    states = []
    with hip.HipSolver(p, [n]) as s:
        s.upload(pos, vel)
        for _ in range(steps):
            st = s.download()
            states.append((st.pos.copy(), st.vel.copy()))
            s.step(1)
        fin = s.download()
    tgt = fp64_target(p, pos, vtg, steps)
    diff = fin.pos.astype(np.float64) - tgt
    seed_x = (2.0 / n) * diff
    _, dx_h, dv_h, dg_h = hip_gradients(p, pos, vel, steps, tgt)
    xb = torch.tensor(seed_x.T.copy())
    vb = torch.zeros_like(xb)
    gb = torch.zeros(3, dtype=torch.float64)
    K = torch_pbf._consts(p)
    for sx, sv in reversed(states):
        x = torch.tensor(sx.astype(np.float64).T.copy(), requires_grad=True)
        v = torch.tensor(sv.astype(np.float64).T.copy(), requires_grad=True)
        g = torch.tensor(np.asarray(p.external_forces, np.float32).astype(np.float64),
                         requires_grad=True)
        x1, v1 = torch_pbf.substep(x, v, g, p, K)
        gx, gv, gg = torch.autograd.grad((x1, v1), (x, v, g), (xb, vb))
        xb, vb, gb = gx.detach(), gv.detach(), gb + gg.detach()
    # Tolerance: 1e-3 relative in 2-norm (1e-2 in max-norm) with the
    # reference's coefficients.  With the dam-break coefficients each
    # substep's Jacobian is ill-conditioned (the adjoint grows ~10^2 per
    # substep) and the float32 reverse pass loses ~2e-4 relative per
    # substep: 2e-3 over those 6 substeps.
    tol = 1e-3 if kind == "reference" else 2e-3
    for a, b in ((dx_h, xb.numpy().T), (dv_h, vb.numpy().T), (dg_h, gb.numpy())):
        r2 = float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))
        print(f"chain: rel2 {r2:.2e} relinf {rel_inf(a, b):.2e} max|b| {np.max(np.abs(b)):.3e}")
        assert r2 < tol, r2
        assert rel_inf(a, b) < 1e-2, rel_inf(a, b)


def test_checkpointing_is_bitwise_invisible():
    p, pos, vel, vtg = make_scene("stated", (8, 8, 8))
    steps = 25
    tgt = fp64_target(p, pos, vtg, steps)
    ref = hip_gradients(p, pos, vel, steps, tgt, every=steps)
    for every in (1, 7, 10, 100):
        got = hip_gradients(p, pos, vel, steps, tgt, every=every)
        for a, b in zip(got[1:], ref[1:]):
            assert same_bits_or_nan(np.asarray(a), np.asarray(b)), every


@pytest.mark.parametrize("kind,restitution", [("reference", True), ("stated", True),
                                               ("smooth", False)])
def test_recorded_tiles_give_the_same_gradients(kind, restitution):
    """Recorded substeps with LDS tiles (their lists stored for the reverse
    pass, their iterates, S/Q sums and plane decisions written by the tile
    kernels) give bitwise the gradients of recorded substeps without tiles,
    with and without checkpoints (a non-finite gradient: NaN at the same
    places)."""
    p, pos, vel, vtg = make_scene(kind, (16, 12, 10), restitution=restitution)
    steps = 12
    tgt = fp64_target(p, pos, vtg, steps)
    off = hip_gradients(p, pos, vel, steps, tgt, every=steps, tile_gather=2)
    for every in (steps, 5):
        on = hip_gradients(p, pos, vel, steps, tgt, every=every, tile_gather=1)
        assert on[0] == off[0]
        for a, b in zip(on[1:], off[1:]):
            assert same_bits_or_nan(np.asarray(a), np.asarray(b)), every
    with hip.HipSolver(p, [pos.shape[1]], tile_gather=1) as s:
        s.upload(pos, vel)
        s.step(1)
        assert s.stats()["tiles"] > 0


def test_recorded_tiles_in_a_large_batch_give_the_same_gradients():
    """From 2^18 particles the tile kernels run on the side stream and the
    tile lists (stored as slot words for the reverse pass) are built beside
    the lanes' lists, both appending to one list buffer: the gradients stay
    bitwise those of recorded substeps without tiles."""
    p, pos, vel, _ = make_scene("reference", (80, 58, 57))
    n = pos.shape[1]
    assert n >= 1 << 18
    steps = 2
    tgt = pos.astype(np.float64) + 0.01
    off = hip_gradients(p, pos, vel, steps, tgt, every=steps, tile_gather=2)
    for every in (steps, 1):
        on = hip_gradients(p, pos, vel, steps, tgt, every=every, tile_gather=1)
        assert on[0] == off[0]
        for a, b in zip(on[1:], off[1:]):
            assert same_bits_or_nan(np.asarray(a), np.asarray(b)), every
    with hip.HipSolver(p, [n], tile_gather=1) as s:
        s.upload(pos, vel)
        s.step(1)
        st = s.stats()
    print(f"{n} particles: {st['tiles']} tiles, {st['tiled_particles']} tiled", flush=True)
    assert st["tiled_particles"] > n // 2


def test_backward_repeatable_and_state_restored():
    p, pos, vel, vtg = make_scene("stated", (8, 8, 8))
    n = pos.shape[1]
    with hip.HipSolver(p, [n]) as s:
        s.upload(pos, vel)
        s.forward_record(23, 5)
        fin = s.download()
        seed = np.random.default_rng(0).standard_normal((3, n)).astype(np.float32)
        a = s.backward(seed, seed)
        assert s.stats()["recomputed"] == 20  # four of five segments re-run
        again = s.download()
        b = s.backward(seed, seed)
        for u, w in zip(a, b):
            assert np.array_equal(u.view(np.uint32), w.view(np.uint32))
        assert np.array_equal(fin.pos.view(np.uint32), again.pos.view(np.uint32))
        assert np.array_equal(fin.vel.view(np.uint32), again.vel.view(np.uint32))
        assert again.time == fin.time
        # the recorded forward equals plain stepping, bitwise
        s.upload(pos, vel)
        s.step(23)
        plain = s.download()
        assert np.array_equal(plain.pos.view(np.uint32), fin.pos.view(np.uint32))
        assert np.array_equal(plain.vel.view(np.uint32), fin.vel.view(np.uint32))


def test_batched_adjoint_equals_per_scene():
    ens = scenes.ensemble(seed=31, n_scenes=3, lo=5, hi=9)
    for q in ens.params:
        q.enable_vorticity = False
    counts = [st.n for st in ens.states]
    pos = np.concatenate([st.pos for st in ens.states], axis=1)
    vel = np.concatenate([st.vel for st in ens.states], axis=1)
    n = pos.shape[1]
    steps = 8
    seed_x = np.random.default_rng(1).standard_normal((3, n)).astype(np.float32)
    with hip.HipSolver(ens.params, counts) as b:
        b.upload(pos, vel)
        b.forward_record(steps, 3)
        dx, dv, dg = b.backward(seed_x, None)
    off = np.concatenate([[0], np.cumsum(counts)])
    for k, (q, st) in enumerate(zip(ens.params, ens.states)):
        sl = slice(off[k], off[k + 1])
        with hip.HipSolver(q, [st.n]) as one:
            one.upload(st.pos, st.vel)
            one.forward_record(steps, 3)
            ex, ev, eg = one.backward(seed_x[:, sl], None)
        assert np.array_equal(dx[:, sl].view(np.uint32), ex.view(np.uint32)), k
        assert np.array_equal(dv[:, sl].view(np.uint32), ev.view(np.uint32)), k
        assert np.array_equal(dg[k].view(np.uint32), eg[0].view(np.uint32)), k


def test_record_errors():
    p, pos, vel, _ = make_scene("stated", (4, 4, 4))
    n = pos.shape[1]
    with hip.HipSolver(p, [n]) as s:
        s.upload(pos, vel)
        with pytest.raises(hip.HipError) as e:
            s.backward(np.zeros((3, n), np.float32))
        assert e.value.code == hip.ESTATE
        s.forward_record(3, 2)
        s.step(1)  # any later step discards the record
        with pytest.raises(hip.HipError) as e:
            s.backward(np.zeros((3, n), np.float32))
        assert e.value.code == hip.ESTATE
        with pytest.raises(hip.HipError) as e:
            s.forward_record(3, 0)
        assert e.value.code == hip.EINVAL
    q = scenes.base_params((4.0, 3.0, 2.0), True)  # vorticity on
    with hip.HipSolver(q, [n]) as s:
        s.upload(pos, vel)
        with pytest.raises(hip.HipError) as e:
            s.forward_record(3, 2)
        assert e.value.code == hip.EUNSUPPORTED


class HipRollout(torch.autograd.Function):
    """The adjoint as a torch.autograd.Function (test harness only): forward
    runs the recorded rollout on the GPU, backward the hand-written adjoint."""

    @staticmethod
    def forward(ctx, x0, v0, g, params, steps, every):
        q = params.copy()
        q.external_forces = tuple(float(t) for t in g.detach().numpy())
        s = hip.HipSolver(q, [x0.shape[1]])
        s.upload(x0.detach().numpy().astype(np.float32), v0.detach().numpy().astype(np.float32))
        s.forward_record(steps, every)
        out = s.download()
        ctx.solver = s
        return torch.from_numpy(out.pos.astype(np.float64)), torch.from_numpy(
            out.vel.astype(np.float64))

    @staticmethod
    def backward(ctx, gx, gv):
        s = ctx.solver
        dx, dv, dg = s.backward(gx.numpy(), gv.numpy())
        s.close()
        return (torch.from_numpy(dx.astype(np.float64)), torch.from_numpy(dv.astype(np.float64)),
                torch.from_numpy(dg[0].astype(np.float64)), None, None, None)


def test_autograd_function_wrapper():
    p, pos, vel, vtg = make_scene("smooth", (8, 8, 8))
    steps = 10
    tgt = torch.from_numpy(fp64_target(p, pos, vtg, steps))
    x0 = torch.tensor(pos.astype(np.float64), requires_grad=True)
    v0 = torch.tensor(vel.astype(np.float64), requires_grad=True)
    g = torch.tensor(np.asarray(p.external_forces, np.float32).astype(np.float64),
                     requires_grad=True)
    xT, vT = HipRollout.apply(x0, v0, g, p, steps, 4)
    L = ((xT - tgt) ** 2).sum(0).mean() + 1e-3 * (vT ** 2).sum(0).mean()
    L.backward()
    # the same loss through torch's own unrolled float64 PBF
    x1 = torch.tensor(pos.astype(np.float64).T.copy(), requires_grad=True)
    v1 = torch.tensor(vel.astype(np.float64).T.copy(), requires_grad=True)
    g1 = torch.tensor(np.asarray(p.external_forces, np.float32).astype(np.float64),
                      requires_grad=True)
    xr, vr = torch_pbf.rollout(x1, v1, g1, p, steps)
    Lr = ((xr - tgt.T) ** 2).sum(1).mean() + 1e-3 * (vr ** 2).sum(1).mean()
    Lr.backward()
    assert rel_inf(x0.grad.numpy(), x1.grad.numpy().T) < 1e-3
    assert rel_inf(v0.grad.numpy(), v1.grad.numpy().T) < 1e-3
    assert rel_inf(g.grad.numpy(), g1.grad.numpy()) < 1e-3
