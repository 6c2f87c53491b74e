"""An unrolled PBF substep in PyTorch (float64, CPU) for gradient checks.

Test infrastructure only.  It restates the reference step
(core/src/core.cpp:150-612, vorticity off) with tensor ops so that
torch.autograd can differentiate it:

  predict          core.cpp:150-161   va = v + g dt, p0 = x + va dt
  neighbor lists   core.cpp:205-247   pairs i != j with |p0_i - p0_j|^2 < h^2,
                                      fixed for the whole substep (no gradient
                                      flows through the selection)
  lambda pass      core.cpp:281-329   rho, C, the gradient sums, lambda
  delta pass       core.cpp:334-398   s_corr, plane projection in plane order
  apply            core.cpp:403-407
  commit           core.cpp:414-421
  XSPH             core.cpp:423-466   rho of the last lambda pass
  restitution      core.cpp:573-612

Sums run in a different order than the reference's, so values agree with the
CPU oracle to rounding, not bitwise; branch decisions (pair selection,
r < min_r clamp, plane pushes, contacts) are taken on this float64
trajectory.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _consts(p):
    f32 = lambda v: float(np.float32(v))
    h = f32(p.h)
    pi = math.pi  # the FP64 path's constants, as the CPU oracle's double instantiation
    min_r = f32(0.01) * h
    return dict(
        dt=f32(p.dt), h=h, h2=h * h, min_r2=min_r * min_r,
        c6=315.0 / (64.0 * pi * h ** 9), cs=-45.0 / (pi * h ** 6),
        rho0=f32(p.density), m=f32(p.particle_mass), eps=f32(p.epsilon),
        gs=f32(p.particle_mass) / f32(p.density), iters=int(p.solver_iterations),
        scorr=bool(p.enable_scorr), k=f32(p.scorr_k), n=int(p.scorr_n),
        dq=f32(p.scorr_dq_coeff), xsph=bool(p.enable_xsph) and f32(p.visc_c) != 0.0,
        c=f32(p.visc_c), e=f32(p.plane_restitution), f=f32(p.plane_friction),
        planes=np.asarray(p.planes, np.float32).reshape(-1, 4).astype(np.float64),
        # (restitution, friction) of each plane: per plane when given, else the globals
        props=(np.asarray(p.plane_props, np.float32).reshape(-1, 2).astype(np.float64)
               if getattr(p, "plane_props", None) else
               np.tile([f32(p.plane_restitution), f32(p.plane_friction)],
                       (len(p.planes), 1)).astype(np.float64)),
    )


def _pairs(p0, h2):
    """Brute-force neighbor pairs (i, j), i != j, |p0_i - p0_j|^2 < h^2."""
    with torch.no_grad():
        d = p0[:, None, :] - p0[None, :, :]
        r2 = (d * d).sum(-1)
        n = p0.shape[0]
        m = (r2 < h2) & ~torch.eye(n, dtype=torch.bool)
        i, j = torch.nonzero(m, as_tuple=True)
    return i, j


def substep(x, v, g, p, K=None):
    """One substep.  x, v: (n, 3) float64 tensors, g: (3,) tensor."""
    K = K or _consts(p)
    n = x.shape[0]
    dt, h, h2 = K["dt"], K["h"], K["h2"]
    va = v + g * dt
    pk = x + va * dt
    i, j = _pairs(pk, h2)
    planes = torch.tensor(K["planes"], dtype=x.dtype)
    w0 = K["c6"] * h2 ** 3
    # s_corr W(dq), with the reference's fallback (core.cpp:143-148)
    dq = K["dq"] * h if K["dq"] > 0.0 else 0.3 * h
    wdq = K["c6"] * (h2 - dq * dq) ** 3 if dq * dq <= h2 else 0.0
    inv_wdq = 1.0 / wdq if wdq > 0.0 else 0.0
    rho = None
    for _ in range(K["iters"]):
        d = pk[i] - pk[j]
        r2 = (d * d).sum(1)
        W = torch.where(r2 <= h2, K["c6"] * (h2 - r2) ** 3, torch.zeros_like(r2))
        rho = K["m"] * (w0 + torch.zeros(n, dtype=x.dtype).index_add(0, i, W))
        act = (r2 < h2).detach()
        r = torch.sqrt(torch.clamp(r2, min=K["min_r2"]))
        F = torch.where(act, K["cs"] * (h - r) ** 2, torch.zeros_like(r))
        G = F[:, None] * d
        S = torch.zeros(n, 3, dtype=x.dtype).index_add(0, i, G)
        Q = K["gs"] ** 2 * (torch.zeros(n, dtype=x.dtype).index_add(0, i, (G * G).sum(1))
                            + (S * S).sum(1))
        C = rho / K["rho0"] - 1.0
        lam = -C / (Q + K["eps"])
        s = lam[i] + lam[j]
        if K["scorr"] and inv_wdq > 0.0:
            s = s - K["k"] * (W * inv_wdq) ** K["n"]
        coef = torch.where(act, s * F, torch.zeros_like(F))
        delta = torch.zeros(n, 3, dtype=x.dtype).index_add(0, i, coef[:, None] * d) / K["rho0"]
        q = pk + delta
        for pl in planes:
            nrm, dd = pl[:3], pl[3]
            pen = -(q @ nrm - dd)
            push = (pen > 0).detach()
            q = q + torch.where(push, pen, torch.zeros_like(pen))[:, None] * nrm
        pk = q
    vc = (pk - x) / dt
    x1 = pk
    vx = vc
    if K["xsph"] and rho is not None:
        d = x1[i] - x1[j]
        r2 = (d * d).sum(1)
        act = (r2 < h2).detach()
        W = torch.where(act, K["c6"] * (h2 - r2) ** 3, torch.zeros_like(r2))
        mu = torch.where(rho > 0, K["m"] / rho, torch.zeros_like(rho))
        dv = torch.zeros(n, 3, dtype=x.dtype).index_add(0, i, (W * mu[j])[:, None] * (vc[j] - vc[i]))
        vx = vc + K["c"] * dv
    if bool((K["props"] > 0.0).any()) and len(planes):
        for pl, (e, f) in zip(planes, K["props"]):
            nrm, dd = pl[:3], pl[3]
            contact = ((x1 @ nrm - dd) <= 0).detach()
            vn = vx @ nrm
            vn_new = torch.where(vn < 0, -float(e) * vn, vn)
            vt = vx - vn[:, None] * nrm
            out = vt * (1.0 - float(f)) + vn_new[:, None] * nrm
            vx = torch.where(contact[:, None], out, vx)
    return x1, vx


def rollout(x0, v0, g, p, nsteps):
    K = _consts(p)
    x, v = x0, v0
    for _ in range(nsteps):
        x, v = substep(x, v, g, p, K)
    return x, v


def gradients(p, pos0, vel0, gravity, nsteps, target):
    """Autograd gradients of L = mean_i |x_T,i - target_i|^2 with respect to
    x0, v0 (arrays of shape (3, n)) and g.  Returns (L, dx0, dv0, dg)."""
    x0 = torch.tensor(np.asarray(pos0, np.float64).T.copy(), requires_grad=True)
    v0 = torch.tensor(np.asarray(vel0, np.float64).T.copy(), requires_grad=True)
    g = torch.tensor(np.asarray(gravity, np.float64), requires_grad=True)
    xT, _ = rollout(x0, v0, g, p, nsteps)
    tgt = torch.tensor(np.asarray(target, np.float64).T.copy())
    L = ((xT - tgt) ** 2).sum(1).mean()
    L.backward()
    return (float(L.detach()), x0.grad.numpy().T.copy(), v0.grad.numpy().T.copy(),
            g.grad.numpy().copy())
