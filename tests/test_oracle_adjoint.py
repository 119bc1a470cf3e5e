"""Pinning the oracle's discrete Taylor adjoint (oracle.TaylorAdjoint) on the CPU.

The adjoint restates no reference routine line by line (the reference's optimal-control
gradient, src/opt_control/propagation.F90:1016-1103, is a continuous-time formula), so it
is pinned by what fixes a gradient:

* central finite differences (h = 1e-6) of J through the oracle's own forward
  propagation, for every eps_s, for dJ/dv along a smooth direction and at single points,
  and for dJ/dpsi_0 (the returned chi_0) along a random direction;
* the reference's formula 2 dt Im <chi_{s+1}|x pol|psi_s> as the first-order-in-dt limit:
  its relative distance to the exact gradient must halve when dt halves;
* planted mistakes (a wrong sign of the field term, a dropped second-order term) that
  the finite differences must catch.
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_sphere
from octhip import workloads


def tiny(natoms=2):
    return small_sphere(radius=1.8, spacing=0.3, natoms=natoms, nwells=3, rc=0.9)


def setup(nst=4, nsteps=6, dt=0.05, seed_first=0):
    s = tiny()
    vol = s.mesh.vol_pp
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=1.5, first=seed_first),
                                 vol)
    tg = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=1.0, first=100 + seed_first),
                                  vol)
    psi0 = workloads.pack(z, s.np_part)
    phi = workloads.pack(tg, s.np_part)
    xyz = s.mesh.x[: s.np]
    t = (np.arange(nsteps) + 0.5) * dt
    eps = 0.3 * np.sin(0.9 * t + 0.2) + 0.1
    return s, vol, psi0, phi, xyz, eps


def make(s, xyz, dt, vol, v=None, order=4):
    h = O.Hamiltonian(s.op, s.v.copy() if v is None else v.copy(), s.proj, xyz=xyz, field=[0.0, 0.0, 0.0])
    return O.TaylorAdjoint(h, xyz, [1.0, 0.0, 0.0], dt, order=order, vol=vol)


def objective(adj, psi0, eps, phi, vol, rows):
    psiN = adj.forward(psi0, eps)
    return O.overlap_objective(phi, psiN, rows, vol)


def gradients(s, vol, psi0, phi, xyz, eps, dt):
    adj = make(s, xyz, dt, vol)
    J, chiN = objective(adj, psi0, eps, phi, vol, s.np)
    chi0, deps, dv = adj.backward(chiN)
    return J, chi0, deps, dv


def test_deps_matches_central_differences():
    s, vol, psi0, phi, xyz, eps = setup()
    dt = 0.05
    J, chi0, deps, dv = gradients(s, vol, psi0, phi, xyz, eps, dt)
    h = 1e-6
    fd = np.zeros_like(deps)
    for i in range(len(eps)):
        ep, em = eps.copy(), eps.copy()
        ep[i] += h
        em[i] -= h
        Jp, _ = objective(make(s, xyz, dt, vol), psi0, ep, phi, vol, s.np)
        Jm, _ = objective(make(s, xyz, dt, vol), psi0, em, phi, vol, s.np)
        fd[i] = (Jp - Jm) / (2 * h)
    assert np.max(np.abs(deps - fd)) <= 1e-6 * np.max(np.abs(fd))


def test_dv_matches_central_differences():
    s, vol, psi0, phi, xyz, eps = setup()
    dt = 0.05
    J, chi0, deps, dv = gradients(s, vol, psi0, phi, xyz, eps, dt)
    h = 1e-6
    x = xyz
    direction = np.cos(0.7 * x[:, 0] + 0.3) * np.exp(-np.sum(x * x, axis=1) / 4.0)
    vp, vm = s.v + h * direction, s.v - h * direction
    Jp, _ = objective(make(s, xyz, dt, vol, v=vp), psi0, eps, phi, vol, s.np)
    Jm, _ = objective(make(s, xyz, dt, vol, v=vm), psi0, eps, phi, vol, s.np)
    fd = (Jp - Jm) / (2 * h)
    assert abs(np.dot(direction, dv) - fd) <= 1e-6 * abs(fd)
    # single points where the gradient is largest: a looser bar (the change of J is tiny)
    for r in np.argsort(-np.abs(dv))[:3]:
        vp, vm = s.v.copy(), s.v.copy()
        vp[r] += 1e-4
        vm[r] -= 1e-4
        Jp, _ = objective(make(s, xyz, dt, vol, v=vp), psi0, eps, phi, vol, s.np)
        Jm, _ = objective(make(s, xyz, dt, vol, v=vm), psi0, eps, phi, vol, s.np)
        assert abs(dv[r] - (Jp - Jm) / 2e-4) <= 1e-6 * np.max(np.abs(dv))


def test_chi0_is_the_initial_state_gradient():
    s, vol, psi0, phi, xyz, eps = setup()
    dt = 0.05
    J, chi0, deps, dv = gradients(s, vol, psi0, phi, xyz, eps, dt)
    rnd = np.random.default_rng(5).standard_normal(psi0.shape)
    rnd[s.np:] = 0.0
    h = 1e-6
    Jp, _ = objective(make(s, xyz, dt, vol), psi0 + h * rnd, eps, phi, vol, s.np)
    Jm, _ = objective(make(s, xyz, dt, vol), psi0 - h * rnd, eps, phi, vol, s.np)
    fd = (Jp - Jm) / (2 * h)
    an = 2.0 * vol * np.sum(chi0[: s.np] * rnd[: s.np])  # 2 Re <chi_0|d psi_0> on packed reals
    assert abs(an - fd) <= 1e-6 * abs(fd)


def test_reference_formula_is_the_first_order_limit():
    """|exact - 2 dt Im<chi|x|psi>| / |exact| is O(dt): halving dt halves it.  The limit
    needs dt ||H|| << 1 (||H|| ~ 100 Ha at h = 0.3), hence the short run."""
    T = 0.04
    rels = []
    for nsteps in (16, 32):
        dt = T / nsteps
        s, vol, psi0, phi, xyz, eps = setup(nsteps=nsteps, dt=dt)
        adj = make(s, xyz, dt, vol)
        J, chiN = objective(adj, psi0, eps, phi, vol, s.np)
        chi0, deps, dv = adj.backward(chiN)
        # chi_{s+1} for every step: rerun the backward recursion keeping the states
        chis = [None] * nsteps
        chi = O.to_complex(chiN, s.np)
        for st in range(nsteps - 1, -1, -1):
            chis[st] = chi
            adj._set(eps[st])
            yc = adj._powers(chi, adj.order + 1)
            chi = sum(np.conj(adj.zf[m]) * yc[m] for m in range(adj.order + 1))
        oct_ = np.array([2.0 * dt * np.imag(vol * np.sum(np.conj(chis[st]) * xyz[:, 0][:, None] * adj.psis[st]))
                         for st in range(nsteps)])
        rels.append(np.max(np.abs(deps - oct_)) / np.max(np.abs(deps)))
    ratio = rels[0] / rels[1]
    assert 1.8 < ratio < 2.3, (rels, ratio)


@pytest.mark.parametrize("mistake", ["field_sign", "drop_second_order"])
def test_planted_adjoint_mistakes_are_caught(mistake):
    s, vol, psi0, phi, xyz, eps = setup(nsteps=3)
    dt = 0.05
    adj = make(s, xyz, dt, vol)
    J, chiN = objective(adj, psi0, eps, phi, vol, s.np)
    if mistake == "field_sign":
        adj.xp = -adj.xp
    else:
        adj.zf[2] = 0.0
    _, deps, _ = adj.backward(chiN)
    h = 1e-6
    fd = np.zeros_like(deps)
    for i in range(len(eps)):
        ep, em = eps.copy(), eps.copy()
        ep[i] += h
        em[i] -= h
        Jp, _ = objective(make(s, xyz, dt, vol), psi0, ep, phi, vol, s.np)
        Jm, _ = objective(make(s, xyz, dt, vol), psi0, em, phi, vol, s.np)
        fd[i] = (Jp - Jm) / (2 * h)
    assert np.max(np.abs(deps - fd)) > 1e-4 * np.max(np.abs(fd))
