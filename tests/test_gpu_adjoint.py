"""The device Taylor adjoint (octhip_adjoint_*) against the oracle's adjoint, against
central finite differences of the device forward, and through the torch.autograd harness.

Bars: the gradients (dJ/deps, dJ/dv, chi_0) within 1e-10 of the oracle's relative to their
largest entry (both sides accumulate many H-applies in different orders); 1e-6 against
central differences with h = 1e-6; bitwise independence of the checkpoint interval (the
recomputation replays the same kernels on the same inputs).
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_sphere
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def problem(nst=8, nsteps=12, dt=0.01):
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    vol = s.mesh.vol_pp
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=2.0), vol)
    tg = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=1.5, first=64), vol)
    t = (np.arange(nsteps) + 0.5) * dt
    eps = 0.05 * np.sin(0.4 * t + 0.3) + 0.02
    return s, vol, workloads.pack(z, s.np_part), workloads.pack(tg, s.np_part), eps


def device_run(gpu, s, vol, psi0, phi, eps, dt, K, v=None, backward=True, order=4):
    from octhip.autograd import overlap_objective
    hm = api.HamiltonianDevice.from_system(s)
    if v is not None:
        hm.set_potential(v)
    nst = psi0.shape[1] // 2
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], int(np.log2(psi0.shape[1])), len(eps), dt, vol,
                      order=order, checkpoint_every=K)
    psib = api.Batch(nst, s.np_part, True, ff=gpu.from_numpy(psi0.copy()).cuda())
    phib = api.Batch(nst, s.np_part, True, ff=gpu.from_numpy(phi.copy()).cuda())
    adj.forward(psib, eps)
    J, chib = overlap_objective(phib, psib, s.np, vol)
    if not backward:
        return float(J)
    deps, dv = adj.backward(chib)
    gpu.cuda.synchronize()
    return float(J), psib.ff.cpu().numpy(), chib.ff.cpu().numpy(), deps, dv.cpu().numpy()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# nsteps 12, 10 and 11 leave chi_0 in each of the three buffers the backward rotates through
# (the caller's, then the two internal ones copied back); orders other than 4 take the
# generic accumulation kernel
@pytest.mark.parametrize("K,nsteps,order", [(5, 12, 4), (12, 12, 4), (4, 10, 4), (3, 11, 4), (4, 10, 2),
                                            (5, 11, 3), (4, 7, 6)])
def test_adjoint_matches_oracle(gpu, K, nsteps, order):
    print(f"[adjoint vs oracle K={K} nsteps={nsteps} order={order}]", flush=True)
    s, vol, psi0, phi, eps = problem(nsteps=nsteps)
    dt = 0.01
    J, psiN, chi0, deps, dv = device_run(gpu, s, vol, psi0, phi, eps, dt, K, order=order)
    h = O.Hamiltonian(s.op, s.v.copy(), s.proj, xyz=s.mesh.x[: s.np], field=[0.0, 0.0, 0.0])
    oa = O.TaylorAdjoint(h, s.mesh.x[: s.np], [1.0, 0.0, 0.0], dt, order=order, vol=vol)
    psiN_ref = oa.forward(psi0, eps)
    J_ref, chiN_ref = O.overlap_objective(phi, psiN_ref, s.np, vol)
    chi0_ref, deps_ref, dv_ref = oa.backward(chiN_ref)
    assert rel(psiN[: s.np], psiN_ref[: s.np]) <= 1e-12
    assert abs(J - J_ref) <= 1e-12 * abs(J_ref)
    assert rel(deps, deps_ref) <= 1e-10
    assert rel(dv, dv_ref) <= 1e-10
    assert rel(chi0[: s.np], chi0_ref[: s.np]) <= 1e-10


def test_checkpoint_interval_does_not_change_gradients(gpu):
    print("[adjoint checkpoint invariance]", flush=True)
    s, vol, psi0, phi, eps = problem(nsteps=9)
    out = [device_run(gpu, s, vol, psi0, phi, eps, 0.01, K) for K in (1, 4, 9)]
    for o in out[1:]:
        assert np.array_equal(o[3], out[0][3]) and np.array_equal(o[4], out[0][4])
        assert np.array_equal(o[2], out[0][2])


def test_backward_twice_gives_the_same_gradients(gpu):
    """A second backward on the same adjoint (same forward) reuses its streams, events and
    buffers and must reproduce the first bit for bit."""
    print("[adjoint backward twice]", flush=True)
    from octhip.autograd import overlap_objective
    s, vol, psi0, phi, eps = problem(nsteps=11)
    hm = api.HamiltonianDevice.from_system(s)
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], int(np.log2(psi0.shape[1])), len(eps), 0.01, vol,
                      checkpoint_every=4)
    psib = api.Batch(8, s.np_part, True, ff=gpu.from_numpy(psi0.copy()).cuda())
    phib = api.Batch(8, s.np_part, True, ff=gpu.from_numpy(phi.copy()).cuda())
    adj.forward(psib, eps)
    out = []
    for _ in range(2):
        J, chib = overlap_objective(phib, psib, s.np, vol)
        deps, dv = adj.backward(chib)
        gpu.cuda.synchronize()
        out.append((np.array(deps), dv.cpu().numpy(), chib.ff.cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_adjoint_matches_central_differences(gpu):
    print("[adjoint vs finite differences]", flush=True)
    s, vol, psi0, phi, eps = problem(nsteps=10)
    dt = 0.01
    J, psiN, chi0, deps, dv = device_run(gpu, s, vol, psi0, phi, eps, dt, 4)
    h = 1e-6
    for i in (0, 4, 9):
        ep, em = eps.copy(), eps.copy()
        ep[i] += h
        em[i] -= h
        fd = (device_run(gpu, s, vol, psi0, phi, ep, dt, 4, backward=False)
              - device_run(gpu, s, vol, psi0, phi, em, dt, 4, backward=False)) / (2 * h)
        assert abs(deps[i] - fd) <= 1e-6 * np.max(np.abs(deps)), (i, deps[i], fd)
    x = s.mesh.x[: s.np]
    d = np.cos(0.5 * x[:, 1] + 0.1) * np.exp(-np.sum(x * x, axis=1) / 6.0)
    fd = (device_run(gpu, s, vol, psi0, phi, eps, dt, 4, v=s.v + h * d, backward=False)
          - device_run(gpu, s, vol, psi0, phi, eps, dt, 4, v=s.v - h * d, backward=False)) / (2 * h)
    assert abs(np.dot(d, dv) - fd) <= 1e-6 * abs(fd)


def test_torch_autograd_harness(gpu):
    print("[adjoint autograd harness]", flush=True)
    from octhip.autograd import taylor_overlap_objective
    s, vol, psi0, phi, eps = problem(nsteps=6)
    dt = 0.01
    hm = api.HamiltonianDevice.from_system(s)
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], int(np.log2(psi0.shape[1])), len(eps), dt, vol,
                      checkpoint_every=3)
    psib = api.Batch(8, s.np_part, True, ff=gpu.from_numpy(psi0.copy()).cuda())
    phib = api.Batch(8, s.np_part, True, ff=gpu.from_numpy(phi.copy()).cuda())
    v = gpu.tensor(s.v, dtype=gpu.float64, device="cuda", requires_grad=True)
    e = gpu.tensor(eps, dtype=gpu.float64, requires_grad=True)
    J = taylor_overlap_objective(v, e, adj, psib, phib, vol)
    (3.0 * J).backward()
    # psi0 untouched by the harness
    assert np.array_equal(psib.ff.cpu().numpy(), psi0)
    J2, psiN, chi0, deps, dv = device_run(gpu, s, vol, psi0, phi, eps, dt, 3)
    assert float(J.detach()) == J2
    assert np.array_equal(e.grad.numpy(), 3.0 * deps)
    assert np.array_equal(v.grad.cpu().numpy(), 3.0 * dv)
