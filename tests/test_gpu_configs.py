"""Oracle parity at the full sizes of the three propagation workloads.

* Taylor graph: 128 orthonormalised complex states (4 batches of 32, batch b on graph
  branch b mod 4) on the R=12, h=0.25 sphere with 16 wells and 16 KB atoms; 200
  graph-captured Taylor-4 steps at dt = 0.02 against the oracle's propagator
  (src/electrons/exponential.F90:677-757 restated).  Bars: state error and norm drift
  <= 1e-10 relative after 200 steps.  dt = 0.02 is beyond the order-4 Taylor stability
  limit at h = 0.25 (theta = dt E_max = 3.12 > 2 sqrt 2), so the norms grow by ~1e100 in
  both propagators; the error is relative to the grown states.
* Uneven systems: the eight spheres R = 6..14 (8..256 states) through one batched launch
  sequence: one H-apply and 5 Taylor-4 steps against the oracle (<= 1e-12 relative per
  state), then 50 steps bitwise against per-pair launches.
* Adjoint: 100 Taylor-4 steps on the R=10, h=0.3 sphere, 32 states, dipole field of five
  sinusoids, checkpoints every 10 steps: dJ/deps and dJ/dv against the oracle's exact
  discrete adjoint (<= 1e-10), and against central differences of the device forward
  (h = 1e-6) on three eps entries and one v direction (<= 1e-6).
"""
import time

import numpy as np
import pytest

import oracle as O
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def state_error_and_drift(A, B, vol):
    """A, B: (nst, np) complex.  Worst per-state max|A-B|/max|B| and relative norm drift."""
    err = float(np.max(np.max(np.abs(A - B), axis=1) / np.max(np.abs(B), axis=1)))
    nA = np.sum(np.abs(A) ** 2, axis=1) * vol
    nB = np.sum(np.abs(B) ** 2, axis=1) * vol
    return err, float(np.max(np.abs(nA - nB) / nB)), nB


def test_taylor_graph_200_steps_128_states_R12(gpu):
    note("[taylor graph: R=12 h=0.25, 128 states, 200 steps, dt=0.02]")
    s = workloads.hamiltonian_system()
    vol = s.mesh.vol_pp
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, 128), vol)
    dt, nsteps = 0.02, 200
    batches = [api.Batch.from_states(z[32 * b:32 * (b + 1)], s.np_part) for b in range(4)]
    g = api.TaylorGraph(api.Exponential(4), api.HamiltonianDevice.from_system(s), batches, dt)
    for _ in range(nsteps):
        g.launch()
    gpu.cuda.synchronize()
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    worst_err = worst_drift = 0.0
    for b in range(4):
        t0 = time.perf_counter()
        psi = workloads.pack(z[32 * b:32 * (b + 1)], s.np_part)
        for k in range(nsteps):
            oh.taylor(psi, dt)
        A = batches[b].to_states(s.np)
        B = workloads.unpack(psi, 32, s.np, True)
        err, drift, norms = state_error_and_drift(A, B, vol)
        note(f"  batch {b}: state error {err:.3e}, norm drift {drift:.3e}, norms {norms.min():.3e}.."
             f"{norms.max():.3e} (oracle {time.perf_counter() - t0:.0f} s)")
        worst_err, worst_drift = max(worst_err, err), max(worst_drift, drift)
    assert worst_err <= 1e-10 and worst_drift <= 1e-10


def uneven_pairs(systems, pack=32):
    pairs, meta = [], []
    for s, nst in systems:
        hm = api.HamiltonianDevice.from_system(s)
        z = workloads.gaussian_envelope_states(s, nst)
        for first in range(0, nst, pack):
            zz = z[first:first + pack]
            pairs.append((hm, api.Batch.from_states(zz, s.np_part)))
            meta.append((s, zz))
    return pairs, meta


def test_uneven_systems_batched_vs_oracle_and_plain(gpu):
    note("[uneven systems: 8 spheres R=6..14, 8..256 states, batched launch]")
    systems = workloads.uneven_systems()
    dt = 0.02
    pairs, meta = uneven_pairs(systems)
    plain = [(hm, b.clone()) for hm, b in pairs]
    bp = api.BatchedPropagator(pairs, dt)
    bp.apply()
    for _ in range(5):
        bp.taylor_step()
    gpu.cuda.synchronize()
    # oracle: the same H-apply and 5 steps, pair by pair
    hams = {}
    worst_h = worst_t = 0.0
    t0 = time.perf_counter()
    for i, ((s, zz), hp, (_, b)) in enumerate(zip(meta, bp.hpsi, pairs)):
        oh = hams.setdefault(id(s), O.Hamiltonian(s.op, s.v, s.proj))
        psi = workloads.pack(zz, s.np_part)
        ref = workloads.unpack(oh.apply(psi.copy()), len(zz), s.np, True)
        got = hp.to_states(s.np)
        worst_h = max(worst_h, float(np.max(np.max(np.abs(got - ref), axis=1) / np.max(np.abs(ref), axis=1))))
        for _ in range(5):
            oh.taylor(psi, dt)
        ref = workloads.unpack(psi, len(zz), s.np, True)
        got = b.to_states(s.np)
        worst_t = max(worst_t, float(np.max(np.max(np.abs(got - ref), axis=1) / np.max(np.abs(ref), axis=1))))
    note(f"  H-apply {worst_h:.3e}, 5 Taylor steps {worst_t:.3e} over {len(pairs)} pairs "
         f"(oracle {time.perf_counter() - t0:.0f} s)")
    assert worst_h <= 1e-12 and worst_t <= 1e-12
    # the plain per-pair launches run the same 1 + 5 steps, then both run 45 more: bitwise
    te = api.Exponential(4)
    for hm, b in plain:
        out = api.Batch(b.nst, hm.np, True, b.pack_size)
        api.hamiltonian_elec_apply_batch(hm, b.clone(), out)
        for _ in range(5):
            te.apply_batch(hm, b, dt)
    for _ in range(45):
        bp.taylor_step()
        for hm, b in plain:
            te.apply_batch(hm, b, dt)
    gpu.cuda.synchronize()
    for (hm, b), (_, p) in zip(pairs, plain):
        assert gpu.equal(b.ff[: hm.np], p.ff[: hm.np])


def test_adjoint_R10_100_steps_checkpoints_10(gpu):
    from octhip.autograd import overlap_objective
    note("[adjoint: R=10 h=0.3, 32 states, 100 steps, K=10]")
    s = workloads.adjoint_system()
    nsteps, dt, K = 100, 0.02, 10
    z, tg, eps = workloads.adjoint_problem(s, 32, nsteps, dt)
    vol = s.mesh.vol_pp
    psi0 = workloads.pack(z, s.np_part)
    phi = workloads.pack(tg, s.np_part)
    hm = api.HamiltonianDevice.from_system(s)
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], 6, nsteps, dt, vol, checkpoint_every=K)
    phib = api.Batch(32, s.np_part, True, ff=gpu.from_numpy(phi.copy()).cuda())

    def forward_J(e, v=None):
        if v is not None:
            hm.set_potential(v)
        p = api.Batch(32, s.np_part, True, ff=gpu.from_numpy(psi0.copy()).cuda())
        adj.forward(p, e)
        J, chi = overlap_objective(phib, p, s.np, vol)
        if v is not None:
            hm.set_potential(s.v)
        return float(J), p, chi

    J, psib, chib = forward_J(eps)
    deps, dv = adj.backward(chib)
    gpu.cuda.synchronize()
    dv = dv.cpu().numpy()
    psiN = psib.ff.cpu().numpy()
    # oracle: exact discrete adjoint keeping every state
    t0 = time.perf_counter()
    h = O.Hamiltonian(s.op, s.v.copy(), s.proj, xyz=s.mesh.x[: s.np], field=[0.0, 0.0, 0.0])
    oa = O.TaylorAdjoint(h, s.mesh.x[: s.np], [1.0, 0.0, 0.0], dt, order=4, vol=vol)
    psiN_ref = oa.forward(psi0, eps)
    J_ref, chiN_ref = O.overlap_objective(phi, psiN_ref, s.np, vol)
    _, deps_ref, dv_ref = oa.backward(chiN_ref)
    del oa

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    note(f"  psi_N {rel(psiN[: s.np], psiN_ref[: s.np]):.3e}, J {abs(J - J_ref) / abs(J_ref):.3e}, "
         f"dJ/deps {rel(deps, deps_ref):.3e}, dJ/dv {rel(dv, dv_ref):.3e} (oracle {time.perf_counter() - t0:.0f} s)")
    assert rel(psiN[: s.np], psiN_ref[: s.np]) <= 1e-12
    assert abs(J - J_ref) <= 1e-12 * abs(J_ref)
    assert rel(deps, deps_ref) <= 1e-10
    assert rel(dv, dv_ref) <= 1e-10
    # central differences of the device forward, h = 1e-6
    hstep = 1e-6
    for k in (0, nsteps // 2, nsteps - 1):
        ep, em = eps.copy(), eps.copy()
        ep[k] += hstep
        em[k] -= hstep
        fd = (forward_J(ep)[0] - forward_J(em)[0]) / (2 * hstep)
        note(f"  dJ/deps[{k}]: adjoint {deps[k]:.9e}, central difference {fd:.9e}")
        assert abs(fd - deps[k]) <= 1e-6 * max(abs(deps[k]), np.max(np.abs(deps)))
    x = s.mesh.x[: s.np]
    d = np.exp(-np.einsum("ij,ij->i", x - 1.0, x - 1.0) / 4.0)  # a smooth direction in v
    fd = (forward_J(eps, s.v + hstep * d)[0] - forward_J(eps, s.v - hstep * d)[0]) / (2 * hstep)
    ad = float(np.dot(dv, d))
    note(f"  dJ/dv . d: adjoint {ad:.9e}, central difference {fd:.9e}")
    assert abs(fd - ad) <= 1e-6 * abs(ad)
