"""Device propagation extras against the oracle: exponential_apply_batch with psib2/deltat2,
inh_psib and the density of the propagated states (src/electrons/exponential.F90:581-757,
src/td/propagation_ops_elec.F90:261-333), the exponential-midpoint and ETRS steps for
independent particles (src/td/propagator_expmid.F90:52-100, propagator_etrs.F90:68-150)
and density_calc_accumulate (src/states/density.F90:375-417).

Bars: states <= 1e-12 relative per state (projector sums in MFMA order), densities
<= 1e-12 relative to their largest value (sums over states in a different order).
Orders 2..4 run the fused last pass (second accumulator and density partials in its
epilogue); orders 1 and 5+ and the inhomogeneous series run plain passes plus axpy
kernels.  Both stencil kernels (z-march and tiled gather) are covered.
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, small_sphere
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def system():
    return small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)


def rho_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def batch_of(gpu, host, nst):
    return api.Batch(nst, host.shape[0], True, ff=gpu.from_numpy(host.copy()).cuda())


@pytest.mark.parametrize("order,kernel", [(4, 0), (3, 0), (2, 0), (4, 1), (1, 0), (5, 0), (6, 1)])
def test_psib2_and_density(gpu, order, kernel):
    print(f"[psib2 + density order={order} kernel={kernel}]", flush=True)
    s = system()
    nst = 12
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    psi = workloads.pack(z, s.np_part)
    w = np.linspace(0.5, 2.0, nst)
    hm = api.HamiltonianDevice.from_system(s)
    if kernel:
        hm.tune(kernel=kernel)
    psib = batch_of(gpu, psi, nst)
    psib2 = batch_of(gpu, np.zeros_like(psi), nst)
    rho = gpu.full((s.np,), 0.25, dtype=gpu.float64, device="cuda")
    api.Exponential(order).apply_batch(hm, psib, 0.02, psib2=psib2, deltat2=0.01, density=rho, weights=w)
    gpu.cuda.synchronize()
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    ref = psi.copy()
    ref2 = np.zeros_like(psi)
    oh.taylor_ex(ref, 0.02, order, psi2=ref2, deltat2=0.01)
    rho_ref = np.full(s.np, 0.25)
    O.density_accumulate(ref2, s.np, nst, True, w, rho_ref)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    assert max_rel_err(psib2.ff.cpu().numpy()[: s.np], ref2[: s.np]) <= 1e-12
    assert not psib2.ff[s.np:].any()  # only the inner rows of psib2 are written
    assert rho_err(rho.cpu().numpy(), rho_ref) <= 1e-12


@pytest.mark.parametrize("order", [4, 3, 1])
def test_density_without_psib2(gpu, order):
    print(f"[density of psib order={order}]", flush=True)
    s = system()
    nst = 16
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    psi = workloads.pack(z, s.np_part)
    w = np.r_[np.full(10, 2.0), np.zeros(6)]
    hm = api.HamiltonianDevice.from_system(s)
    psib = batch_of(gpu, psi, nst)
    rho = gpu.zeros((s.np,), dtype=gpu.float64, device="cuda")
    api.Exponential(order).apply_batch(hm, psib, 0.015, density=rho, weights=w)
    gpu.cuda.synchronize()
    ref = psi.copy()
    O.Hamiltonian(s.op, s.v, s.proj).taylor(ref, 0.015, order)
    rho_ref = np.zeros(s.np)
    O.density_accumulate(ref, s.np, nst, True, w, rho_ref)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    assert rho_err(rho.cpu().numpy(), rho_ref) <= 1e-12


@pytest.mark.parametrize("with_psib2", [False, True])
def test_inhomogeneous_term(gpu, with_psib2):
    print(f"[inh_psib psib2={with_psib2}]", flush=True)
    s = system()
    nst = 8
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    g = workloads.gaussian_envelope_states(s, nst, width=1.5, first=40)
    psi, inh = workloads.pack(z, s.np_part), workloads.pack(g, s.np_part)
    inh[s.np:] = 1.0  # boundary rows: the H-apply of inh zeroes them, as in the reference
    w = np.ones(nst)
    hm = api.HamiltonianDevice.from_system(s)
    psib, inhb = batch_of(gpu, psi, nst), batch_of(gpu, inh, nst)
    psib2 = batch_of(gpu, np.zeros_like(psi), nst) if with_psib2 else None
    rho = gpu.zeros((s.np,), dtype=gpu.float64, device="cuda")
    api.Exponential(4).apply_batch(hm, psib, 0.02, psib2=psib2, deltat2=0.005 if with_psib2 else None,
                                   inh_psib=inhb, density=rho, weights=w)
    gpu.cuda.synchronize()
    ref, ref_inh = psi.copy(), inh.copy()
    ref2 = np.zeros_like(psi) if with_psib2 else None
    O.Hamiltonian(s.op, s.v, s.proj).taylor_ex(ref, 0.02, 4, psi2=ref2, deltat2=0.005, inh=ref_inh)
    rho_ref = np.zeros(s.np)
    O.density_accumulate(ref2 if with_psib2 else ref, s.np, nst, True, w, rho_ref)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    if with_psib2:
        assert max_rel_err(psib2.ff.cpu().numpy()[: s.np], ref2[: s.np]) <= 1e-12
    assert np.array_equal(inhb.ff.cpu().numpy(), ref_inh)
    assert rho_err(rho.cpu().numpy(), rho_ref) <= 1e-12


def test_imaginary_time_with_psib2(gpu):
    print("[imaginary time psib2]", flush=True)
    s = system()
    z = workloads.gaussian_envelope_states(s, 4, width=2.0)
    psi = workloads.pack(z, s.np_part)
    hm = api.HamiltonianDevice.from_system(s)
    psib, psib2 = batch_of(gpu, psi, 4), batch_of(gpu, np.zeros_like(psi), 4)
    api.Exponential(4).apply_batch(hm, psib, 0.01, True, psib2=psib2, deltat2=0.02)
    gpu.cuda.synchronize()
    ref, ref2 = psi.copy(), np.zeros_like(psi)
    O.Hamiltonian(s.op, s.v, s.proj).taylor_ex(ref, 0.01, 4, True, psi2=ref2, deltat2=0.02)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    assert max_rel_err(psib2.ff.cpu().numpy()[: s.np], ref2[: s.np]) <= 1e-12


@pytest.mark.parametrize("driver", ["expmid", "etrs"])
def test_propagator_steps(gpu, driver):
    print(f"[propagator {driver}]", flush=True)
    s = system()
    x = s.mesh.x[: s.np]
    z = workloads.gaussian_envelope_states(s, 20, width=2.0)
    chunks = [z[0:8], z[8:16], z[16:20]]
    hosts = [workloads.pack(c, s.np_part, 8) for c in chunks]
    weights = [np.full(len(c), 2.0) for c in chunks]
    hm = api.HamiltonianDevice.from_system(s)
    hm.set_dipole_field([0.0, 0.0, 0.0], positions=x)
    batches = [api.Batch(len(c), h.shape[0], True, 8, ff=gpu.from_numpy(h.copy()).cuda())
               for h, c in zip(hosts, chunks)]
    rho = gpu.full((s.np,), 7.0, dtype=gpu.float64, device="cuda")  # zeroed by the driver
    oh = O.Hamiltonian(s.op, s.v.copy(), s.proj, xyz=x, field=[0.0, 0.0, 0.0])
    te = api.Exponential(4)
    refs = [h.copy() for h in hosts]
    for step in range(3):
        t = 0.02 * (step + 1)
        e = lambda tt: np.array([0.05 * np.sin(3.0 * tt), 0.0, 0.02])  # noqa: E731
        if driver == "expmid":
            api.propagator_expmid(te, hm, batches, 0.02, e(t - 0.01), weights, rho)
            rho_ref = O.expmid_step(oh, refs, 0.02, e(t - 0.01), weights=weights)
        else:
            api.propagator_etrs(te, hm, batches, 0.02, e(t - 0.02), e(t), weights, rho)
            rho_ref = O.etrs_step(oh, refs, 0.02, e(t - 0.02), e(t), weights=weights)
    gpu.cuda.synchronize()
    for b, r in zip(batches, refs):
        assert max_rel_err(b.ff.cpu().numpy()[: s.np], r[: s.np]) <= 1e-12
    assert rho_err(rho.cpu().numpy(), rho_ref) <= 1e-12


def test_density_accumulate_real_and_complex(gpu):
    print("[density_calc_accumulate]", flush=True)
    s = system()
    z = workloads.gaussian_envelope_states(s, 6, width=2.0)
    w = np.array([1.0, 2.0, 0.5, 0.25, 3.0, 1.5])
    for f, cplx in ((z, True), (np.ascontiguousarray(z.real), False)):
        host = workloads.pack(f, s.np_part, 8)
        b = api.Batch(6, s.np_part, cplx, 8, ff=gpu.from_numpy(host.copy()).cuda())
        rho = gpu.full((s.np,), 0.5, dtype=gpu.float64, device="cuda")
        api.density_calc_accumulate(b, s.np, w, rho)
        ref = np.full(s.np, 0.5)
        O.density_accumulate(host, s.np, 6, cplx, w, ref)
        assert rho_err(rho.cpu().numpy(), ref) <= 1e-12
