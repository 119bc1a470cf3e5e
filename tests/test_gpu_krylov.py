"""Device mesh reductions and the Lanczos / Chebyshev exponentials against the oracle
(src/grid/mesh_batch_inc.F90:379-549, 1168-1274; src/electrons/exponential.F90:766-971).

Bars: dot products and norms <= 1e-13 relative (deterministic sums in a different order
from numpy's); orthogonalised batches and overlaps <= 1e-12; Lanczos and Chebyshev states
<= 1e-12 relative per state, with the same Krylov dimension as the oracle.  The Lanczos
projected problem is solved on the host in both (tql2 on the device side, LAPACK zheev
through numpy on the oracle side), so the expansion coefficients agree to rounding.
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, small_sphere
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def system():
    return small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)


def dev(gpu, host, nst, cplx=True, pack=None):
    return api.Batch(nst, host.shape[0], cplx, pack, ff=gpu.from_numpy(host.copy()).cuda())


def test_dotp_and_nrm2(gpu):
    print("[mesh_batch_dotp_vector / nrm2]", flush=True)
    s = system()
    vol = s.mesh.vol_pp
    a = workloads.gaussian_envelope_states(s, 12, width=2.0)
    b = workloads.gaussian_envelope_states(s, 12, width=1.5, first=50)
    pa, pb = workloads.pack(a, s.np_part), workloads.pack(b, s.np_part)
    da, db = dev(gpu, pa, 12), dev(gpu, pb, 12)
    got = api.mesh_batch_dotp_vector(vol, da, db, s.np)
    ref = O.mesh_dotp(pa, pb, s.np, 12, vol)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    nrm = api.mesh_batch_nrm2(vol, da, s.np)
    assert np.max(np.abs(nrm - np.sqrt(O.mesh_dotp(pa, pa, s.np, 12, vol).real))) <= 1e-13 * np.max(nrm)
    # real batch: one dot per lane
    ra, rb = np.ascontiguousarray(a.real), np.ascontiguousarray(b.imag)
    pra, prb = workloads.pack(ra, s.np_part, 16), workloads.pack(rb, s.np_part, 16)
    got = api.mesh_batch_dotp_vector(vol, dev(gpu, pra, 12, False, 16), dev(gpu, prb, 12, False, 16), s.np)
    ref = vol * np.sum(ra * rb, axis=1)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_orthogonalization(gpu):
    print("[mesh_batch_orthogonalization]", flush=True)
    s = system()
    vol = s.mesh.vol_pp
    q = [workloads.orthonormalise(workloads.gaussian_envelope_states(s, 8, width=2.0, first=10 * k), vol)
         for k in range(3)]
    phi = workloads.gaussian_envelope_states(s, 8, width=1.7, first=99)
    hq = [workloads.pack(x, s.np_part) for x in q]
    hphi = workloads.pack(phi, s.np_part)
    dphi = dev(gpu, hphi, 8)
    ov, nrm = api.mesh_batch_orthogonalization(vol, [dev(gpu, h, 8) for h in hq], dphi, s.np, normalize=True)
    gpu.cuda.synchronize()
    # classical Gram-Schmidt on the host
    ss = [O.mesh_dotp(h, hphi, s.np, 8, vol) for h in hq]
    ref = phi.copy()
    for x, o in zip(q, ss):
        ref -= o[:, None] * x
    rn = np.sqrt(vol * np.sum(np.abs(ref) ** 2, axis=1))
    assert np.max(np.abs(ov - np.array(ss))) <= 1e-12 * np.max(np.abs(ss))
    assert np.max(np.abs(nrm - rn)) <= 1e-12 * np.max(rn)
    got = dphi.to_states(s.np)
    assert max_rel_err(got.T, (ref / rn[:, None]).T) <= 1e-12


@pytest.mark.parametrize("case", ["plain", "imag", "inh", "psib2"])
def test_lanczos(gpu, case):
    print(f"[lanczos {case}]", flush=True)
    s = system()
    vol = s.mesh.vol_pp
    nst = 12
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    psi = workloads.pack(z, s.np_part)
    hm = api.HamiltonianDevice.from_system(s)
    te = api.Exponential(16, method="lanczos", lanczos_tol=1e-8)
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    dt = 0.02
    psib = dev(gpu, psi, nst)
    ref = psi.copy()
    if case == "plain":
        te.apply_batch(hm, psib, dt)
        it, res = O.lanczos_batch(oh, ref, nst, dt, 16, 1e-8, vol)
    elif case == "imag":
        te.apply_batch(hm, psib, -dt, imag_time=True)
        it, res = O.lanczos_batch(oh, ref, nst, 1j * -dt, 16, 1e-8, vol)
    elif case == "inh":
        g = workloads.pack(workloads.gaussian_envelope_states(s, nst, width=1.5, first=77), s.np_part)
        te.apply_batch(hm, psib, dt, inh_psib=dev(gpu, g, nst))
        O.lanczos_batch(oh, ref, nst, dt, 16, 1e-8, vol)
        it, res = O.lanczos_batch(oh, ref, nst, dt, 16, 1e-8, vol, inh=g)
    else:
        psib2 = dev(gpu, np.zeros_like(psi), nst)
        te.apply_batch(hm, psib, dt, psib2=psib2, deltat2=0.5 * dt)
        ref2 = psi.copy()
        O.lanczos_batch(oh, ref, nst, dt, 16, 1e-8, vol)
        it, res = O.lanczos_batch(oh, ref2, nst, 0.5 * dt, 16, 1e-8, vol)
        assert max_rel_err(psib2.ff.cpu().numpy()[: s.np], ref2[: s.np]) <= 1e-12
    gpu.cuda.synchronize()
    print(f"  Krylov dimension {te.last_iterations} (oracle {it}), residual {te.last_residual:.2e}", flush=True)
    assert te.last_iterations == it
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12


@pytest.mark.parametrize("with_psib2", [False, True])
def test_chebyshev(gpu, with_psib2):
    print(f"[chebyshev psib2={with_psib2}]", flush=True)
    s = system()
    nst = 8
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    psi = workloads.pack(z, s.np_part)
    hm = api.HamiltonianDevice.from_system(s)
    # spectrum inside mid +- span: the order-4 star's radius 3 * 6.5016 / (2 h^2), the
    # wells (>= -2 Ha) and a margin for the projectors
    emax = 3 * 6.5016 / (2 * 0.3 ** 2) + 5.0
    hm.spectral_middle_point, hm.spectral_half_span = (emax - 7.0) / 2, (emax + 7.0) / 2
    te = api.Exponential(24, method="chebyshev")
    dt = 0.01
    psib = dev(gpu, psi, nst)
    psib2 = dev(gpu, np.zeros_like(psi), nst) if with_psib2 else None
    te.apply_batch(hm, psib, dt, psib2=psib2, deltat2=0.5 * dt if with_psib2 else None)
    gpu.cuda.synchronize()
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    ref = psi.copy()
    O.cheby_batch(oh, ref, dt, 24, hm.spectral_middle_point, hm.spectral_half_span)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    if with_psib2:
        ref2 = psi.copy()
        O.cheby_batch(oh, ref2, 0.5 * dt, 24, hm.spectral_middle_point, hm.spectral_half_span)
        assert max_rel_err(psib2.ff.cpu().numpy()[: s.np], ref2[: s.np]) <= 1e-12
