"""200 Taylor-4 steps at the headline spacing h = 0.25 on the device vs the oracle.

The order-4 star's spectral radius at h = 0.25 is 156 Ha, so dt = 0.02 gives
theta = dt E_max = 3.12 > 2 sqrt(2): Taylor-4 amplifies the top modes (up to 1.95x per
step) and the norms blow up (to ~1e105 here after 200 steps), for the reference's
propagator as for this one.  Rounding differences grow with the same modes as the
states themselves, so the relative state error and norm drift stay at the rounding level
(measured 2.6e-15 and 1.6e-15 after 200 steps on MI355X); both runs carry the 1e-10 bar.
The stable companion dt = 0.015 (theta = 2.34) also checks that the norm does not grow.
"""
import numpy as np
import pytest

import oracle as O
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def run(gpu, dt, nsteps=200):
    s = workloads.hamiltonian_system(radius=5.0, spacing=0.25, seed=77, nwells=6, natoms=4, rc=1.5,
                                     name="taylor-long")
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, 16, width=2.0), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    hm = api.HamiltonianDevice.from_system(s)
    psib = api.Batch.from_states(z, s.np_part)
    g = api.TaylorGraph(api.Exponential(4), hm, [psib], dt)
    for _ in range(nsteps):
        oh.taylor(psi, dt)
        g.launch()
    gpu.cuda.synchronize()
    A = workloads.unpack(psib.ff.cpu().numpy(), 16, s.np, True)
    B = workloads.unpack(psi, 16, s.np, True)
    err = float(np.max(np.max(np.abs(A - B), axis=1) / np.max(np.abs(B), axis=1)))
    nA = np.sum(np.abs(A) ** 2, axis=1) * s.mesh.vol_pp
    nB = np.sum(np.abs(B) ** 2, axis=1) * s.mesh.vol_pp
    drift = float(np.max(np.abs(nA - nB) / nB))
    print(f"dt={dt}: state error {err:.3e}, norm drift {drift:.3e}, norms {nB.min():.3e}..{nB.max():.3e}",
          flush=True)
    return err, drift, nB


def test_taylor_200_steps_stable_dt(gpu):
    err, drift, norms = run(gpu, 0.015)
    assert err <= 1e-10 and drift <= 1e-10
    assert np.all(norms <= 1.0 + 1e-12)  # theta < 2 sqrt(2): norm non-increasing


def test_taylor_200_steps_headline_dt_is_unstable(gpu):
    err, drift, norms = run(gpu, 0.02)
    assert np.all(norms > 1e50)          # the instability, in both propagators
    assert err <= 1e-10 and drift <= 1e-10  # relative to the grown states
