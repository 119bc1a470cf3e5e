"""KB projector matrices of any projector count through the device H-apply and Taylor step.

The synthetic KB sets of the workloads have 4 projectors per matrix (l <= 1); real sets
have 1, 8, 9, 16 or more.  These flat build_accel arrays (random columns, overlapping
spheres) exercise every projector code path: the second 16-projector block of the MFMA
bra, the projector masking of the MFMA ket when a count is not a multiple of 4, and the
separate bra reduce with the plain ket for matrices of more than 32 projectors.  Bar:
max relative error <= 1e-12 per state against the oracle (the device sums sphere points
in its own blocked order).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, small_sphere, synthetic_projectors
from octhip import api, workloads

pytestmark = pytest.mark.gpu

CENTERS = [(0.0, 0.0, 0.0), (0.6, 0.0, 0.0), (0.0, 0.7, 0.3), (-0.9, -0.3, 0.6), (0.3, -0.8, -0.6),
           (1.1, 0.9, -0.2)]
COUNTS = {"one": [1] * 3, "nine": [9] * 3, "seventeen": [17] * 3, "thirtythree": [33] * 3,
          "mixed": [1, 9, 17, 33, 4, 2]}


def system_with(counts, seed):
    s = small_sphere(radius=3.0, spacing=0.3, natoms=0)
    proj = synthetic_projectors(s.mesh, CENTERS[: len(counts)], 1.2, counts, seed)
    return dataclasses.replace(s, proj=proj)


@pytest.mark.parametrize("bra", [0, 1])
@pytest.mark.parametrize("nst", [4, 16])
@pytest.mark.parametrize("name", list(COUNTS))
def test_h_apply_projector_counts(gpu, name, nst, bra):
    print(f"[projectors {name} nst={nst} bra={bra}]", flush=True)
    s = system_with(COUNTS[name], seed=len(name) * 7 + nst)
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=bra)
    psib = api.Batch.from_states(z, s.np_part)
    hpsib = api.Batch(nst, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib)
    gpu.cuda.synchronize()
    ref = O.Hamiltonian(s.op, s.v, s.proj).apply(workloads.pack(z, s.np_part))
    assert max_rel_err(hpsib.ff.cpu().numpy(), ref) <= 1e-12
    # the nonlocal term alone, so that a wrong projector term cannot hide behind the kinetic one
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib, api.TERM_NON_LOCAL_POTENTIAL)
    gpu.cuda.synchronize()
    ref_nl = O.Hamiltonian(s.op, s.v, s.proj).apply(workloads.pack(z, s.np_part), O.TERM_NON_LOCAL)
    assert np.max(np.abs(ref_nl)) > 0.0
    assert max_rel_err(hpsib.ff.cpu().numpy(), ref_nl) <= 1e-12


@pytest.mark.parametrize("bra", [0, 1])
@pytest.mark.parametrize("name", ["nine", "thirtythree", "mixed"])
def test_taylor_projector_counts(gpu, name, bra):
    print(f"[projectors taylor {name} bra={bra}]", flush=True)
    s = system_with(COUNTS[name], seed=11 + len(name))
    z = workloads.gaussian_envelope_states(s, 16, width=2.0)
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=bra)
    psib = api.Batch.from_states(z, s.np_part)
    psi = workloads.pack(z, s.np_part)
    ham = O.Hamiltonian(s.op, s.v, s.proj)
    for _ in range(3):
        api.Exponential(4).apply_batch(hm, psib, 0.01)
        ham.taylor(psi, 0.01, 4)
    gpu.cuda.synchronize()
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-12
