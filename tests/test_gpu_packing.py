"""A state's result does not depend on the batch it is packed in.

In the reference a batch (batch_t, src/grid/batch.F90:47-116) only groups states: every
state is applied independently (the stencil per lane, operate_inc.c:29-342; the projector
gemms per column, hamiltonian_elec_base_inc.F90:770-772, 916-918).  Here a 16-state
complex batch has 32-lane rows, whose MFMA bra and ket run on 32-lane tiles (two columns
per lane), and an 8-state batch 16-lane rows (16-lane tiles); each column's sums are
formed in the same order in both, so H applied to the 16-state batch equals, bit for bit,
H applied to its two halves packed as 8-state batches -- for the H-apply and for Taylor
steps.  The spheres overlap, so the ket sums are covered.  (The plain bra, BRA_MODE 0,
splits sphere points over BLOCK / row-length slices, so its sums are ordered by the
packing; it is held to the oracle's 1e-12 bar, not to bit equality across packings.)
"""
import numpy as np
import pytest

from helpers import small_sphere
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.fixture(scope="module")
def system():
    return small_sphere(radius=3.5, spacing=0.25, natoms=5, rc=1.5)


def test_h_apply_independent_of_packing(gpu, system):
    print("[packing h-apply]", flush=True)
    s = system
    z = workloads.gaussian_envelope_states(s, 16, width=2.0)
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=1)

    def apply(zz):
        psib = api.Batch.from_states(zz, s.np_part)
        out = api.Batch(zz.shape[0], s.np, True)
        api.hamiltonian_elec_apply_batch(hm, psib, out)
        gpu.cuda.synchronize()
        return out.to_states(s.np)

    whole = apply(z)
    halves = np.concatenate([apply(z[:8]), apply(z[8:])])
    assert np.max(np.abs(whole)) > 0.0
    assert same_bits(whole, halves)


def test_taylor_independent_of_packing(gpu, system):
    print("[packing taylor]", flush=True)
    s = system
    z = workloads.gaussian_envelope_states(s, 16, width=2.0)
    hm = api.HamiltonianDevice.from_system(s)

    def steps(zz):
        psib = api.Batch.from_states(zz, s.np_part)
        for _ in range(3):
            api.Exponential(4).apply_batch(hm, psib, 0.01)
        gpu.cuda.synchronize()
        return psib.to_states(s.np)

    whole = steps(z)
    halves = np.concatenate([steps(z[:8]), steps(z[8:])])
    assert same_bits(whole, halves)
