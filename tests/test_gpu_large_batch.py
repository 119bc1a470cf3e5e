"""Batches beyond 4 GiB: the z-march kernel addresses rows with 32-bit byte offsets, so a
batch whose rows do not fit ((np_part + 1) * row bytes > 2^32) is sent to the gather
kernel instead.  Both kernels sum the stencil terms in the same order, so the first 64
lanes of a 1024-lane batch on the R=12 sphere (4.6 GB, gather kernel) must equal a
64-lane batch of the same values (z-march) bit for bit; a Taylor step likewise, to
rounding of the complex epilogue.
"""
import pytest

from octhip import api, workloads

pytestmark = pytest.mark.gpu


def test_rows_beyond_32bit_offsets_take_the_gather_kernel(gpu):
    torch = gpu
    print("[large batch]", flush=True)
    s = workloads.hamiltonian_system(natoms=4, name="large-batch")
    assert (s.np_part + 1) * 1024 * 8 > 2 ** 32
    hm = api.HamiltonianDevice.from_system(s)
    gen = torch.Generator(device="cuda").manual_seed(77)
    big = api.Batch(512, s.np_part, True)
    assert big.lanes == 1024
    big.ff.copy_(torch.randn(big.ff.shape, dtype=torch.float64, device="cuda", generator=gen))
    small = api.Batch(32, s.np_part, True, ff=big.ff[:, :64].contiguous())
    hbig = api.Batch(512, s.np, True)
    hsmall = api.Batch(32, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, big, hbig)
    api.hamiltonian_elec_apply_batch(hm, small, hsmall)
    torch.cuda.synchronize()
    assert torch.equal(hbig.ff[:, :64], hsmall.ff)
    del hbig, hsmall
    te = api.Exponential(4)
    te.apply_batch(hm, big, 0.002)
    te.apply_batch(hm, small, 0.002)
    torch.cuda.synchronize()
    a, b = big.ff[: s.np, :64], small.ff[: s.np]
    scale = b.abs().max()
    assert float((a - b).abs().max() / scale) <= 1e-13
