"""Device pack/unpack (octhip_batch_pack/unpack) against the host packer: the transposes of
batch_write_unpacked_to_device / batch_read_device_to_unpacked (src/grid/batch.F90:625-760,
share/opencl/pack.cl:23-67).  Pure data movement, so bit-exact: real and complex batches,
ragged sizes (points not a multiple of the 64-point tile, states not a multiple of 16),
states placed at an offset inside the row, a leading dimension larger than np, lanes of
other states left untouched, and an empty state range."""
import numpy as np
import pytest

from octhip import api, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cplx,nst,np_,ld,first,pack", [
    (True, 32, 4099, 4099, 0, 32), (True, 5, 1000, 1033, 3, 8), (False, 8, 777, 800, 0, 8),
    (False, 37, 2050, 2050, 20, 64), (True, 1, 65, 70, 0, 1), (True, 17, 300, 300, 15, 32)])
def test_pack_unpack_roundtrip(gpu, cplx, nst, np_, ld, first, pack):
    r = np.random.default_rng(nst * 1000 + np_)
    if cplx:
        u = r.standard_normal((nst, ld)) + 1j * r.standard_normal((nst, ld))
    else:
        u = r.standard_normal((nst, ld))
    nrows = np_ + 11  # np_part-like rows beyond np stay untouched
    lanes = pack * (2 if cplx else 1)
    init = r.standard_normal((nrows, lanes))
    b = api.Batch(pack, nrows, cplx, pack, ff=gpu.from_numpy(init.copy()).cuda())
    ud = gpu.from_numpy(u.copy()).cuda()
    b.nrows = np_  # pack only the np inner rows
    api.batch_pack(ud, b, first)
    gpu.cuda.synchronize()
    got = b.ff.cpu().numpy()
    ref = init.copy()
    host = workloads.pack(u[:, :np_], np_, pack) if first == 0 else None
    if cplx:
        v = ref.view(np.complex128)
        v[:np_, first:first + nst] = u[:, :np_].T
    else:
        ref[:np_, first:first + nst] = u[:, :np_].T
    assert np.array_equal(got, ref)
    if host is not None:  # the same layout as the host packer (batch_t's pack_size rows)
        lanes_used = 2 * nst if cplx else nst
        assert np.array_equal(got[:np_, :lanes_used], host[:, :lanes_used])
    out = gpu.zeros((nst, ld), dtype=gpu.complex128 if cplx else gpu.float64, device="cuda")
    out[:, np_:] = 5.0  # elements beyond np are not written
    api.batch_unpack(b, out, first)
    gpu.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.array_equal(o[:, :np_], u[:, :np_])
    assert np.all(o[:, np_:] == 5.0)


def test_pack_rejects_states_outside_the_row(gpu):
    b = api.Batch(8, 100, True, 8)
    u = gpu.zeros((4, 100), dtype=gpu.complex128, device="cuda")
    with pytest.raises(api._lib.OcthipError):
        api.batch_pack(u, b, 5)
    api.batch_pack(u[:0], b, 8)  # an empty range is a no-op, even at the end of the row
