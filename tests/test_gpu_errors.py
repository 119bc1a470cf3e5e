"""Error paths of the C ABI on a device: invalid input is refused on the host with a
status and a message (OcthipError), as the reference turns device errors into
messages_fatal; nothing is launched for it."""
import ctypes as C

import numpy as np
import pytest

from helpers import small_sphere
from octhip import _lib, api, workloads
from octhip.stencil import star_points

pytestmark = pytest.mark.gpu


def test_operator_create_rejects_bad_index(gpu):
    s = small_sphere(radius=2.5, spacing=0.3)
    op = s.op
    h = C.c_void_p()
    w = np.ascontiguousarray(op.w, np.float64)
    ri = np.ascontiguousarray(op.ri, np.int32).ravel().copy()
    rinv = np.ascontiguousarray(op.rimap_inv, np.int32).copy()
    ri[0] = 10 ** 8  # neighbour far outside np_part
    with pytest.raises(_lib.OcthipError, match="reaches outside"):
        _lib.call("octhip_nl_operator_create", op.n, _lib.as_ptr(w, C.c_double), None, op.nri,
                  _lib.as_ptr(ri, C.c_int), _lib.as_ptr(rinv, C.c_int), op.np, op.np_part, None, C.byref(h))
    ri = np.ascontiguousarray(op.ri, np.int32).ravel().copy()
    rinv[-1] = op.np + 1  # runs must end at np
    with pytest.raises(_lib.OcthipError, match="rimap_inv"):
        _lib.call("octhip_nl_operator_create", op.n, _lib.as_ptr(w, C.c_double), None, op.nri,
                  _lib.as_ptr(ri, C.c_int), _lib.as_ptr(rinv, C.c_int), op.np, op.np_part, None, C.byref(h))


def test_apply_rejects_bad_arguments(gpu):
    s = small_sphere(radius=2.5, spacing=0.3)
    op = api.NLOperatorDevice(s.op, star_points(4))
    fi = api.Batch(4, s.np_part, False)
    fo = api.Batch(4, s.np, False)
    with pytest.raises(_lib.OcthipError, match="OP_INNER, OP_OUTER, OP_ALL"):
        api.nl_operator_operate_batch(op, fi, fo, points=0)
    with pytest.raises(_lib.OcthipError, match="in place"):
        _lib.call("octhip_nl_operator_operate_batch", op.handle, fi.ptr(), fi.ptr(), fi.ldf, 1.0, _lib.OP_ALL,
                  None)
    hm = api.HamiltonianDevice.from_system(s)
    with pytest.raises(_lib.OcthipError, match="ldf"):
        _lib.call("octhip_hamiltonian_apply_batch", hm.handle, fi.ptr(), fo.ptr(), 11, 0, _lib.TERM_ALL, 1, None)
    with pytest.raises(_lib.OcthipError, match="complex"):
        _lib.call("octhip_exponential_taylor_batch", hm.handle, fi.ptr(), 0, 0.01, 4, 0, None)
    with pytest.raises(_lib.OcthipError, match="order"):
        _lib.call("octhip_exponential_taylor_batch", hm.handle, fi.ptr(), 2, 0.01, 0, 0, None)


def test_tuning_rejects_oversized_lds_and_bad_values(gpu):
    s = small_sphere(radius=2.5, spacing=0.3)
    hm = api.HamiltonianDevice.from_system(s)
    with pytest.raises(_lib.OcthipError):
        hm.tune(lane_chunk=3)
    with pytest.raises(_lib.OcthipError):
        hm.tune(zseg=65)
    hm.tune(kernel=2, lane_chunk=32, zseg=64)  # accepted, but the launch needs > 160 KB of LDS
    z = workloads.gaussian_envelope_states(s, 32, width=2.0)
    psib = api.Batch.from_states(z, s.np_part)
    hpsib = api.Batch(32, s.np, True)
    with pytest.raises(_lib.OcthipError, match="LDS"):
        api.hamiltonian_elec_apply_batch(hm, psib, hpsib)
    hm.tune(zseg=0)  # automatic again: the same call works
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib)
    gpu.cuda.synchronize()


def test_release_stream_rejects_null_handle(gpu):
    with pytest.raises(_lib.OcthipError):
        _lib.call("octhip_hamiltonian_release_stream", None, None)


def test_adjoint_backward_needs_forward(gpu):
    s = small_sphere(radius=2.5, spacing=0.3)
    hm = api.HamiltonianDevice.from_system(s)
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], 4, 3, 0.01, s.mesh.vol_pp)
    chi = api.Batch(8, s.np_part, True)
    with pytest.raises(_lib.OcthipError, match="before adjoint_forward"):
        adj.backward(chi)
