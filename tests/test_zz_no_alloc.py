"""octhip_mesh_batch_dotp_vector_noalloc (no allocation: its reduction scratch is a fixed
device array of the library) against octhip_mesh_batch_dotp_vector (which allocates its
scratch per call), through the C ABI: same kernels and partition, so the results must be
bit-identical, on two calls in a row with the second one smaller (the array keeps the
first call's contents), for complex and real rows, and the size limit must be enforced."""
import ctypes as C

import numpy as np
import pytest

from octhip import _lib, api, workloads

pytestmark = pytest.mark.gpu


def both(gpu, a, b, ldf, np_rows, cplx, vol):
    L = 1 << ldf
    n = L // 2 if cplx else L
    host = np.zeros(2 * n if cplx else n)
    _lib.call("octhip_mesh_batch_dotp_vector", C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), ldf, np_rows,
              int(cplx), vol, host.ctypes.data_as(C.POINTER(C.c_double)), None)
    dev = gpu.full((2 * n,), np.nan, dtype=gpu.float64, device="cuda")
    _lib.call("octhip_mesh_batch_dotp_vector_noalloc", C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), ldf,
              np_rows, int(cplx), vol, C.c_void_p(dev.data_ptr()), None)
    gpu.cuda.synchronize()
    d = dev.cpu().numpy()
    if not cplx:
        assert np.all(d[1::2] == 0.0)
        d = d[0::2]
    return host, d


def test_noalloc_dotp_matches_allocating_entry_point(gpu):
    print("[no-alloc dot product]", flush=True)
    r = np.random.default_rng(5)
    # first call: 64-lane complex rows over 300k points; second: 16 lanes over 1000 points
    for (rows, lanes, cplx) in ((300_000, 64, True), (1000, 16, True), (200_000, 1024, False), (777, 8, False)):
        a = gpu.from_numpy(r.standard_normal((rows, lanes))).cuda()
        b = gpu.from_numpy(r.standard_normal((rows, lanes))).cuda()
        ldf = int(np.log2(lanes))
        host, dev = both(gpu, a, b, ldf, rows, cplx, 0.015625)
        assert np.array_equal(host, dev), (rows, lanes, cplx)
        # and the values are the mesh inner products
        A, B = a.cpu().numpy(), b.cpu().numpy()
        if cplx:
            ref = 0.015625 * np.sum(np.conj(A.view(np.complex128)) * B.view(np.complex128), axis=0)
            got = host.view(np.complex128)
        else:
            ref, got = 0.015625 * np.sum(A * B, axis=0), host
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_noalloc_rejects_bad_arguments(gpu):
    a = gpu.zeros((10, 4), dtype=gpu.float64, device="cuda")
    dev = gpu.zeros((8,), dtype=gpu.float64, device="cuda")
    with pytest.raises(_lib.OcthipError):  # ldf beyond the largest row the scratch holds
        _lib.call("octhip_mesh_batch_dotp_vector_noalloc", C.c_void_p(a.data_ptr()), C.c_void_p(a.data_ptr()), 11,
                  10, 0, 1.0, C.c_void_p(dev.data_ptr()), None)
    with pytest.raises(_lib.OcthipError):
        _lib.call("octhip_mesh_batch_dotp_vector_noalloc", None, C.c_void_p(a.data_ptr()), 2, 10, 0, 1.0,
                  C.c_void_p(dev.data_ptr()), None)
