"""Device allocation failures, injected through the library's allocator hook
(octhip_inject_alloc_failure): each operation runs with its k-th allocation failing, for
k = 0, 1, ... until it succeeds.  Every failed call must return OCTHIP_ERR_ALLOC with the
injected message (the reference turns such an error into messages_fatal), must leave no
device memory behind once its objects are destroyed (or trimmed), and the handles must
compute the right result afterwards.  Nothing is launched on a buffer that failed to
allocate: every call site checks the status before it uses the buffer.
"""
import ctypes as C
import gc

import pytest

import oracle as O
from helpers import max_rel_err, small_sphere
from octhip import _lib, api, workloads
from octhip.stencil import star_points

pytestmark = pytest.mark.gpu


def until_success(fn, cleanup=lambda: None, limit=200):
    """Run fn with the k-th allocation failing, k = 0, 1, ...; returns (k, fn's result)."""
    for k in range(limit):
        _lib.call("octhip_inject_alloc_failure", k)
        try:
            out = fn()
        except _lib.OcthipError as e:
            assert e.code == _lib.ERR_ALLOC and "injected" in str(e), str(e)
            out = None
            ok = False
        else:
            ok = True
        finally:
            _lib.call("octhip_inject_alloc_failure", -1)
        if ok:
            return k, out
        cleanup()
    raise AssertionError(f"still failing after {limit} allocations")


def free_bytes(gpu):
    """Free device memory with torch's cached blocks returned (batches are torch tensors)."""
    gc.collect()
    gpu.cuda.synchronize()
    gpu.cuda.empty_cache()
    return gpu.cuda.mem_get_info()[0]


def test_injected_failures_leave_nothing_behind(gpu):
    print("[allocation failures]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    z = workloads.gaussian_envelope_states(s, 8, width=2.0)
    ref_h = O.Hamiltonian(s.op, s.v, s.proj)
    psi_host = workloads.pack(z, s.np_part)
    ref_apply = ref_h.apply(psi_host.copy())
    ref_taylor = psi_host.copy()
    ref_h.taylor(ref_taylor, 0.01, 4)

    def every_operation(hm):
        b = api.Batch.from_states(z, s.np_part)
        api.hamiltonian_elec_apply_batch(hm, b, api.Batch(8, s.np, True))
        api.Exponential(4).apply_batch(hm, b, 0.01)
        g = api.TaylorGraph(api.Exponential(4), hm, [b], 0.01)
        api.Exponential(8, method="lanczos", lanczos_tol=1e-8).apply_batch(hm, b, 0.01)
        bp = api.BatchedPropagator([(hm, api.Batch.from_states(z, s.np_part))], 0.01)
        adj = api.Adjoint(hm, s.mesh.x[: s.np], [1.0, 0.0, 0.0], b.ldf, 6, 0.01, s.mesh.vol_pp,
                          checkpoint_every=3)
        api.mesh_batch_dotp_vector(s.mesh.vol_pp, b, b, s.np)
        gpu.cuda.synchronize()
        del g, bp, adj

    # one untroubled pass first: the runtime's own first-use allocations (code objects of
    # kernels loaded on first launch, stream and graph bookkeeping) are not the library's
    warm = api.HamiltonianDevice.from_system(s)
    every_operation(warm)
    del warm
    free0 = free_bytes(gpu)

    # operator and Hamiltonian creation (uploads of the index, projector and march tables)
    n_op, kin = until_success(lambda: api.NLOperatorDevice(s.op, star_points(4)))
    assert n_op >= 2
    n_h, hm = until_success(lambda: api.HamiltonianDevice(kin, s.v, s.proj))
    assert n_h >= 2
    hm.vol_pp = float(s.mesh.vol_pp)

    # first H-apply (the stream's projector work buffers) and first Taylor step (its
    # intermediates): the failed calls leave partial sets, retired on the next growth
    psib = api.Batch.from_states(z, s.np_part)
    out = api.Batch(8, s.np, True)
    n_a, _ = until_success(lambda: api.hamiltonian_elec_apply_batch(hm, psib, out))
    gpu.cuda.synchronize()
    assert max_rel_err(out.ff.cpu().numpy(), ref_apply) <= 1e-12

    def taylor():
        b = api.Batch.from_states(z, s.np_part)
        api.Exponential(4).apply_batch(hm, b, 0.01)
        return b
    n_t, b = until_success(taylor)
    gpu.cuda.synchronize()
    assert max_rel_err(b.ff.cpu().numpy()[: s.np], ref_taylor[: s.np]) <= 1e-12

    # a Taylor graph (its own work sets), a Lanczos step (Krylov vectors), a batched
    # launch (per-pair buffers and descriptors), an adjoint (checkpoints), a dot product
    bs = [api.Batch.from_states(z, s.np_part)]
    until_success(lambda: api.TaylorGraph(api.Exponential(4), hm, bs, 0.01))
    until_success(lambda: api.Exponential(8, method="lanczos", lanczos_tol=1e-8).apply_batch(
        hm, api.Batch.from_states(z, s.np_part), 0.01))
    until_success(lambda: api.BatchedPropagator([(hm, api.Batch.from_states(z, s.np_part))], 0.01))
    until_success(lambda: api.Adjoint(hm, s.mesh.x[: s.np], [1.0, 0.0, 0.0], psib.ldf, 6, 0.01,
                                      s.mesh.vol_pp, checkpoint_every=3))
    until_success(lambda: api.mesh_batch_dotp_vector(s.mesh.vol_pp, psib, psib, s.np))
    gc.collect()

    # everything still computes the same
    api.hamiltonian_elec_apply_batch(hm, psib, out)
    gpu.cuda.synchronize()
    assert max_rel_err(out.ff.cpu().numpy(), ref_apply) <= 1e-12

    # once the handles are gone (retired buffers with them), the device memory is back
    del hm, kin, psib, out, b, bs
    gc.collect()
    free1 = free_bytes(gpu)
    assert free0 - free1 < 8 << 20, f"{(free0 - free1) / 2**20:.1f} MB not returned"


def test_hook_fails_the_next_allocation_and_turns_off(gpu):
    p = C.c_void_p()
    _lib.call("octhip_inject_alloc_failure", 1)
    _lib.call("octhip_malloc", C.byref(p), 64)        # the one allowed allocation
    with pytest.raises(_lib.OcthipError, match="injected"):
        _lib.call("octhip_malloc", C.byref(C.c_void_p()), 64)
    _lib.call("octhip_inject_alloc_failure", -1)
    q = C.c_void_p()
    _lib.call("octhip_malloc", C.byref(q), 64)
    _lib.call("octhip_free", p)
    _lib.call("octhip_free", q)
