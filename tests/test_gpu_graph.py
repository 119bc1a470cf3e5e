"""Taylor steps captured into a HIP graph, and the dipole-field potential.

* A graph replay must give bit-identical states to the same steps launched one by one
  (same kernels, same arguments), on every replay, with unrelated device work and host
  synchronisations between replays.
* With a dipole field set before capture, changing the field between replays must be
  seen by the graph (the potential buffer is updated in place).
* The dipole-field H-apply is checked against the oracle's hamiltonian (<= 1e-12 per
  state: the projector sums differ in order).
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, small_sphere
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def _batches(s, nst_list, seed=3):
    out = []
    for i, nst in enumerate(nst_list):
        z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=2.0, first=seed + 16 * i),
                                     s.mesh.vol_pp)
        out.append(api.Batch.from_states(z, s.np_part, pack_size=8))
    return out


@pytest.mark.parametrize("kernel,lpt", [(1, 0), (2, 2), (2, 1)])
def test_taylor_graph_replays_bitwise(gpu, kernel, lpt):
    print(f"[taylor graph kernel={kernel} lpt={lpt}]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.3, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(kernel=kernel, lanes_per_thread=lpt)
    te = api.Exponential(4)
    g_b = _batches(s, [8, 8, 5])
    e_b = [b.clone() for b in g_b]
    graph = api.TaylorGraph(te, hm, g_b, 0.02)
    other = _batches(s, [8], seed=11)[0]
    hother = api.Batch(8, s.np, True, pack_size=8)
    for rep in range(4):
        graph.launch()
        for b in e_b:
            te.apply_batch(hm, b, 0.02)
        # unrelated eager work on the same stream and a host sync between replays
        api.hamiltonian_elec_apply_batch(hm, other, hother)
        hother.ff.mul_(0.5)
        gpu.cuda.synchronize()
        for gb, eb in zip(g_b, e_b):
            assert gpu.equal(gb.ff[: s.np], eb.ff[: s.np]), f"replay {rep} differs from eager steps"
    # and the eager path itself against the oracle after the 4 steps
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, 8, width=2.0, first=3), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    for _ in range(4):
        oh.taylor(psi, 0.02)
    assert max_rel_err(g_b[0].ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-11


@pytest.mark.parametrize("order,imag", [(1, False), (2, True), (3, False), (6, False), (4, True)])
def test_taylor_graph_other_orders_and_imaginary_time(gpu, order, imag):
    """Graphs of other Taylor orders (the deferred accumulation covers orders 2-4, order 1
    and orders above 4 accumulate order by order) and in imaginary time replay bitwise as
    the eager steps, which match the oracle."""
    print(f"[taylor graph order={order} imag={imag}]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    te = api.Exponential(order)
    dt = -0.01 if imag else 0.01
    g_b = _batches(s, [8, 4])
    e_b = [b.clone() for b in g_b]
    graph = api.TaylorGraph(te, hm, g_b, dt, imag_time=imag)
    for rep in range(3):
        graph.launch()
        for b in e_b:
            te.apply_batch(hm, b, dt, imag_time=imag)
        gpu.cuda.synchronize()
        for gb, eb in zip(g_b, e_b):
            assert gpu.equal(gb.ff[: s.np], eb.ff[: s.np]), f"order {order}: replay {rep} differs"
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, 8, width=2.0, first=3), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    for _ in range(3):
        oh.taylor(psi, dt, order, imag)
    assert max_rel_err(g_b[0].ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-11


def test_taylor_graph_sees_field_updates(gpu):
    print("[taylor graph field]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.3, natoms=3, rc=1.2)
    xyz = s.mesh.x[: s.np]
    hm = api.HamiltonianDevice.from_system(s)
    te = api.Exponential(4)
    g_b = _batches(s, [8, 8])
    e_b = [b.clone() for b in g_b]
    hm.set_dipole_field([0.0, 0.0, 0.0], positions=xyz)
    graph = api.TaylorGraph(te, hm, g_b, 0.02)
    oh = O.Hamiltonian(s.op, s.v, s.proj, xyz=xyz, field=[0.0, 0.0, 0.0])
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, 8, width=2.0, first=3), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    for step in range(5):
        f = [0.05 * np.sin(0.7 * step + 0.3), -0.02 * step, 0.01]
        hm.set_dipole_field(f)
        oh.set_field(f)
        graph.launch()
        for b in e_b:
            te.apply_batch(hm, b, 0.02)
        oh.taylor(psi, 0.02)
        gpu.cuda.synchronize()
        for gb, eb in zip(g_b, e_b):
            assert gpu.equal(gb.ff[: s.np], eb.ff[: s.np]), f"step {step}: graph differs from eager"
    assert max_rel_err(g_b[0].ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-11


def test_graph_rejects_mixed_batches(gpu):
    s = small_sphere(radius=3.0, spacing=0.3)
    hm = api.HamiltonianDevice.from_system(s)
    b8 = _batches(s, [8])[0]
    b2 = api.Batch.from_states(workloads.gaussian_envelope_states(s, 2, width=2.0), s.np_part)
    with pytest.raises(ValueError):
        api.TaylorGraph(api.Exponential(4), hm, [b8, b2], 0.02)


@pytest.mark.parametrize("bra", [0, 1])
def test_hamiltonian_dipole_field(gpu, bra):
    print(f"[H-apply dipole field bra={bra}]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.3, natoms=3, rc=1.2)
    xyz = s.mesh.x[: s.np]
    f = [0.03, -0.011, 0.007]
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=bra)
    hm.set_dipole_field(f, positions=xyz)
    z = workloads.gaussian_envelope_states(s, 8, width=2.0)
    psib = api.Batch.from_states(z, s.np_part)
    hpsib = api.Batch(8, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib)
    gpu.cuda.synchronize()
    ref = O.Hamiltonian(s.op, s.v, s.proj, xyz=xyz, field=f).apply(workloads.pack(z, s.np_part))
    assert max_rel_err(hpsib.ff.cpu().numpy(), ref) <= 1e-12
    # local part alone is bit-identical (same fused multiply-add order)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib, terms=api.TERM_LOCAL_POTENTIAL)
    gpu.cuda.synchronize()
    ref = O.Hamiltonian(s.op, s.v, s.proj, xyz=xyz, field=f).apply(workloads.pack(z, s.np_part),
                                                                    terms=O.TERM_LOCAL)
    assert np.array_equal(hpsib.ff.cpu().numpy(), ref)


def test_taylor_graph_create_destroy_frees_work_sets(gpu):
    """Each Taylor graph owns the work sets it was captured with and frees them when it is
    destroyed: rebuilding graphs (a new dt, new batches) does not grow device memory."""
    import gc
    print("[taylor graph create/destroy memory]", flush=True)
    s = small_sphere(radius=5.0, spacing=0.25, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    te = api.Exponential(4)
    bs = [api.Batch.from_states(workloads.gaussian_envelope_states(s, 16, width=2.0, first=16 * i), s.np_part)
          for i in range(4)]
    # first cycle builds the operator's z-march plans, which live as long as the operator
    g = api.TaylorGraph(te, hm, bs, 0.01)
    g.launch()
    del g
    gc.collect()
    gpu.cuda.synchronize()
    free0, _ = gpu.cuda.mem_get_info()
    for dt in (0.011, 0.012, 0.013, 0.014):
        g = api.TaylorGraph(te, hm, bs, dt)
        g.launch()
        gpu.cuda.synchronize()
        del g
        gc.collect()
    gpu.cuda.synchronize()
    free1, _ = gpu.cuda.mem_get_info()
    # four branches, each with three Taylor buffers of np rows of 32 doubles: far above 16 MB
    set_bytes = 4 * 3 * s.np * 32 * 8
    assert set_bytes > 16 << 20
    assert free0 - free1 < 16 << 20, f"device memory grew by {(free0 - free1) / 2**20:.1f} MB"
    # a graph outliving its Hamiltonian: destroying the Hamiltonian frees the sets, the
    # graph then destroys cleanly
    g = api.TaylorGraph(te, hm, bs, 0.01)
    hm_handle = hm.handle
    hm.handle = None
    from octhip import _lib
    _lib.call("octhip_hamiltonian_destroy", hm_handle)
    del g
    gc.collect()


def test_memory_report_and_trim(gpu):
    """octhip_hamiltonian_memory names what a handle holds: growing the row length
    retires the smaller work buffers, a graph's sets are reported as the graph's, trim
    frees the retired ones and the handle still computes the same H psi afterwards."""
    import gc
    print("[hamiltonian memory / trim]", flush=True)
    s = small_sphere(radius=4.0, spacing=0.25, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    m0 = hm.memory()
    assert m0["data"] > 0 and m0["work"] == 0 and m0["graphs"] == 0 and m0["retired"] == 0
    small = api.Batch.from_states(workloads.gaussian_envelope_states(s, 4, width=2.0), s.np_part)
    big = api.Batch.from_states(workloads.gaussian_envelope_states(s, 32, width=2.0), s.np_part)
    out_s, out_b = api.Batch(4, s.np, True), api.Batch(32, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, small, out_s)
    api.Exponential(4).apply_batch(hm, small.clone(), 0.01)
    m1 = hm.memory()
    assert m1["work"] > 0 and m1["retired"] == 0
    api.hamiltonian_elec_apply_batch(hm, big, out_b)
    api.Exponential(4).apply_batch(hm, big.clone(), 0.01)
    m2 = hm.memory()
    assert m2["retired"] > 0 and m2["work"] > m1["work"]
    g = api.TaylorGraph(api.Exponential(4), hm, [big.clone()], 0.01)
    m3 = hm.memory()
    assert m3["graphs"] > 0
    del g
    gc.collect()
    gpu.cuda.synchronize()
    assert hm.memory()["graphs"] == 0
    ref_b = out_b.ff.clone()
    hm.trim()
    m4 = hm.memory()
    assert m4["retired"] == 0 and m4["work"] == m2["work"] and m4["data"] == m0["data"]
    api.hamiltonian_elec_apply_batch(hm, big, out_b)
    api.hamiltonian_elec_apply_batch(hm, small, out_s)
    gpu.cuda.synchronize()
    assert gpu.equal(out_b.ff, ref_b)
    ref = O.Hamiltonian(s.op, s.v, s.proj).apply(workloads.pack(workloads.gaussian_envelope_states(s, 4, width=2.0),
                                                                s.np_part))
    assert max_rel_err(out_s.ff.cpu().numpy(), ref) <= 1e-12
