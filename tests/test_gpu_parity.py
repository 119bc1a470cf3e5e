"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars (the tolerance of each check is written in it):
* stencil and local potential: bit-identical -- both sides sum the 25 terms in stencil
  order with fused multiply-adds;
* anything with the KB projectors: max relative error <= 1e-12 per state (max|x-y|/max|y|),
  since the device bra sums sphere points in a different (blocked / MFMA) order;
* Taylor runs: per-state error and norm drift <= 1e-10 against the oracle's propagator.
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, random_packed, small_sphere
from octhip import api, workloads
from octhip.stencil import star_points

pytestmark = pytest.mark.gpu


def to_dev(gpu, a):
    return gpu.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_operate(gpu, op_dev, fin, factor):
    L = fin.shape[1]
    fi = api.Batch(L, fin.shape[0], False, ff=to_dev(gpu, fin))
    fo = api.Batch(L, op_dev.np, False)
    api.nl_operator_operate_batch(op_dev, fi, fo, factor)
    gpu.cuda.synchronize()
    return fo.ff.cpu().numpy()


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("kernel", ["ri", "gather", "zmarch"])
def test_stencil_bit_exact_sphere(gpu, lanes, kernel):
    print(f"[stencil sphere lanes={lanes} kernel={kernel}]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.25)
    op = api.NLOperatorDevice(s.op, None if kernel == "ri" else star_points(4))
    assert op.uses_lattice == (kernel != "ri")
    if kernel != "ri":
        if kernel == "gather":
            op.tune(kernel=1, gather_tile=3 if lanes == 8 else 4)
        else:
            op.tune(kernel=2, zseg=5 if lanes == 8 else 16, lanes_per_thread=1 if lanes in (4, 64) else 2)
    f = random_packed(s.np_part, lanes, seed=lanes)
    got = run_operate(gpu, op, f, -0.5)
    ref = O.operate_ri_vec(s.op, f, -0.5)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("knobs", [dict(kernel=1, xcd_chunk=1), dict(kernel=1, xcd_chunk=64, gather_tile=2),
                                   dict(kernel=2), dict(kernel=2, zseg=7, lane_chunk=4),
                                   dict(kernel=2, lanes_per_thread=1)])
def test_stencil_bit_exact_cube_workload(gpu, knobs):
    """The 64^3 cube workload (8 Gaussian sums, pack 8) at full size; kernels and
    schedules only reorder work, so every choice is bit-identical."""
    print(f"[stencil cube64 {knobs}]", flush=True)
    s = workloads.laplacian_system()
    f = workloads.gaussian_functions(s, 8)
    packed = workloads.pack(f, s.np_part, 8)
    op = api.NLOperatorDevice(s.op, star_points(4))
    op.tune(**knobs)
    got = run_operate(gpu, op, packed, 1.0)
    ref = O.operate_ri_vec(s.op, packed, 1.0)
    assert np.array_equal(got, ref)


def run_h(gpu, s, z, terms=api.TERM_ALL, set_bc=True, lattice=True, bra=1, pack_size=None, kernel=0):
    hm = api.HamiltonianDevice.from_system(s, lattice=lattice)
    hm.tune(bra_mode=bra)
    if lattice:
        hm.tune(kernel=kernel)
    psib = api.Batch.from_states(z, s.np_part, pack_size)
    hpsib = api.Batch(psib.nst, s.np, psib.is_complex, psib.pack_size)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib, terms, set_bc)
    gpu.cuda.synchronize()
    return hpsib.ff.cpu().numpy()


def oracle_h(s, z, terms=O.TERM_ALL, set_bc=True, pack_size=None):
    psi = workloads.pack(z, s.np_part, pack_size)
    return O.Hamiltonian(s.op, s.v, s.proj).apply(psi, terms, set_bc)


def states(s, nst, seed=0):
    return workloads.gaussian_envelope_states(s, nst, width=2.0, first=seed)


@pytest.mark.parametrize("terms", [api.TERM_KINETIC, api.TERM_LOCAL_POTENTIAL,
                                   api.TERM_KINETIC | api.TERM_LOCAL_POTENTIAL])
@pytest.mark.parametrize("kernel", [1, 2])
def test_hamiltonian_local_terms_bit_exact(gpu, terms, kernel):
    print(f"[hamiltonian terms={terms} kernel={kernel}]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.25, natoms=4)
    z = states(s, 16)
    assert np.array_equal(run_h(gpu, s, z, terms, kernel=kernel), oracle_h(s, z, terms))


@pytest.mark.parametrize("bra", [0, 1])
@pytest.mark.parametrize("nst", [1, 3, 16, 32, 64])
def test_hamiltonian_full(gpu, bra, nst):
    print(f"[hamiltonian full bra={bra} nst={nst}]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.25, natoms=5, rc=1.5)
    z = states(s, nst)
    got, ref = run_h(gpu, s, z, bra=bra), oracle_h(s, z)
    assert max_rel_err(got, ref) <= 1e-12


def test_hamiltonian_real_batch(gpu):
    print("[hamiltonian real]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.25, natoms=5, rc=1.5)
    z = states(s, 8).real.copy()
    assert max_rel_err(run_h(gpu, s, z), oracle_h(s, z)) <= 1e-12


def test_hamiltonian_boundary_rows_as_given(gpu):
    """set_bc = .false.: boundary rows are read as the caller left them."""
    print("[hamiltonian set_bc=0]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.25, natoms=3)
    z = states(s, 8)
    psi = workloads.pack(z, s.np_part)
    psi[s.np:] = np.random.default_rng(1).standard_normal(psi[s.np:].shape)
    hm = api.HamiltonianDevice.from_system(s)
    psib = api.Batch(8, s.np_part, True, ff=to_dev(gpu, psi))
    hpsib = api.Batch(8, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib, api.TERM_ALL, set_bc=False)
    gpu.cuda.synchronize()
    ref = O.Hamiltonian(s.op, s.v, s.proj).apply(psi.copy(), O.TERM_ALL, set_bc=False)
    assert max_rel_err(hpsib.ff.cpu().numpy(), ref) <= 1e-12
    # set_bc = .true. zeroes psi's boundary rows, like boundaries_set
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib, api.TERM_ALL, set_bc=True)
    gpu.cuda.synchronize()
    assert not np.any(psib.ff.cpu().numpy()[s.np:])


def test_hamiltonian_no_projectors(gpu):
    print("[hamiltonian nmat=0]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.25, natoms=0)
    z = states(s, 4)
    assert np.array_equal(run_h(gpu, s, z), oracle_h(s, z))


def test_hamiltonian_workload_batch(gpu):
    """One 32-state batch of the R=12, h=0.25 sphere workload (462,781 points), through
    both apply kernels; the two kernels agree bit for bit."""
    print("[hamiltonian sphere R=12 batch]", flush=True)
    s = workloads.hamiltonian_system()
    z = workloads.gaussian_envelope_states(s, 32)
    ref = oracle_h(s, z)
    got_z, got_g = run_h(gpu, s, z, kernel=2), run_h(gpu, s, z, kernel=1)
    assert max_rel_err(got_z, ref) <= 1e-12
    assert np.array_equal(got_z, got_g)


def test_taylor_step_workload_batch(gpu):
    """One Taylor-4 step (dt 0.02) of a 32-state batch of the R=12, h=0.25 sphere workload
    with the production launch shapes (z-march Taylor epilogue, automatic segments)."""
    print("[taylor sphere R=12 batch]", flush=True)
    s = workloads.hamiltonian_system()
    z = workloads.gaussian_envelope_states(s, 32)
    psi = workloads.pack(z, s.np_part)
    O.Hamiltonian(s.op, s.v, s.proj).taylor(psi, 0.02)
    hm = api.HamiltonianDevice.from_system(s)
    psib = api.Batch.from_states(z, s.np_part)
    api.Exponential(4).apply_batch(hm, psib, 0.02)
    gpu.cuda.synchronize()
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("imag_time", [False, True])
@pytest.mark.parametrize("kernel,lpt", [(1, 0), (2, 2), (2, 1)])
def test_taylor_step(gpu, order, imag_time, kernel, lpt):
    print(f"[taylor order={order} imag={imag_time} kernel={kernel} lpt={lpt}]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    z = workloads.orthonormalise(states(s, 8), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    O.Hamiltonian(s.op, s.v, s.proj).taylor(psi, 0.01, order, imag_time)
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(kernel=kernel, lanes_per_thread=lpt)
    psib = api.Batch.from_states(z, s.np_part)
    api.Exponential(order).apply_batch(hm, psib, 0.01, imag_time)
    gpu.cuda.synchronize()
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-12


def test_taylor_run_drift(gpu):
    """50 Taylor-4 steps at a stable dt: state error and norm drift <= 1e-10."""
    print("[taylor 50 steps]", flush=True)
    s = small_sphere(radius=4.0, spacing=0.3, natoms=4, rc=1.5)
    z = workloads.orthonormalise(states(s, 16), s.mesh.vol_pp)
    psi = workloads.pack(z, s.np_part)
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    hm = api.HamiltonianDevice.from_system(s)
    psib = api.Batch.from_states(z, s.np_part)
    te = api.Exponential(4)
    for _ in range(50):
        oh.taylor(psi, 0.02)
        te.apply_batch(hm, psib, 0.02)
    gpu.cuda.synchronize()
    got = workloads.unpack(psib.ff.cpu().numpy(), 16, s.np, True)
    ref = workloads.unpack(psi, 16, s.np, True)
    err = np.max(np.abs(got - ref), axis=1) / np.max(np.abs(ref), axis=1)
    assert np.max(err) <= 1e-10
    ng, nr = np.sum(np.abs(got) ** 2, axis=1), np.sum(np.abs(ref) ** 2, axis=1)
    assert np.max(np.abs(ng - nr) / nr) <= 1e-10


@pytest.mark.parametrize("ldf", [0, 1, 2, 3, 4, 5, 6])
def test_operate_ri_vec_stateless_bit_exact(gpu, ldf):
    """octhip_operate_ri_vec, the device counterpart of doperate_ri_vec, on Octopus's own
    run-compressed index (ri, rimap_inv, rimap_inv_max as device buffers): bit-identical
    to the oracle leaf, runs split in two calls as two OpenMP threads would split them."""
    import ctypes as C
    from octhip import _lib
    print(f"[operate_ri_vec ldf={ldf}]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.25)
    op = s.op
    L = 1 << ldf
    f = random_packed(s.np_part, L, seed=100 + ldf)
    w = np.ascontiguousarray(op.w * -0.5, np.float64)
    ri = to_dev(gpu, np.ascontiguousarray(op.ri, np.int32).ravel())
    rinv = np.ascontiguousarray(op.rimap_inv, np.int32)
    imin, imax = to_dev(gpu, rinv[:-1].copy()), to_dev(gpu, rinv[1:].copy())
    fi = to_dev(gpu, f)
    fo = gpu.full((s.np, L), np.nan, dtype=gpu.float64, device="cuda")
    half = op.nri // 2
    stream = C.c_void_p(gpu.cuda.current_stream().cuda_stream)
    for l0, l1 in ((0, half), (half, op.nri)):
        _lib.call("octhip_operate_ri_vec", op.n, _lib.as_ptr(w, C.c_double), l1 - l0,
                  C.c_void_p(ri.data_ptr() + 4 * op.n * l0), C.c_void_p(imin.data_ptr() + 4 * l0),
                  C.c_void_p(imax.data_ptr() + 4 * l0), C.c_void_p(fi.data_ptr()), ldf,
                  C.c_void_p(fo.data_ptr()), stream)
    gpu.cuda.synchronize()
    ref = O.operate_ri_vec(op, f, -0.5)
    assert np.array_equal(fo.cpu().numpy(), ref)


@pytest.mark.parametrize("kernel", [1, 2])
def test_widest_rows(gpu, kernel):
    """512 complex states in one batch (1024 lanes, the longest row the library takes): the
    H-apply and one Taylor-4 step against the oracle, through the gather and the z-march
    kernels (64 lane chunks of 16)."""
    print(f"[widest rows kernel={kernel}]", flush=True)
    s = small_sphere(radius=2.2, spacing=0.3, natoms=2, rc=1.0)
    z = states(s, 512)
    got, ref = run_h(gpu, s, z, kernel=kernel), oracle_h(s, z)
    assert got.shape[1] == 1024
    assert max_rel_err(got, ref) <= 1e-12
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(kernel=kernel)
    psib = api.Batch.from_states(z, s.np_part)
    api.Exponential(4).apply_batch(hm, psib, 0.01)
    gpu.cuda.synchronize()
    refp = workloads.pack(z, s.np_part)
    O.Hamiltonian(s.op, s.v, s.proj).taylor(refp, 0.01, 4)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], refp[: s.np]) <= 1e-12
