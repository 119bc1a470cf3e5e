"""Periodic cells and k-points on the device against the oracle's periodic apply
(src/hamiltonian/hamiltonian_elec_inc.F90:116-154, 243-247; boundaries_set's periodic
branch, src/grid/boundaries_inc.F90:366-743; Bloch phases and projector phases,
src/hamiltonian/hamiltonian_elec.F90:1178-1300).

A 24^3 cell at h = 0.3 with Gaussian wells and three KB atoms, one of whose spheres crosses
two faces of the cell; k = (0.3, -0.2, 0.45).  Bars: <= 1e-12 relative per state for the
H-apply (both routes: a batch without the phase and one that carries it), the Gamma-point
periodic apply without phases, Taylor-4 and Lanczos steps on phased batches, and the
exponential-midpoint driver over two batches with the phase set and unset around each.
Both stencil kernels (z-march and tiled gather) are covered.
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err
from octhip import api, periodic as PER, workloads

pytestmark = pytest.mark.gpu

K = np.array([0.3, -0.2, 0.45])


@pytest.fixture(scope="module")
def cell():
    return PER.periodic_system(n=24, spacing=0.3, natoms=3, nwells=3, rc=1.5, seed=4242)


def make(s, k=K, kernel=0, phases=True):
    from octhip.stencil import star_points
    kin = api.NLOperatorDevice(s.op, star_points(4))
    hm = api.HamiltonianDevice(kin, s.v, s.proj)
    hm.vol_pp = s.mesh.vol_pp
    if kernel:
        hm.tune(kernel=kernel)
    hm.set_periodic(s.bnd_src)
    if phases:
        phase, corr = s.phases(k)
        hm.set_phases(phase, corr, s.pphase_of(k))
    return hm


def states(s, nst, seed):
    r = np.random.default_rng(seed)
    z = np.zeros((nst, s.np), np.complex128)
    x = s.mesh.x[: s.np]
    env = np.exp(-np.einsum("ij,ij->i", x, x) / 9.0) + 0.2
    z[:] = (r.standard_normal((nst, s.np)) + 1j * r.standard_normal((nst, s.np))) * env
    return z


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("has_phase", [False, True])
def test_periodic_apply(gpu, cell, kernel, has_phase):
    print(f"[periodic apply kernel={kernel} has_phase={has_phase}]", flush=True)
    s = cell
    hm = make(s, kernel=kernel)
    z = states(s, 12, 5)
    phase, corr = s.phases(K)
    if has_phase:
        z = z * phase[None, : s.np]
    psi = workloads.pack(z, s.np_part)
    psib = api.Batch.from_states(z, s.np_part)
    out = api.Batch(12, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, out, has_phase=has_phase)
    gpu.cuda.synchronize()
    ref = O.periodic_apply(O.Hamiltonian(s.op, s.v, None), psi, s.bnd_src, s.proj, phase, corr, s.pphase_of(K),
                           has_phase=has_phase)
    assert max_rel_err(out.ff.cpu().numpy(), ref) <= 1e-12
    # boundary rows were filled from the images, as boundaries_set does
    assert max_rel_err(psib.ff.cpu().numpy()[s.np:], psi[s.np:]) <= 1e-15


def test_gamma_point_periodic_without_phases(gpu, cell):
    print("[periodic gamma]", flush=True)
    s = cell
    hm = make(s, phases=False)
    z = states(s, 8, 6)
    psi = workloads.pack(z, s.np_part)
    psib = api.Batch.from_states(z, s.np_part)
    out = api.Batch(8, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, out)
    gpu.cuda.synchronize()
    ref = O.periodic_apply(O.Hamiltonian(s.op, s.v, None), psi, s.bnd_src, s.proj)
    assert max_rel_err(out.ff.cpu().numpy(), ref) <= 1e-12


def test_periodic_taylor_and_lanczos(gpu, cell):
    print("[periodic taylor + lanczos]", flush=True)
    s = cell
    hm = make(s)
    phase, corr = s.phases(K)
    pp = s.pphase_of(K)
    z = states(s, 8, 7) * phase[None, : s.np]  # batches inside the propagation carry the phase
    oh = O.Hamiltonian(s.op, s.v, None)

    def apply(p):
        return O.periodic_apply(oh, p.copy(), s.bnd_src, s.proj, phase, corr, pp, has_phase=True)

    psib = api.Batch.from_states(z, s.np_part)
    api.Exponential(4).apply_batch(hm, psib, 0.01)
    gpu.cuda.synchronize()
    ref = workloads.pack(z, s.np_part)
    O.periodic_taylor(apply, ref, s.np, 0.01)
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    # Lanczos with the same H
    psib = api.Batch.from_states(z, s.np_part)
    te = api.Exponential(12, method="lanczos", lanczos_tol=1e-8)
    te.apply_batch(hm, psib, 0.01)
    gpu.cuda.synchronize()

    class H:
        np, np_part = s.np, s.np_part

        @staticmethod
        def apply(p):
            return apply(p)

    ref = workloads.pack(z, s.np_part)
    it, _ = O.lanczos_batch(H, ref, 8, 0.01, 12, 1e-8, s.mesh.vol_pp)
    assert te.last_iterations == it
    assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12


def test_periodic_expmid_driver(gpu, cell):
    print("[periodic expmid]", flush=True)
    s = cell
    hm = make(s)
    x = s.mesh.x[: s.np]
    hm.set_dipole_field([0.0, 0.0, 0.0], positions=x)
    phase, corr = s.phases(K)
    pp = s.pphase_of(K)
    zs = [states(s, 8, 10), states(s, 8, 11)]
    batches = [api.Batch.from_states(z, s.np_part) for z in zs]
    w = [np.full(8, 2.0), np.full(8, 1.0)]
    rho = gpu.zeros((s.np,), dtype=gpu.float64, device="cuda")
    field = np.array([0.02, 0.0, -0.01])
    api.propagator_expmid(api.Exponential(4), hm, batches, 0.01, field, w, rho)
    gpu.cuda.synchronize()
    oh = O.Hamiltonian(s.op, s.v.copy(), None, xyz=x, field=field)
    rho_ref = np.zeros(s.np)
    for z, b, ww in zip(zs, batches, w):
        ref = workloads.pack(z * phase[None, : s.np], s.np_part)  # set_phase_corr
        O.periodic_taylor(lambda p: O.periodic_apply(oh, p.copy(), s.bnd_src, s.proj, phase, corr, pp,
                                                     has_phase=True), ref, s.np, 0.01)
        zz = ref.view(np.complex128)
        zz[: s.np] *= np.conj(phase[: s.np, None])  # unset_phase_corr
        O.density_accumulate(ref, s.np, 8, True, ww, rho_ref)
        assert max_rel_err(b.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
    assert float(np.max(np.abs(rho.cpu().numpy() - rho_ref)) / np.max(rho_ref)) <= 1e-12


def test_periodic_handles_refuse_fused_paths(gpu, cell):
    s = cell
    hm = make(s)
    b = api.Batch.from_states(states(s, 4, 12), s.np_part)
    with pytest.raises(api._lib.OcthipError):
        api.TaylorGraph(api.Exponential(4), hm, [b], 0.01)


@pytest.mark.parametrize("kernel", [0, 1])
def test_gamma_point_real_states(gpu, cell, kernel):
    """Real states at the Gamma point: boundary rows copied from their images (no phase),
    then the plain H-apply on the filled rows (the oracle's apply with set_bc off)."""
    print(f"[periodic real states kernel={kernel}]", flush=True)
    s = cell
    hm = make(s, kernel=kernel, phases=False)
    z = np.ascontiguousarray(states(s, 8, 13).real)
    psib = api.Batch.from_states(z, s.np_part)
    out = api.Batch(8, s.np, False)
    api.hamiltonian_elec_apply_batch(hm, psib, out)
    gpu.cuda.synchronize()
    ref_in = workloads.pack(z, s.np_part)
    ref_in[s.np:] = ref_in[s.bnd_src]
    ref = O.Hamiltonian(s.op, s.v, s.proj).apply(ref_in.copy(), set_bc=False)
    assert max_rel_err(out.ff.cpu().numpy(), ref) <= 1e-12
    assert np.array_equal(psib.ff.cpu().numpy()[s.np:], ref_in[s.np:])


def test_boundaries_cannot_change_under_a_graph(gpu):
    """A Taylor graph keeps its zero-boundary launches: changing the handle's boundaries or
    phases while it exists is refused, and allowed again once it is destroyed."""
    import gc
    s = PER.periodic_system(n=16, spacing=0.3, natoms=1, nwells=2, rc=1.2, seed=77)
    from octhip.stencil import star_points
    hm = api.HamiltonianDevice(api.NLOperatorDevice(s.op, star_points(4)), s.v, s.proj)
    b = api.Batch.from_states(states(s, 4, 14), s.np_part)
    g = api.TaylorGraph(api.Exponential(4), hm, [b], 0.01)
    with pytest.raises(api._lib.OcthipError, match="Taylor graphs"):
        hm.set_periodic(s.bnd_src)
    phase, corr = s.phases(K)
    with pytest.raises(api._lib.OcthipError, match="Taylor graphs"):
        hm.set_phases(phase, corr, s.pphase_of(K))
    del g
    gc.collect()
    hm.set_periodic(s.bnd_src)
    hm.set_phases(phase, corr, s.pphase_of(K))
