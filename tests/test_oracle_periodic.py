"""Pinning the oracle's periodic apply (Bloch phases of one k-point, periodic boundaries:
src/hamiltonian/hamiltonian_elec_inc.F90:116-154, 243-247) before it checks the device.

* plane waves: with v = 0 and no projectors, u = exp(i G.r) (G a reciprocal lattice
  vector) is an eigenfunction of the periodic H_k: H u = (1/2) sum_d s(G_d - k_d) u with
  s(q) = -(c_0 + 2 sum_j c_j cos(j q h)) / h^2, the symbol of the order-4 star;
* the two routes agree: H on a batch that carries the phase (boundary rows times
  phase_corr) equals phase x (H on the batch without it), on the inner points;
* H_k with wells and KB atoms (one sphere crossing the cell boundary) is Hermitian, and
  moving an atom by a lattice vector leaves H_k unchanged (its projector phases change by
  exp(-i k.R) in the bra and back in the ket); dropping the conjugate in the ket (a
  plausible slip) breaks the Hermiticity.
"""
import numpy as np

import oracle as O
from octhip import periodic as PER, workloads


def small(natoms=2, n=10, h=0.4, edge=True):
    return PER.periodic_system(n=n, spacing=h, natoms=natoms, nwells=2, rc=1.2, seed=77, edge_atom=edge)


def dense(s, k, proj=True, pphase_of=None):
    phase, corr = s.phases(k)
    pp = (pphase_of or s.pphase_of)(k) if (proj and s.proj is not None) else None
    h = O.Hamiltonian(s.op, s.v, None)
    n = s.np
    H = np.zeros((n, n), np.complex128)
    for c0 in range(0, n, 64):
        cols = min(64, n - c0)
        z = np.zeros((s.np_part, 64), np.complex128)  # rows of a power-of-two length
        z[np.arange(c0, c0 + cols), np.arange(cols)] = 1.0
        out = O.periodic_apply(h, z.view(np.float64), s.bnd_src, s.proj if proj else None, phase, corr, pp)
        H[:, c0:c0 + cols] = out.view(np.complex128)[:, :cols]
    return H


def test_plane_waves_are_eigenfunctions():
    s = small(natoms=0)
    n, h = s.n, s.mesh.spacing
    Lc = n * h
    k = np.array([0.21, -0.37, 0.5])
    x = s.mesh.x[: s.np]
    c = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]

    def sym(q):
        return -(c[0] + 2 * sum(c[j] * np.cos(j * q * h) for j in range(1, 5))) / h ** 2

    op0 = O.Hamiltonian(s.op, np.zeros(s.np), None)
    phase, corr = s.phases(k)
    for G in ([0, 0, 0], [1, 0, 0], [2, -1, 3]):
        Gv = 2 * np.pi / Lc * np.asarray(G, np.float64)
        u = np.zeros((s.np_part, 1), np.complex128)
        u[: s.np, 0] = np.exp(1j * (x @ Gv))
        out = O.periodic_apply(op0, u.view(np.float64), s.bnd_src, None, phase, corr).view(np.complex128)[:, 0]
        lam = 0.5 * sum(sym(Gv[d] - k[d]) for d in range(3))
        assert np.max(np.abs(out - lam * u[: s.np, 0])) <= 1e-12 * abs(lam)


def test_routes_with_and_without_the_phase_agree():
    s = small()
    k = np.array([0.3, 0.1, -0.2])
    phase, corr = s.phases(k)
    pp = s.pphase_of(k)
    r = np.random.default_rng(3)
    z = np.zeros((s.np_part, 3), np.complex128)
    z[: s.np] = r.standard_normal((s.np, 3)) + 1j * r.standard_normal((s.np, 3))
    h = O.Hamiltonian(s.op, s.v, None)
    a = O.periodic_apply(h, z.copy().view(np.float64), s.bnd_src, s.proj, phase, corr, pp).view(np.complex128)
    zp = z.copy()
    zp[: s.np] *= phase[: s.np, None]
    b = O.periodic_apply(h, zp.view(np.float64), s.bnd_src, s.proj, phase, corr, pp, has_phase=True)
    b = b.view(np.complex128)
    assert np.max(np.abs(b - phase[: s.np, None] * a)) <= 1e-12 * np.max(np.abs(a))


def test_hk_is_hermitian_and_translation_invariant():
    s = small()
    k = np.array([0.4, -0.25, 0.1])
    H = dense(s, k)
    assert np.max(np.abs(H - H.conj().T)) <= 1e-12 * np.max(np.abs(H))
    # the same atoms moved by lattice vectors: identical H_k
    Lc = s.n * s.mesh.spacing
    from octhip import rng
    natoms = 2
    alphas = rng.uniform(s.seed, 3, 2 * natoms, 0.6, 1.2).reshape(natoms, 2)
    energies = rng.uniform(s.seed, 4, 2 * natoms, -1.0, 1.0).reshape(natoms, 2)
    proj, diffs = PER.periodic_kb_projectors(s, s.proj.centers + np.array([[Lc, 0, 0], [0, -Lc, Lc]]), 1.2,
                                             alphas, energies)
    assert np.array_equal(proj.maps, s.proj.maps)
    s2 = PER.PeriodicSystem(mesh=s.mesh, op=s.op, n=s.n, v=s.v, proj=proj, pphase_of=None, bnd_src=s.bnd_src,
                            seed=s.seed)
    H2 = dense(s2, k, pphase_of=lambda kk: np.exp(-1j * (diffs @ kk)))
    assert np.max(np.abs(H2 - H)) <= 1e-12 * np.max(np.abs(H))
    # without the conjugate in the ket, H_k is not Hermitian
    phase, corr = s.phases(k)
    pp = s.pphase_of(k)
    P, mp, sc = s.proj.matrix(0)
    e = np.zeros(s.np, np.complex128)
    e[mp[0] - 1] = 1.0
    ket_wrong = (P @ (sc * (P.T @ (pp[: len(mp)] * e[mp - 1])))) * pp[: len(mp)]
    ket_right = (P @ (sc * (P.T @ (pp[: len(mp)] * e[mp - 1])))) * np.conj(pp[: len(mp)])
    assert np.max(np.abs(ket_wrong - ket_right)) > 1e-6 * np.max(np.abs(ket_right))
