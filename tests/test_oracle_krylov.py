"""Pinning the oracle's Lanczos and Chebyshev exponentials (src/electrons/exponential.F90:
766-971) before they check the device.  Both approximate exp(-i dt H) psi, which on a tiny
mesh is known exactly from scipy's expm of the dense H (assembled independently of the
oracle's H-apply in test_oracle.dense_parts): with enough Krylov steps / Chebyshev terms
they must agree with it to near rounding; with too few they must not (so a truncated or
mis-signed expansion fails).  The Lanczos phi-function branch (inhomogeneous term) is
checked against the top-right block of expm([[-i dt H, dt I], [0, 0]]).
"""
import numpy as np
import scipy.linalg as sla

import oracle as O
from helpers import small_sphere
from octhip import workloads


def setup():
    from test_oracle import dense_parts
    s = small_sphere(radius=1.5, spacing=0.3, natoms=3, nwells=3, rc=0.9)
    K, V, Vnl = dense_parts(s)
    H = K + V + Vnl
    r = np.random.default_rng(31)
    z = r.standard_normal((3, s.np)) + 1j * r.standard_normal((3, s.np))
    return s, H, z


def test_lanczos_converges_to_expm():
    s, H, z = setup()
    dt = 0.02
    psi = workloads.pack(z, s.np_part, 4)
    it, res = O.lanczos_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, 3, dt, 40, 1e-13, s.mesh.vol_pp)
    ref = (sla.expm(-1j * dt * H) @ z.T).T
    got = workloads.unpack(psi, 3, s.np, True)
    assert res < 1e-13 and it < 40
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    # two Krylov steps are far from it (the expansion really is truncated at `order`)
    psi = workloads.pack(z, s.np_part, 4)
    O.lanczos_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, 3, dt, 2, 1e-13, s.mesh.vol_pp)
    assert np.max(np.abs(workloads.unpack(psi, 3, s.np, True) - ref)) > 1e-4 * np.max(np.abs(ref))


def test_lanczos_imaginary_time_and_inhomogeneous_term():
    s, H, z = setup()
    dt = 0.01
    # imaginary time: deltat -> i deltat, so exp(-i (i deltat) H) = exp(deltat H); the
    # reference's imaginary-time caller passes deltat = -tau (eigen_evolution_inc.F90:87)
    psi = workloads.pack(z, s.np_part, 4)
    O.lanczos_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, 3, 1j * -dt, 40, 1e-13, s.mesh.vol_pp)
    ref = (sla.expm(-dt * H) @ z.T).T
    assert np.max(np.abs(workloads.unpack(psi, 3, s.np, True) - ref)) <= 1e-11 * np.max(np.abs(ref))
    g = np.conj(z[::-1])
    psi = workloads.pack(z, s.np_part, 4)
    inh = workloads.pack(g, s.np_part, 4)
    O.lanczos_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, 3, dt, 40, 1e-13, s.mesh.vol_pp, inh=inh)
    n = s.np
    A = np.zeros((2 * n, 2 * n), np.complex128)
    A[:n, :n] = -1j * dt * H
    A[:n, n:] = dt * np.eye(n)
    E = sla.expm(A)
    ref = z + (E[:n, n:] @ g.T).T  # psi += dt phi(-i dt H) inh
    assert np.max(np.abs(workloads.unpack(psi, 3, s.np, True) - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_chebyshev_converges_to_expm():
    s, H, z = setup()
    dt = 0.02
    w = np.linalg.eigvalsh(H)
    mid, span = (w[-1] + w[0]) / 2, (w[-1] - w[0]) / 2 * 1.01
    psi = workloads.pack(z, s.np_part, 4)
    O.cheby_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, dt, 30, mid, span)
    ref = (sla.expm(-1j * dt * H) @ z.T).T
    got = workloads.unpack(psi, 3, s.np, True)
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    psi = workloads.pack(z, s.np_part, 4)
    O.cheby_batch(O.Hamiltonian(s.op, s.v, s.proj), psi, dt, 3, mid, span)
    assert np.max(np.abs(workloads.unpack(psi, 3, s.np, True) - ref)) > 1e-3 * np.max(np.abs(ref))
