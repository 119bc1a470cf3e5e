"""Pinning the oracle's propagation extras before they check the device:

* imaginary time (src/electrons/exponential.F90:609-619, deltat -> i deltat): for a
  diagonal H the Taylor-4 step multiplies each point by p(dt v) = sum_k (dt v)^k / k!;
* psib2/deltat2 (exponential.F90:700, 740-743): psi2 must equal a separate propagation
  of the same psi at deltat2, and psi must be unchanged by the presence of psi2;
* inh_psib (exponential.F90:705-731, 639-646): the result is the exact solution of
  d psi/dt = -i H psi + inh, psi(dt) = e^{-iH dt} psi + dt phi(-iH dt) inh with
  phi(x) = (e^x - 1)/x, up to the Taylor truncation.  Checked pointwise for a diagonal H
  (closed form) and with a dense H through scipy's expm of the augmented matrix
  [[-iH dt, dt I], [0, 0]] (its top-right block is dt phi(-iH dt)); a wrong denominator
  (k instead of k+1) or a missing first term fails both.  With psib2 the reference copies
  the propagated psi into psib2 again before the inhomogeneous series (700), so psi2 =
  e^{-iH dt} psi + dt2 phi(-iH dt2) inh: checked in the diagonal closed form;
* density (src/states/density.F90:375-417, share/opencl/density.cl:24-61): the weighted
  sum of |psi|^2 against numpy, for complex and real batches.
"""
import numpy as np
import scipy.linalg as sla

import oracle as O
from helpers import small_sphere
from octhip import stencil as S, workloads


def tiny(natoms=3):
    return small_sphere(radius=1.5, spacing=0.3, natoms=natoms, nwells=3, rc=0.9)


def diagonal_ham(s):
    """The oracle Hamiltonian with zero stencil weights and no projectors: H = diag(v)."""
    op0 = S.NLOperator(**{**s.op.__dict__, "w": np.zeros_like(s.op.w)})
    return O.Hamiltonian(op0, s.v, None)


def states(s, n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal((n, s.np)) + 1j * r.standard_normal((n, s.np))


def test_imaginary_time_diagonal_closed_form():
    s = tiny(natoms=0)
    z = states(s, 2, 3)
    psi = workloads.pack(z, s.np_part, 2)
    dt = 0.21
    diagonal_ham(s).taylor(psi, dt, 4, imag_time=True)
    got = workloads.unpack(psi, 2, s.np, True)
    a = dt * s.v  # (-i)(i dt) v
    p = 1 + a + a ** 2 / 2 + a ** 3 / 6 + a ** 4 / 24
    assert np.max(np.abs(got - p[None, :] * z)) <= 1e-14 * np.max(np.abs(z))
    # and it is the real-exponential decay e^{dt v} up to the truncation
    assert np.max(np.abs(got - np.exp(a)[None, :] * z)) <= np.max(np.abs(a)) ** 5 / 120 * 1.2 * np.max(np.abs(z))


def test_psi2_is_a_second_propagation_at_deltat2():
    s = tiny(natoms=3)
    z = states(s, 3, 4)
    psi = workloads.pack(z, s.np_part, 4)
    psi2 = np.zeros_like(psi)
    ref1, ref2 = psi.copy(), psi.copy()
    h = O.Hamiltonian(s.op, s.v, s.proj)
    h.taylor_ex(psi, 0.02, 4, psi2=psi2, deltat2=0.013)
    h.taylor(ref1, 0.02, 4)
    h.taylor(ref2, 0.013, 4)
    assert np.array_equal(psi[: s.np], ref1[: s.np])
    assert np.array_equal(psi2[: s.np], ref2[: s.np])
    assert not np.any(psi2[s.np:])  # boundary rows of psi2 are not written


def phi(x):
    return np.where(np.abs(x) < 1e-8, 1.0 + x / 2, np.expm1(x) / np.where(x == 0, 1, x))


def test_inhomogeneous_term_diagonal_closed_form():
    s = tiny(natoms=0)
    z = states(s, 2, 5)
    g = states(s, 2, 6)
    dt, dt2 = 1.1e-4, 2.3e-4  # |x| <= 1.2e-3: the truncation (x^5/120) is below rounding
    psi = workloads.pack(z, s.np_part, 2)
    inh = workloads.pack(g, s.np_part, 2)
    psi2 = np.zeros_like(psi)
    diagonal_ham(s).taylor_ex(psi, dt, 4, psi2=psi2, deltat2=dt2, inh=inh)
    x, x2 = -1j * dt * s.v, -1j * dt2 * s.v
    exact = np.exp(x)[None, :] * z + (dt * phi(x))[None, :] * g
    exact2 = np.exp(x)[None, :] * z + (dt2 * phi(x2))[None, :] * g
    scale = np.max(np.abs(z)) + np.max(np.abs(g))
    assert np.max(np.abs(workloads.unpack(psi, 2, s.np, True) - exact)) <= 1e-15 * scale
    assert np.max(np.abs(workloads.unpack(psi2, 2, s.np, True) - exact2)) <= 1e-15 * scale
    # the homogeneous-only alternative for psi2 (e^{-i dt2 v} psi) is far from it
    alt = np.exp(x2)[None, :] * z + (dt2 * phi(x2))[None, :] * g
    assert np.max(np.abs(alt - exact2)) > 1e-6 * scale


def test_inhomogeneous_term_dense_series():
    from test_oracle import dense_parts
    s = tiny(natoms=3)
    K, V, Vnl = dense_parts(s)
    H = K + V + Vnl
    z = states(s, 1, 7)[0]
    g = states(s, 1, 8)[0]
    dt = 0.001  # theta = dt E_max ~ 0.11: truncation ~1e-7 relative, a wrong denominator ~1e-5
    psi = workloads.pack(z[None, :], s.np_part, 1)
    inh = workloads.pack(g[None, :], s.np_part, 1)
    O.Hamiltonian(s.op, s.v, s.proj).taylor_ex(psi, dt, 4, inh=inh)
    got = workloads.unpack(psi, 1, s.np, True)[0]
    n = s.np
    A = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    A[:n, :n] = -1j * dt * H
    A[:n, n:] = dt * np.eye(n)
    E = sla.expm(A)
    exact = E[:n, :n] @ z + E[:n, n:] @ g
    th = dt * np.max(np.abs(np.linalg.eigvalsh(H)))
    bound = (th ** 5 / 120) * 1.5 * (np.linalg.norm(z) + dt * np.linalg.norm(g))
    assert np.linalg.norm(got - exact) <= bound + 1e-12 * np.linalg.norm(exact)
    # the same series with denominators k instead of k + 1 (a plausible slip) is caught
    wrong = z.astype(np.complex128).copy()
    e = sla.expm(-1j * dt * H)
    wrong = e @ z + dt * (e @ g)  # phi replaced by exp
    assert np.linalg.norm(wrong - exact) > 20 * bound
    # inh's boundary rows were set to zero by its H-apply (boundaries_set of the input)
    assert not np.any(inh[s.np:])


def test_density_complex_and_real():
    s = tiny(natoms=0)
    z = states(s, 3, 9)
    w = np.array([2.0, 0.5, 1.25])
    psi = workloads.pack(z, s.np_part, 4)
    rho = np.full(s.np, 0.75)
    O.density_accumulate(psi, s.np, 3, True, w, rho)
    ref = 0.75 + np.sum(w[:, None] * np.abs(z) ** 2, axis=0)
    assert np.max(np.abs(rho - ref)) <= 1e-14 * np.max(ref)
    f = z.real
    pr = workloads.pack(f, s.np_part, 4)
    rho = np.zeros(s.np)
    O.density_accumulate(pr, s.np, 3, False, w, rho)
    assert np.max(np.abs(rho - np.sum(w[:, None] * f ** 2, axis=0))) <= 1e-14 * np.max(rho)


def test_expmid_and_etrs_diagonal_closed_forms():
    """With H = diag(v) + E(t) x the propagators are pointwise phases: exponential midpoint
    e^{-i dt (v + E_mid x)}, ETRS e^{-i dt/2 (v + E_new x)} e^{-i dt/2 (v + E_old x)}, up
    to the Taylor truncation (below rounding at these dt); the density is sum w |psi|^2."""
    s = tiny(natoms=0)
    op0 = S.NLOperator(**{**s.op.__dict__, "w": np.zeros_like(s.op.w)})
    x = s.mesh.x[: s.np]
    h = O.Hamiltonian(op0, s.v.copy(), None, xyz=x, field=[0.0, 0.0, 0.0])
    z = [states(s, 2, 20), states(s, 2, 21)]
    w = [np.array([1.0, 0.5]), np.array([2.0, 0.0])]
    dt = 2e-4
    e_mid, e_old, e_new = np.array([0.3, 0.0, -0.2]), np.array([0.1, 0.2, 0.0]), np.array([-0.4, 0.1, 0.3])
    psis = [workloads.pack(zz, s.np_part, 2) for zz in z]
    rho = O.expmid_step(h, psis, dt, e_mid, weights=w)
    ph = np.exp(-1j * dt * (s.v + x @ e_mid))
    for zz, p in zip(z, psis):
        assert np.max(np.abs(workloads.unpack(p, 2, s.np, True) - ph[None, :] * zz)) <= 1e-15 * np.max(np.abs(zz))
    ref = sum(np.sum(ww[:, None] * np.abs(ph[None, :] * zz) ** 2, axis=0) for ww, zz in zip(w, z))
    assert np.max(np.abs(rho - ref)) <= 1e-14 * np.max(ref)
    psis = [workloads.pack(zz, s.np_part, 2) for zz in z]
    rho = O.etrs_step(h, psis, dt, e_old, e_new, weights=w)
    ph = np.exp(-0.5j * dt * (s.v + x @ e_new)) * np.exp(-0.5j * dt * (s.v + x @ e_old))
    for zz, p in zip(z, psis):
        assert np.max(np.abs(workloads.unpack(p, 2, s.np, True) - ph[None, :] * zz)) <= 1e-15 * np.max(np.abs(zz))
    # one field for both halves (a plausible slip) is far outside that bound
    bad = np.exp(-1j * dt * (s.v + x @ e_old))
    assert np.max(np.abs(bad - ph)) > 1e-6
