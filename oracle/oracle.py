"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker
and by bench.py's cpu_baseline leg.  The product path (liboctopus_hip.so and the octhip
package) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

TERM_KINETIC, TERM_LOCAL, TERM_NON_LOCAL, TERM_ALL = 1, 2, 4, 0x7FFFFFFF


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_lib = None
# ORACLE_LIB: another build of the same source (the sanitizer build of the memory-safety test)
_path = Path(os.environ["ORACLE_LIB"]) if os.environ.get("ORACLE_LIB") else LIB
FAST = HERE / "_fast" / "liboracle_fast.so"


def use_timing_build() -> str:
    """Switch to the timing build of the same source (-O3 -march=native, built here for this
    host by `make fast`; its results equal the parity build's).  Returns the flags used, or
    the parity build's when the timing build cannot be made."""
    global _lib, _path
    r = subprocess.run(["make", "-s", "-C", str(HERE), "fast"], capture_output=True, text=True)
    if r.returncode == 0 and FAST.exists():
        _lib, _path = None, FAST
        return "-O3 -march=native -fopenmp"
    return "-O3 -march=x86-64-v3 -fopenmp (timing build failed: " + (r.stderr.strip()[-200:] or "?") + ")"


def lib():
    global _lib
    if _lib is None:
        if not _path.exists():
            build()
        l = C.CDLL(str(_path))
        l.oracle_taylor.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int,
                                    C.c_int]
        l.oracle_taylor_ex.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
        l.oracle_density_accumulate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        l.oracle_daxpy.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        l.oracle_hamiltonian_apply.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.c_int, C.c_int, C.c_int]
        _lib = l
    return _lib


def _p(a, t=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(C.POINTER(t))


class OracleHam(C.Structure):
    _fields_ = [("np", C.c_int), ("np_part", C.c_int), ("n", C.c_int), ("nri", C.c_int),
                ("w", C.POINTER(C.c_double)), ("ri", C.POINTER(C.c_int)),
                ("rimap_inv", C.POINTER(C.c_int)), ("mass", C.c_double),
                ("v", C.POINTER(C.c_double)), ("xyz", C.POINTER(C.c_double)),
                ("field", C.POINTER(C.c_double)), ("nmat", C.c_int),
                ("offsets", C.POINTER(C.c_int)), ("matrices", C.POINTER(C.c_double)),
                ("maps", C.POINTER(C.c_int)), ("scals", C.POINTER(C.c_double))]


class Hamiltonian:
    """Keeps the numpy arrays alive for as long as the C struct points at them."""

    def __init__(self, op, v=None, proj=None, mass=1.0, xyz=None, field=None):
        self._keep = dict(
            w=np.ascontiguousarray(op.w, np.float64), ri=np.ascontiguousarray(op.ri, np.int32).ravel(),
            rinv=np.ascontiguousarray(op.rimap_inv, np.int32),
            v=None if v is None else np.ascontiguousarray(v, np.float64),
            xyz=None if xyz is None else np.ascontiguousarray(xyz, np.float64).ravel(),
            field=None if field is None else np.ascontiguousarray(field, np.float64))
        k = self._keep
        nmat = 0
        if proj is not None and proj.nmat > 0:
            k.update(offs=np.ascontiguousarray(proj.offsets, np.int32).ravel(),
                     mats=np.ascontiguousarray(proj.matrices, np.float64),
                     maps=np.ascontiguousarray(proj.maps, np.int32),
                     scals=np.ascontiguousarray(proj.scals, np.float64))
            nmat = proj.nmat
        else:
            k.update(offs=None, mats=None, maps=None, scals=None)
        self.np, self.np_part = op.np, op.np_part
        self.s = OracleHam(op.np, op.np_part, op.n, op.nri, _p(k["w"]), _p(k["ri"], C.c_int),
                           _p(k["rinv"], C.c_int), mass, _p(k["v"]), _p(k["xyz"]), _p(k["field"]),
                           nmat, _p(k["offs"], C.c_int), _p(k["mats"]), _p(k["maps"], C.c_int),
                           _p(k["scals"]))

    def set_potential(self, v):
        """Overwrite the local potential in place (the C struct keeps pointing at it)."""
        self._keep["v"][:] = np.asarray(v, np.float64)

    def set_field(self, field):
        self._keep["field"] = None if field is None else np.ascontiguousarray(field, np.float64)
        self.s.field = _p(self._keep["field"])

    def apply(self, psi: np.ndarray, terms=TERM_ALL, set_bc=True) -> np.ndarray:
        """psi: (np_part, L) packed float64 (modified if set_bc); returns (np, L)."""
        L = psi.shape[1]
        ldf = int(np.log2(L))
        out = np.zeros((self.np, L))
        rc = lib().oracle_hamiltonian_apply(C.byref(self.s), _p(psi), _p(out), ldf, int(terms), int(set_bc))
        if rc:
            raise RuntimeError("oracle_hamiltonian_apply failed")
        return out

    def taylor(self, psi: np.ndarray, deltat: float, order=4, imag_time=False) -> None:
        L = psi.shape[1]
        rc = lib().oracle_taylor(C.byref(self.s), _p(psi), int(np.log2(L)), C.c_double(deltat), order,
                                 int(imag_time))
        if rc:
            raise RuntimeError("oracle_taylor failed")

    def taylor_ex(self, psi: np.ndarray, deltat: float, order=4, imag_time=False, psi2=None, deltat2=0.0,
                  inh=None) -> None:
        """exponential_apply_batch with psib2/deltat2 and inh_psib (all (np_part, L), modified
        in place; inh's boundary rows are zeroed by its H-apply, as in the reference)."""
        L = psi.shape[1]
        for a in (psi2, inh):
            assert a is None or (a.shape == psi.shape and a.flags["C_CONTIGUOUS"])
        rc = lib().oracle_taylor_ex(C.byref(self.s), _p(psi), int(np.log2(L)), C.c_double(deltat), order,
                                    int(imag_time), _p(psi2), C.c_double(deltat2), _p(inh))
        if rc:
            raise RuntimeError("oracle_taylor_ex failed")


def density_accumulate(psi: np.ndarray, rows: int, nst: int, is_complex: bool, weights, density: np.ndarray) -> None:
    """density[:rows] += sum_ist weights[ist] |psi(ist, p)|^2 (the device kernels' order)."""
    w = np.ascontiguousarray(weights, np.float64)
    assert density.flags["C_CONTIGUOUS"] and density.dtype == np.float64 and len(w) >= nst
    lib().oracle_density_accumulate(int(rows), int(np.log2(psi.shape[1])), int(nst), int(bool(is_complex)),
                                    _p(w), _p(np.ascontiguousarray(psi)), _p(density))


def exp_apply_batches(h: Hamiltonian, psis: list, dt: float, order: int = 4, weights=None):
    """propagation_ops_elec_exp_apply / _fuse_density_exp_apply for independent particles
    (src/td/propagation_ops_elec.F90:213-333): each packed complex batch propagated by dt in
    turn; with weights (one array per batch, pack-size long), the density of the results
    (density_calc_init zeroes it, density_calc_accumulate adds each batch)."""
    rho = np.zeros(h.np) if weights is not None else None
    for b, psi in enumerate(psis):
        h.taylor(psi, dt, order)
        if rho is not None:
            w = np.asarray(weights[b], np.float64)
            density_accumulate(psi, h.np, len(w), True, w, rho)
    return rho


def expmid_step(h: Hamiltonian, psis: list, dt: float, field_mid=None, order: int = 4, weights=None):
    """exponential_midpoint (src/td/propagator_expmid.F90:52-100), independent particles:
    H at t - dt/2 (the dipole field there), then every batch propagated by dt."""
    if field_mid is not None:
        h.set_field(field_mid)
    return exp_apply_batches(h, psis, dt, order, weights)


def etrs_step(h: Hamiltonian, psis: list, dt: float, field_old=None, field_new=None, order: int = 4,
              weights=None):
    """td_etrs (src/td/propagator_etrs.F90:68-150), independent particles: dt/2 with
    H(t - dt), H updated to t, dt/2 with the density."""
    if field_old is not None:
        h.set_field(field_old)
    exp_apply_batches(h, psis, 0.5 * dt, order)
    if field_new is not None:
        h.set_field(field_new)
    return exp_apply_batches(h, psis, 0.5 * dt, order, weights)


def operate_ri_vec(op, fi: np.ndarray, factor: float = 1.0) -> np.ndarray:
    """fo (np, L) = factor*L fi with Octopus's run-compressed index."""
    L = fi.shape[1]
    w = np.ascontiguousarray(op.w * factor, np.float64)
    ri = np.ascontiguousarray(op.ri, np.int32).ravel()
    rinv = np.ascontiguousarray(op.rimap_inv, np.int32)
    fo = np.zeros((op.np, L))
    fi = np.ascontiguousarray(fi, np.float64)
    lib().oracle_operate_ri_vec(op.n, _p(w), op.nri, _p(ri, C.c_int), _p(rinv, C.c_int),
                                _p(rinv[1:], C.c_int), _p(fi), int(np.log2(L)), _p(fo))
    return fo


def nlocal_start(proj, psi: np.ndarray) -> np.ndarray:
    L = psi.shape[1]
    nps = int((proj.offsets[:, 4] + proj.offsets[:, 1]).max())
    out = np.zeros((nps, L))
    offs = np.ascontiguousarray(proj.offsets, np.int32).ravel()
    lib().oracle_nlocal_start(proj.nmat, _p(offs, C.c_int), _p(proj.matrices), _p(proj.maps, C.c_int),
                              _p(proj.scals), int(np.log2(L)), _p(np.ascontiguousarray(psi)), _p(out))
    return out



# ---------------------------------------------------------------- adjoint --
def to_complex(packed: np.ndarray, rows: int) -> np.ndarray:
    """(rows, pack) complex view of a packed (>= rows, 2*pack) batch (copy)."""
    return np.ascontiguousarray(packed[:rows]).view(np.complex128).copy()


def to_packed(z: np.ndarray, np_part: int) -> np.ndarray:
    out = np.zeros((np_part, 2 * z.shape[1]))
    out[: z.shape[0]] = np.ascontiguousarray(z).view(np.float64)
    return out


class TaylorAdjoint:
    """Exact gradient of the discrete Taylor propagator under a dipole field.

    psi_{s+1} = sum_{n<=order} zf_n H_s^n psi_s, zf_n = (-i dt)^n / n!,
    H_s = H + eps_s (pol . r) (laser_potential sign, src/hamiltonian/lasers.F90:823-853).
    For dJ = 2 Re <chi_N|d psi_N> (<a|b> = vol sum conj(a) b):
      chi_s = sum_m conj(zf_m) H_s^m chi_{s+1}       (p(A)^dagger, H Hermitian)
      rho_s(r) = sum_states sum_{m+k<=order-1} zf_{m+k+1} conj(H^m chi_{s+1}) H^k psi_s
      dJ/dv(r) = 2 vol Re sum_s rho_s(r),  dJ/deps_s = 2 vol sum_r (pol . r) Re rho_s(r).
    The reference's optimal-control gradient (src/opt_control/propagation.F90:1016-1103,
    2 Im <chi|x pol|psi>) is the first-order-in-dt limit of dJ/deps_s / dt.  This
    restatement keeps every psi_s (no checkpoints): it is the checker.
    """

    def __init__(self, ham: Hamiltonian, xyz: np.ndarray, pol, dt: float, order: int = 4, vol: float = 1.0):
        self.h, self.dt, self.order, self.vol = ham, float(dt), int(order), float(vol)
        self.pol = np.asarray(pol, np.float64)
        self.xp = np.asarray(xyz, np.float64)[: ham.np] @ self.pol
        self.zf = [1.0 + 0j]
        for n in range(1, self.order + 1):
            self.zf.append(self.zf[-1] * (-1j * self.dt) / n)

    def _powers(self, z: np.ndarray, count: int) -> list:
        """[z, H z, ..., H^(count-1) z] on complex (np, pack) arrays."""
        out = [z]
        for _ in range(count - 1):
            out.append(to_complex(self.h.apply(to_packed(out[-1], self.h.np_part)), self.h.np))
        return out

    def _set(self, e: float) -> None:
        self.h.set_field(e * self.pol)

    def forward(self, psi0: np.ndarray, eps) -> np.ndarray:
        """psi0: packed (np_part, L); returns psi_N packed; keeps psi_s for backward."""
        self.eps = np.asarray(eps, np.float64)
        z = to_complex(psi0, self.h.np)
        self.psis = []
        for e in self.eps:
            self.psis.append(z)
            self._set(e)
            y = self._powers(z, self.order + 1)
            z = sum(self.zf[n] * y[n] for n in range(self.order + 1))
        return to_packed(z, self.h.np_part)

    def backward(self, chiN: np.ndarray):
        """chiN: packed (np_part, L). Returns (chi_0 packed, dJ/deps (nsteps,), dJ/dv (np,))."""
        chi = to_complex(chiN, self.h.np)
        dv = np.zeros(self.h.np)
        deps = np.zeros(len(self.eps))
        for s in range(len(self.eps) - 1, -1, -1):
            self._set(self.eps[s])
            yp = self._powers(self.psis[s], self.order)
            yc = self._powers(chi, self.order + 1)
            rho = np.zeros(self.h.np)
            for m in range(self.order):
                u = sum(self.zf[m + k + 1] * yp[k] for k in range(self.order - m))
                rho += np.sum((np.conj(yc[m]) * u).real, axis=1)
            dv += 2.0 * self.vol * rho
            deps[s] = 2.0 * self.vol * np.dot(self.xp, rho)
            chi = sum(np.conj(self.zf[m]) * yc[m] for m in range(self.order + 1))
        return to_packed(chi, self.h.np_part), deps, dv


def overlap_objective(phi: np.ndarray, psi: np.ndarray, rows: int, vol: float):
    """J = sum_j |<phi_j|psi_j>|^2 and chi_j = <phi_j|psi_j> phi_j (dJ = 2 Re <chi|dpsi>)."""
    a, b = to_complex(phi, rows), to_complex(psi, rows)
    c = vol * np.sum(np.conj(a) * b, axis=0)
    return float(np.sum(np.abs(c) ** 2)), to_packed(a * c[None, :], phi.shape[0])


# ------------------------------------------------------- Krylov exponentials --
def mesh_dotp(a: np.ndarray, b: np.ndarray, rows: int, nst: int, vol: float) -> np.ndarray:
    """<a_st|b_st> = vol sum_p conj(a) b for packed complex batches (mesh_batch_dotp_vector,
    src/grid/mesh_batch_inc.F90:379-549)."""
    za, zb = to_complex(a, rows)[:, :nst], to_complex(b, rows)[:, :nst]
    return vol * np.sum(np.conj(za) * zb, axis=0)


def _cols_axpy(y: np.ndarray, x: np.ndarray, coef, rows: int) -> None:
    """y(:, st) += coef[st] x(:, st) on the inner rows of packed complex batches (states beyond
    len(coef) untouched, batch_axpy with a_full = .false.)."""
    zy = y[:rows].view(np.complex128)
    zx = x[:rows].view(np.complex128)
    n = len(coef)
    zy[:, :n] += np.asarray(coef)[None, :] * zx[:, :n]


def _cols_scal(x: np.ndarray, coef, rows: int) -> None:
    zx = x[:rows].view(np.complex128)
    n = len(coef)
    zx[:, :n] = np.asarray(coef)[None, :] * zx[:, :n]


def _hermitian_function(A: np.ndarray, pp: complex, phi: bool) -> np.ndarray:
    """zlalg_exp / zlalg_phi, Hermitian branch (src/math/lalg_adv.F90:216-340): eigenpairs of
    the matrix read from its upper triangle (zheev 'U'), f(pp A) = V f(pp D) V^H."""
    w, V = np.linalg.eigh(A, UPLO="U")
    x = pp * w
    f = (np.exp(x) - 1.0) / x if phi else np.exp(x)
    return (V * f[None, :]) @ np.conj(V).T


def lanczos_batch(h: "Hamiltonian", psi: np.ndarray, nst: int, deltat: complex, order: int, tol: float,
                  vol: float, inh: np.ndarray | None = None):
    """exponential_lanczos_batch (src/electrons/exponential.F90:766-900) on a packed complex
    batch, in place; Hermitian H (Lanczos: orthogonalisation against the last two vectors,
    classical Gram-Schmidt).  Returns (Krylov dimension used, largest error estimate)."""
    rows, L = h.np, psi.shape[1]
    eps = np.finfo(np.float64).eps
    vb = [np.zeros((h.np_part, L)) for _ in range(order + 1)]
    vb[0][:rows] = (inh if inh is not None else psi)[:rows]
    beta = np.sqrt(mesh_dotp(vb[0], vb[0], rows, nst, vol).real)
    if np.all(np.abs(beta) <= 1e-12):
        return 0, 0.0
    _cols_scal(vb[0], 1.0 / beta, rows)
    M = order + 1
    hamilt = np.zeros((nst, M, M), np.complex128)
    expo = np.zeros((nst, M), np.complex128)
    pp = -1j * deltat
    res = np.zeros(nst)
    it = 1
    while it <= order:
        vb[it][:rows] = h.apply(vb[it - 1])
        l = max(1, it - 1)
        ss = [mesh_dotp(vb[j - 1], vb[it], rows, nst, vol) for j in range(l, it + 1)]
        for j, o in zip(range(l, it + 1), ss):
            hamilt[:, j - 1, it - 1] = o
        for j, o in zip(range(l, it + 1), ss):
            _cols_axpy(vb[it], vb[j - 1], -o, rows)
        nrm = np.sqrt(mesh_dotp(vb[it], vb[it], rows, nst, vol).real)
        hamilt[:, it, it - 1] = nrm
        for st in range(nst):
            expo[st, :it] = _hermitian_function(hamilt[st, :it, :it], pp, inh is not None)[:, 0]
            res[st] = abs(hamilt[st, it, it - 1] * abs(expo[st, it - 1]))
        if np.all(np.abs(hamilt[:, it, it - 1]) < 1e4 * eps):
            break
        sc = np.where(np.abs(hamilt[:, it, it - 1]) >= 1e4 * eps, 1.0 / np.abs(hamilt[:, it, it - 1]), 1.0)
        _cols_scal(vb[it], sc, rows)
        if it > 3 and np.all(res < tol):
            break
        it += 1
    used = min(it, order)
    if inh is not None:
        for ii in range(1, used + 1):
            _cols_axpy(psi, vb[ii - 1], (np.real(deltat) * beta) * expo[:, ii - 1], rows)
    else:
        _cols_scal(psi, expo[:, 0], rows)
        for ii in range(2, used + 1):
            _cols_axpy(psi, vb[ii - 1], beta * expo[:, ii - 1], rows)
    return used, float(np.max(res))


def cheby_batch(h: "Hamiltonian", psi: np.ndarray, deltat: float, order: int, mid: float, span: float) -> None:
    """exponential_cheby_batch (src/electrons/exponential.F90:903-971), in place: Clenshaw's
    recurrence for sum_j c_j T_j((H - mid)/span), c_j = 2 (-i)^j J_j(span deltat)."""
    from scipy.special import jv
    rows, L = h.np, psi.shape[1]
    psi0 = np.zeros((h.np_part, L))
    psi1 = np.zeros((h.np_part, L))
    psi2 = np.zeros((h.np_part, L))
    zb = to_complex(psi, rows)
    for j in range(order - 1, -1, -1):
        psi2[:rows] = psi1[:rows]
        psi1[:rows] = psi0[:rows]
        psi0[:rows] = h.apply(psi1)
        zfact = 2.0 * (-1j) ** j * jv(j, span * deltat)
        z0 = psi0[:rows].view(np.complex128)
        z0 += -mid * psi1[:rows].view(np.complex128)
        z0 *= 2.0 / span
        z0 += zfact * zb
        z0 += -1.0 * psi2[:rows].view(np.complex128)
    out = psi0[:rows].view(np.complex128) - psi2[:rows].view(np.complex128)
    psi[:rows] = (0.5 * np.exp(-1j * mid * deltat) * out).view(np.float64)


# ------------------------------------------------------- periodic / k-points --
def periodic_apply(h: "Hamiltonian", psi: np.ndarray, bnd_src: np.ndarray, proj=None, phase=None, phase_corr=None,
                   pphase=None, has_phase: bool = False) -> np.ndarray:
    """X(hamiltonian_elec_apply_batch) on a periodic cell, optionally with the Bloch phases of
    one k-point (src/hamiltonian/hamiltonian_elec_inc.F90:116-154, 243-247).  h: the oracle
    Hamiltonian of the kinetic and local terms (no projectors); proj: the projector set,
    applied here with its Bloch phases (hamiltonian_elec_base_inc.F90:683-692 bra,
    955-966 ket).  psi: packed complex (np_part, L), boundary rows set in place
    (boundaries_set: periodic images, times phase_corr when psi carries the phase).
    Returns H psi (np, L)."""
    npt = h.np
    z = psi.view(np.complex128)
    if has_phase:
        z[npt:] = phase_corr[:, None] * z[bnd_src]
    else:
        z[npt:] = z[bnd_src]
    set_phase = phase is not None and not has_phase
    e = (phase[:, None] * z) if set_phase else z.copy()
    ep = np.ascontiguousarray(e).view(np.float64)
    out = h.apply(ep.copy(), TERM_KINETIC | TERM_LOCAL, set_bc=False).view(np.complex128)
    if proj is not None and proj.nmat > 0:
        off = 0
        for a in range(proj.nmat):
            P, mp, sc = proj.matrix(a)
            idx = mp - 1
            ph = pphase[off: off + len(idx)] if pphase is not None else np.ones(len(idx))
            lpsi = e[idx] * ph[:, None]
            pr = sc[:, None] * (P.T @ lpsi)
            out[idx] += (P @ pr) * np.conj(ph)[:, None]
            off += len(idx)
    if set_phase:
        out = np.conj(phase[:npt])[:, None] * out
    return np.ascontiguousarray(out).view(np.float64)


def periodic_taylor(apply, psi: np.ndarray, rows: int, deltat: float, order: int = 4) -> None:
    """exponential_taylor_series_batch (src/electrons/exponential.F90:677-757) with a given
    H-apply (apply(psi_np_part) -> H psi (rows, L)), in place on the inner rows."""
    L = psi.shape[1]
    z = 1.0 + 0j
    psi1 = np.zeros_like(psi)
    for k in range(1, order + 1):
        z = z * (-1j * deltat) / k
        hp = apply(psi if k == 1 else psi1)
        zp = psi[:rows].view(np.complex128)
        zp += z * hp.view(np.complex128)
        if k != order:
            psi1[:rows] = hp
