"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker
and by bench.py's cpu_baseline leg.  The product path (liboctopus_hip.so and the octhip
package) never imports it.
"""
from __future__ import annotations

import ctypes as C
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


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.oracle_taylor.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int,
                                    C.c_int]
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

