"""Host-side mirror of the reference interfaces of the H-apply path.

Names and argument meaning follow Octopus:

* ``Batch``                        -- batch_t in BATCH_DEVICE_PACKED state
                                      (src/grid/batch.F90:47-116)
* ``nl_operator_operate_batch``    -- X(nl_operator_operate_batch)(op, fi, fo, ..., points,
                                      factor) (src/grid/nl_operator_inc.F90:21-28)
* ``hamiltonian_elec_apply_batch`` -- X(hamiltonian_elec_apply_batch)(hm, ..., psib, hpsib,
                                      terms, set_bc) (src/hamiltonian/hamiltonian_elec_inc.F90:77-84)
* ``Exponential.apply_batch``      -- exponential_apply_batch(te, ..., psib, deltat, ...,
                                      imag_time) with TDExponentialMethod = taylor
                                      (src/electrons/exponential.F90:581-592)

Device memory and streams come from torch (plumbing only); every computation goes
through liboctopus_hip.so.  Errors raise OcthipError, as Octopus turns device errors into
messages_fatal.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (OP_ALL, TERM_ALL, TERM_KINETIC, TERM_LOCAL_POTENTIAL,  # noqa: F401
                   TERM_NON_LOCAL_POTENTIAL, as_ptr, call)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the device path needs a GPU (torch.cuda.is_available() is False)")
    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def pad_pow2(n: int) -> int:
    """pad_pow2 (src/math/math.F90:754-779)."""
    return 1 << max(0, int(np.ceil(np.log2(max(n, 1)))))


class Batch:
    """Packed batch on the device: ff (np_rows x 2^ldf) float64, rows = mesh points."""

    def __init__(self, nst: int, nrows: int, is_complex: bool, pack_size: int | None = None, ff=None):
        torch = _torch()
        self.nst = nst
        self.is_complex = bool(is_complex)
        self.pack_size = pack_size or pad_pow2(nst)
        self.lanes = self.pack_size * (2 if self.is_complex else 1)
        self.ldf = int(np.log2(self.lanes))
        self.nrows = nrows
        if ff is None:
            ff = torch.zeros((nrows, self.lanes), dtype=torch.float64, device="cuda")
        assert ff.dtype == torch.float64 and ff.is_cuda and ff.is_contiguous()
        assert ff.shape[0] >= nrows and ff.shape[1] == self.lanes
        self.ff = ff

    @classmethod
    def from_states(cls, states: np.ndarray, nrows: int, pack_size: int | None = None) -> "Batch":
        torch = _torch()
        from .workloads import pack
        host = pack(states, nrows, pack_size)
        b = cls(states.shape[0], nrows, np.iscomplexobj(states), pack_size,
                ff=torch.from_numpy(np.ascontiguousarray(host)).to("cuda"))
        return b

    def to_states(self, npt: int) -> np.ndarray:
        from .workloads import unpack
        return unpack(self.ff.cpu().numpy(), self.nst, npt, self.is_complex)

    def ptr(self):
        return C.c_void_p(self.ff.data_ptr())

    def clone(self) -> "Batch":
        return Batch(self.nst, self.nrows, self.is_complex, self.pack_size, ff=self.ff.clone())


class NLOperatorDevice:
    """nl_operator_t on the device, built from the host arrays of nl_operator_build."""

    def __init__(self, op, stencil_points: np.ndarray | None = None, coords: np.ndarray | None = None):
        _lib.init()
        self.op = op
        self.np, self.np_part = op.np, op.np_part
        h = C.c_void_p()
        w = np.ascontiguousarray(op.w, dtype=np.float64)
        ri = np.ascontiguousarray(op.ri, dtype=np.int32).ravel()
        rinv = np.ascontiguousarray(op.rimap_inv, dtype=np.int32)
        pts = None if stencil_points is None else np.ascontiguousarray(stencil_points, dtype=np.int32).ravel()
        crd = None if coords is None else np.ascontiguousarray(coords, dtype=np.int32).ravel()
        call("octhip_nl_operator_create", op.n, as_ptr(w, C.c_double), as_ptr(pts, C.c_int), op.nri,
             as_ptr(ri, C.c_int), as_ptr(rinv, C.c_int), op.np, op.np_part, as_ptr(crd, C.c_int),
             C.byref(h))
        self.handle = h

    def tune(self, **knobs) -> None:
        """Set tuning knobs by name: kernel (0 auto, 1 gather, 2 z-march), xcd_chunk,
        gather_tile, zseg, lane_chunk, lanes_per_thread."""
        for k, v in knobs.items():
            call("octhip_nl_operator_set_tuning", self.handle, _lib.TUNE[k], int(v))

    def tuning(self, key: str) -> int:
        v = C.c_int()
        call("octhip_nl_operator_get_tuning", self.handle, _lib.TUNE[key], C.byref(v))
        return v.value

    @property
    def uses_lattice(self) -> bool:
        f = C.c_int()
        call("octhip_nl_operator_uses_lattice", self.handle, C.byref(f))
        return bool(f.value)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().octhip_nl_operator_destroy(self.handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self.handle = None


def nl_operator_operate_batch(op: NLOperatorDevice, fi: Batch, fo: Batch, factor: float | None = None,
                              points: int = OP_ALL) -> None:
    """fo = factor * L fi on every inner point; fi's boundary rows are read as set."""
    assert fi.lanes == fo.lanes, "batches are not compatible"
    call("octhip_nl_operator_operate_batch", op.handle, fi.ptr(), fo.ptr(), fi.ldf,
         1.0 if factor is None else float(factor), points, _stream())


class HamiltonianDevice:
    """hamiltonian_elec_t image: kinetic operator, local potential, KB projector matrices."""

    def __init__(self, kinetic: NLOperatorDevice, vlocal: np.ndarray | None, proj=None, mass: float = 1.0):
        self.kinetic = kinetic
        self.np, self.np_part = kinetic.np, kinetic.np_part
        h = C.c_void_p()
        v = None if vlocal is None else np.ascontiguousarray(vlocal, dtype=np.float64)
        if proj is not None and proj.nmat > 0:
            offs = np.ascontiguousarray(proj.offsets, dtype=np.int32).ravel()
            mats = np.ascontiguousarray(proj.matrices, dtype=np.float64)
            maps = np.ascontiguousarray(proj.maps, dtype=np.int32)
            scals = np.ascontiguousarray(proj.scals, dtype=np.float64)
            nmat = proj.nmat
        else:
            offs = mats = maps = scals = None
            nmat = 0
        call("octhip_hamiltonian_create", kinetic.handle, float(mass), as_ptr(v, C.c_double), nmat,
             as_ptr(offs, C.c_int), as_ptr(mats, C.c_double), as_ptr(maps, C.c_int),
             as_ptr(scals, C.c_double), C.byref(h))
        self.handle = h

    @classmethod
    def from_system(cls, system, lattice: bool = True, mass: float = 1.0) -> "HamiltonianDevice":
        from .stencil import star_points
        kin = NLOperatorDevice(system.op, star_points(4) if lattice else None)
        return cls(kin, system.v, system.proj, mass)

    def tune(self, **knobs) -> None:
        """bra_mode (0 plain, 1 FP64 MFMA) plus the kinetic operator's knobs."""
        for k, v in knobs.items():
            call("octhip_hamiltonian_set_tuning", self.handle, _lib.TUNE[k], int(v))

    def autotune(self, ldf: int, reps: int = 5, force: bool = False) -> float:
        """Pick the fastest launch shape for rows of 2^ldf lanes; returns us per H-apply."""
        us = C.c_double()
        call("octhip_hamiltonian_autotune", self.handle, int(ldf), int(reps), int(bool(force)), _stream(),
             C.byref(us))
        return us.value

    def set_potential(self, v: np.ndarray) -> None:
        v = np.ascontiguousarray(v, dtype=np.float64)
        call("octhip_hamiltonian_set_potential", self.handle, as_ptr(v, C.c_double), _stream())

    def set_dipole_field(self, field, positions: np.ndarray | None = None) -> None:
        f = np.ascontiguousarray(field, dtype=np.float64)
        x = None if positions is None else np.ascontiguousarray(positions, dtype=np.float64).ravel()
        call("octhip_hamiltonian_set_dipole_field", self.handle, as_ptr(x, C.c_double),
             as_ptr(f, C.c_double), _stream())

    def release_stream(self) -> None:
        """Free the work buffers this Hamiltonian keeps for the current stream."""
        call("octhip_hamiltonian_release_stream", self.handle, _stream())

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().octhip_hamiltonian_destroy(self.handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self.handle = None


def hamiltonian_elec_apply_batch(hm: HamiltonianDevice, psib: Batch, hpsib: Batch, terms: int = TERM_ALL,
                                 set_bc: bool = True) -> None:
    assert psib.lanes == hpsib.lanes and psib.is_complex == hpsib.is_complex
    call("octhip_hamiltonian_apply_batch", hm.handle, psib.ptr(), hpsib.ptr(), psib.ldf,
         int(psib.is_complex), int(terms), int(bool(set_bc)), _stream())


class Exponential:
    """exponential_t with exp_method = taylor (TDExpOrder default 4, share/varinfo)."""

    def __init__(self, exp_order: int = 4):
        self.exp_order = exp_order

    def apply_batch(self, hm: HamiltonianDevice, psib: Batch, deltat: float, imag_time: bool = False) -> None:
        if not psib.is_complex:
            raise ValueError("exponential_apply_batch needs a complex batch")
        call("octhip_exponential_taylor_batch", hm.handle, psib.ptr(), psib.ldf, float(deltat),
             self.exp_order, int(bool(imag_time)), _stream())


class TaylorGraph:
    """One Taylor step over several batches captured into a HIP graph: the loop over
    batches of propagation_ops_elec_exp_apply (src/td/propagation_ops_elec.F90:232-249),
    replayed with one launch per time step.  The batches must stay alive (and at the same
    device address) for the life of the graph."""

    def __init__(self, te: Exponential, hm: HamiltonianDevice, batches: list[Batch], deltat: float,
                 imag_time: bool = False):
        if not batches:
            raise ValueError("TaylorGraph needs at least one batch")
        ldf = batches[0].ldf
        for b in batches:
            if not b.is_complex or b.ldf != ldf:
                raise ValueError("TaylorGraph batches must be complex with one ldf")
        self.hm, self.batches = hm, list(batches)
        ptrs = (C.c_void_p * len(batches))(*[b.ptr() for b in batches])
        h = C.c_void_p()
        call("octhip_taylor_graph_create", hm.handle, ptrs, len(batches), ldf, float(deltat),
             te.exp_order, int(bool(imag_time)), _stream(), C.byref(h))
        self.handle = h

    def launch(self) -> None:
        call("octhip_taylor_graph_launch", self.handle, _stream())

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().octhip_taylor_graph_destroy(self.handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self.handle = None


def exponential_apply_batch(te: Exponential, hm: HamiltonianDevice, psib: Batch, deltat: float,
                            imag_time: bool = False) -> None:
    te.apply_batch(hm, psib, deltat, imag_time)


class Adjoint:
    """Differentiable Taylor propagation of one complex batch under a dipole field
    H_s = H + eps_s (pol . r): the forward/backward pair of the optimal-control gradient
    (src/opt_control/propagation.F90:579-783, 1016-1103) made exact for the discrete
    propagator, with checkpoints every ``checkpoint_every`` steps (1108-1205)."""

    def __init__(self, hm: HamiltonianDevice, positions: np.ndarray, pol, ldf: int, nsteps: int, deltat: float,
                 vol_pp: float, order: int = 4, checkpoint_every: int = 10):
        x = np.ascontiguousarray(positions, dtype=np.float64)[: hm.np].ravel()
        p = np.ascontiguousarray(pol, dtype=np.float64)
        h = C.c_void_p()
        call("octhip_adjoint_create", hm.handle, as_ptr(x, C.c_double), as_ptr(p, C.c_double), int(ldf),
             int(nsteps), float(deltat), int(order), int(checkpoint_every), float(vol_pp), C.byref(h))
        self.handle, self.hm, self.nsteps, self.ldf = h, hm, int(nsteps), int(ldf)

    def forward(self, psib: Batch, eps) -> None:
        """psib: psi_0 in, psi_N out (in place)."""
        e = np.ascontiguousarray(eps, dtype=np.float64)
        if e.shape != (self.nsteps,) or psib.ldf != self.ldf:
            raise ValueError("eps must hold one value per step and the batch must match ldf")
        call("octhip_adjoint_forward", self.handle, psib.ptr(), as_ptr(e, C.c_double), _stream())

    def backward(self, chib: Batch):
        """chib: chi_N in, chi_0 out (in place).  Returns (dJ/deps numpy (nsteps,),
        dJ/dv torch float64 (np,) on the device)."""
        torch = _torch()
        deps = np.zeros(self.nsteps)
        dv = torch.empty(self.hm.np, dtype=torch.float64, device="cuda")
        call("octhip_adjoint_backward", self.handle, chib.ptr(), as_ptr(deps, C.c_double),
             C.c_void_p(dv.data_ptr()), _stream())
        return deps, dv

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().octhip_adjoint_destroy(self.handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self.handle = None


class BatchedPropagator:
    """Several (Hamiltonian, batch) pairs of uneven size propagated with one launch per
    kernel kind and Taylor order (instead of the per-batch, per-system loops of
    propagation_ops_elec_exp_apply, src/td/propagation_ops_elec.F90:232-249).  The
    batches must stay alive, at the same addresses, for the life of the object."""

    def __init__(self, pairs: list, deltat: float, order: int = 4, with_apply: bool = True):
        """pairs: list of (HamiltonianDevice, Batch) with complex batches."""
        if not pairs:
            raise ValueError("BatchedPropagator needs at least one pair")
        self.pairs = list(pairs)
        self.hpsi = [Batch(b.nst, hm.np, True, b.pack_size) for hm, b in pairs] if with_apply else None
        n = len(pairs)
        hms = (C.c_void_p * n)(*[hm.handle for hm, _ in pairs])
        psi = (C.c_void_p * n)(*[b.ptr() for _, b in pairs])
        hp = (C.c_void_p * n)(*[h.ptr() for h in self.hpsi]) if with_apply else None
        ldf = (C.c_int * n)(*[b.ldf for _, b in pairs])
        for _, b in pairs:
            if not b.is_complex:
                raise ValueError("BatchedPropagator needs complex batches")
        h = C.c_void_p()
        call("octhip_batched_create", n, hms, psi, hp, ldf, float(deltat), int(order), C.byref(h))
        self.handle = h

    def apply(self) -> None:
        """hpsi[i] = H_i psi[i] for every pair (results in self.hpsi)."""
        call("octhip_batched_apply", self.handle, _stream())

    def taylor_step(self) -> None:
        call("octhip_batched_taylor_step", self.handle, _stream())

    def graph(self) -> "TaylorGraph":
        """One batched Taylor step captured as a HIP graph (TaylorGraph.launch())."""
        h = C.c_void_p()
        call("octhip_batched_graph_create", self.handle, _stream(), C.byref(h))
        g = TaylorGraph.__new__(TaylorGraph)
        g.hm, g.batches, g.owner, g.handle = None, [b for _, b in self.pairs], self, h
        return g

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().octhip_batched_destroy(self.handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self.handle = None


def tuning_save(path) -> None:
    call("octhip_tuning_save", str(path).encode())


def tuning_load(path) -> None:
    call("octhip_tuning_load", str(path).encode())
