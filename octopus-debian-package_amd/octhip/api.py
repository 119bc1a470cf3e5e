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


def batch_pack(unpacked, psib: "Batch", first: int = 0) -> None:
    """batch_write_unpacked_to_device (src/grid/batch.F90:625-701): unpacked is a torch
    tensor (nst, ld) on the device, complex128 for complex batches (float64 otherwise),
    one state per row as batch_t%zff_linear's columns; states go to elements
    first..first+nst-1 of psib's rows 0..np-1 (np = min(ld, psib.nrows))."""
    torch = _torch()
    want = torch.complex128 if psib.is_complex else torch.float64
    if unpacked.dtype != want or not unpacked.is_cuda or not unpacked.is_contiguous() or unpacked.dim() != 2:
        raise ValueError("unpacked must be a contiguous (nst, ld) device tensor of the batch's type")
    nst, ld = unpacked.shape
    call("octhip_batch_pack", C.c_void_p(unpacked.data_ptr()), ld, nst, min(ld, psib.nrows), int(psib.is_complex),
         psib.ptr(), psib.ldf, int(first), _stream())


def batch_unpack(psib: "Batch", unpacked, first: int = 0) -> None:
    """batch_read_device_to_unpacked (src/grid/batch.F90:703-760): the reverse of batch_pack."""
    torch = _torch()
    want = torch.complex128 if psib.is_complex else torch.float64
    if unpacked.dtype != want or not unpacked.is_cuda or not unpacked.is_contiguous() or unpacked.dim() != 2:
        raise ValueError("unpacked must be a contiguous (nst, ld) device tensor of the batch's type")
    nst, ld = unpacked.shape
    call("octhip_batch_unpack", psib.ptr(), psib.ldf, int(first), nst, min(ld, psib.nrows), int(psib.is_complex),
         C.c_void_p(unpacked.data_ptr()), ld, _stream())


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

    # mesh volume element (the Lanczos inner products) and the spectral range the Chebyshev
    # expansion assumes (hm%spectral_middle_point / spectral_half_span)
    vol_pp = None
    spectral_middle_point = None
    spectral_half_span = None

    @classmethod
    def from_system(cls, system, lattice: bool = True, mass: float = 1.0) -> "HamiltonianDevice":
        from .stencil import star_points
        kin = NLOperatorDevice(system.op, star_points(4) if lattice else None)
        hm = cls(kin, system.v, system.proj, mass)
        hm.vol_pp = float(system.mesh.vol_pp)
        return hm

    def span(self, delta: float, emin: float) -> None:
        """hamiltonian_elec_span (src/hamiltonian/hamiltonian_elec.F90:898-909)."""
        self.spectral_middle_point = ((np.pi ** 2 / (2 * delta ** 2)) + emin) / 2.0
        self.spectral_half_span = ((np.pi ** 2 / (2 * delta ** 2)) - emin) / 2.0

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

    def autotune_pack(self, nst: int, reps: int = 3, force: bool = False) -> tuple[int, float]:
        """StatesBlockSize for nst states by time (block sizes 8..64); leaves the launch
        shape of that row length configured.  Returns (block size, us per Taylor-4 step of
        all nst states)."""
        pack, us = C.c_int(), C.c_double()
        call("octhip_hamiltonian_autotune_pack", self.handle, int(nst), int(reps), int(bool(force)), _stream(),
             C.byref(pack), C.byref(us))
        return pack.value, us.value

    def set_potential(self, v: np.ndarray) -> None:
        v = np.ascontiguousarray(v, dtype=np.float64)
        call("octhip_hamiltonian_set_potential", self.handle, as_ptr(v, C.c_double), _stream())

    def set_dipole_field(self, field, positions: np.ndarray | None = None) -> None:
        f = np.ascontiguousarray(field, dtype=np.float64)
        x = None if positions is None else np.ascontiguousarray(positions, dtype=np.float64).ravel()
        call("octhip_hamiltonian_set_dipole_field", self.handle, as_ptr(x, C.c_double),
             as_ptr(f, C.c_double), _stream())

    def set_periodic(self, bnd_src) -> None:
        """Periodic boundaries: boundary row np+b takes inner row bnd_src[b] (None: zero
        boundaries again)."""
        b = None if bnd_src is None else np.ascontiguousarray(bnd_src, dtype=np.int32)
        _torch().cuda.synchronize()
        call("octhip_hamiltonian_set_periodic", self.handle, as_ptr(b, C.c_int))

    def set_phases(self, phase, phase_corr=None, projector_phases=None) -> None:
        """Bloch phases of one k-point (complex arrays): hm_base%phase (np_part),
        phase_corr (np_part - np) and the projector phases (total_points); None removes them."""
        def cp(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, np.complex128)).view(np.float64)
        ph, pc, pp = cp(phase), cp(phase_corr), cp(projector_phases)
        _torch().cuda.synchronize()
        call("octhip_hamiltonian_set_phases", self.handle, as_ptr(ph, C.c_double), as_ptr(pc, C.c_double),
             as_ptr(pp, C.c_double))

    def base_phase(self, psib: "Batch", np_rows: int, conjugate: bool, src: "Batch | None" = None) -> None:
        """hamiltonian_elec_base_phase (src/hamiltonian/hamiltonian_elec_base.F90:979-1139);
        np_rows = self.np is set_phase_corr (conjugate False) / unset_phase_corr (True)."""
        call("octhip_hamiltonian_base_phase", self.handle, psib.ptr(), None if src is None else src.ptr(),
             psib.ldf, int(np_rows), int(bool(conjugate)), _stream())

    def memory(self) -> dict:
        """Device bytes held by this handle: data, per-stream work buffers, sets held by
        Taylor graphs, retired buffers (octhip_hamiltonian_memory)."""
        b = (C.c_size_t * 4)()
        call("octhip_hamiltonian_memory", self.handle, b)
        return dict(data=b[0], work=b[1], graphs=b[2], retired=b[3])

    def trim(self) -> None:
        """Free the retired work buffers (synchronises the device)."""
        call("octhip_hamiltonian_trim", self.handle)

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
                                 set_bc: bool = True, has_phase: bool = False) -> None:
    """has_phase: psib%has_phase (the batch carries the Bloch phase of the handle's k-point)."""
    assert psib.lanes == hpsib.lanes and psib.is_complex == hpsib.is_complex
    if has_phase:
        call("octhip_hamiltonian_apply_batch_phase", hm.handle, psib.ptr(), hpsib.ptr(), psib.ldf, int(terms),
             int(bool(set_bc)), 1, _stream())
        return
    call("octhip_hamiltonian_apply_batch", hm.handle, psib.ptr(), hpsib.ptr(), psib.ldf,
         int(psib.is_complex), int(terms), int(bool(set_bc)), _stream())


def _ptr_or_none(b):
    return None if b is None else b.ptr()


class Exponential:
    """exponential_t (src/electrons/exponential.F90:45-194): exp_method taylor (TDExpOrder
    default 4), lanczos (TDLanczosTol default 1e-5) or chebyshev."""

    def __init__(self, exp_order: int = 4, method: str = "taylor", lanczos_tol: float = 1e-5):
        if method not in ("taylor", "lanczos", "chebyshev"):
            raise ValueError(f"unknown exponential method {method}")
        self.exp_order = exp_order
        self.exp_method = method
        self.lanczos_tol = lanczos_tol
        self.last_iterations = None  # Lanczos: Krylov dimension and error estimate of the last call
        self.last_residual = None

    def _lanczos(self, hm, psib, deltat, imag_time, inh_psib=None):
        if hm.vol_pp is None:
            raise ValueError("Lanczos needs the mesh volume element (hm.vol_pp)")
        it, res = C.c_int(), C.c_double()
        call("octhip_exponential_lanczos_batch", hm.handle, psib.ptr(), psib.ldf, psib.nst, float(deltat),
             int(bool(imag_time)), self.exp_order, float(self.lanczos_tol), _ptr_or_none(inh_psib),
             float(hm.vol_pp), C.byref(it), C.byref(res), _stream())
        self.last_iterations, self.last_residual = it.value, res.value
        if res.value > self.lanczos_tol:
            import warnings
            warnings.warn(f"Lanczos exponential expansion did not converge: {res.value:.2e}")

    def _cheby(self, hm, psib, deltat):
        if hm.spectral_half_span is None:
            raise ValueError("Chebyshev needs the spectral range (hm.span(delta, emin))")
        call("octhip_exponential_cheby_batch", hm.handle, psib.ptr(), psib.ldf, float(deltat), self.exp_order,
             float(hm.spectral_middle_point), float(hm.spectral_half_span), _stream())

    def _apply_krylov(self, hm, psib, deltat, imag_time, psib2, deltat2, inh_psib):
        """The lanczos and chebyshev branches of exponential_apply_batch
        (src/electrons/exponential.F90:648-669)."""
        if psib2 is not None:
            call("octhip_memcpy_dtod", psib2.ptr(), psib.ptr(), hm.np * psib.lanes * 8, _stream())
        if self.exp_method == "lanczos":
            self._lanczos(hm, psib, deltat, imag_time)
            if inh_psib is not None:
                self._lanczos(hm, psib, deltat, imag_time, inh_psib)
            if psib2 is not None:
                self._lanczos(hm, psib2, deltat2, imag_time)
                if inh_psib is not None:
                    self._lanczos(hm, psib2, deltat2, imag_time, inh_psib)
            return
        if imag_time:
            raise ValueError("imaginary time needs the taylor or lanczos method")  # exponential.F90:614-618
        if inh_psib is not None:
            raise ValueError("the Chebyshev exponential has no inhomogeneous term")  # exponential.F90:660-664
        self._cheby(hm, psib, deltat)
        if psib2 is not None:
            self._cheby(hm, psib2, deltat2)

    def apply_batch(self, hm: HamiltonianDevice, psib: Batch, deltat: float, imag_time: bool = False, *,
                    psib2: Batch | None = None, deltat2: float | None = None, inh_psib: Batch | None = None,
                    density=None, weights=None) -> None:
        """te%apply_batch(namespace, mesh, hm, psib, deltat, psib2, deltat2, vmagnus, imag_time,
        inh_psib) (src/electrons/exponential.F90:581-592); density (torch float64 (np,) on the
        device) += sum_st weights[st] |final psi|^2 of psib2 when given, else psib (the
        density_calc_accumulate of propagation_ops_elec_fuse_density_exp_apply)."""
        if not psib.is_complex:
            raise ValueError("exponential_apply_batch needs a complex batch")
        if (psib2 is None) != (deltat2 is None):
            raise ValueError("psib2 and deltat2 go together")  # ASSERT at exponential.F90:602
        for b in (psib2, inh_psib):
            if b is not None and (b.ldf != psib.ldf or not b.is_complex):
                raise ValueError("psib2 / inh_psib must match psib")
        if (density is None) != (weights is None):
            raise ValueError("density and weights go together")
        if self.exp_method != "taylor":
            self._apply_krylov(hm, psib, deltat, imag_time, psib2, deltat2, inh_psib)
            if density is not None:
                density_calc_accumulate(psib2 if psib2 is not None else psib, hm.np, weights, density)
            return
        if psib2 is None and inh_psib is None and density is None:
            call("octhip_exponential_taylor_batch", hm.handle, psib.ptr(), psib.ldf, float(deltat),
                 self.exp_order, int(bool(imag_time)), _stream())
            return
        w = None if weights is None else _weights(weights, psib.pack_size)
        call("octhip_exponential_apply_batch", hm.handle, psib.ptr(), psib.ldf, float(deltat), self.exp_order,
             int(bool(imag_time)), _ptr_or_none(psib2), float(deltat2 or 0.0), _ptr_or_none(inh_psib),
             as_ptr(w, C.c_double), None if density is None else C.c_void_p(density.data_ptr()), _stream())


def _weights(weights, pack: int) -> np.ndarray:
    """kweight*occ per state, padded with zeros to the pack size."""
    w = np.zeros(pack)
    v = np.asarray(weights, dtype=np.float64).ravel()
    if len(v) > pack:
        raise ValueError("more weights than states in the batch")
    w[: len(v)] = v
    return w


def mesh_batch_dotp_vector(vol: float, aa: Batch, bb: Batch, rows: int) -> np.ndarray:
    """<aa_st|bb_st> = vol sum_p conj(aa) bb per state (src/grid/mesh_batch_inc.F90:379-549)."""
    if aa.lanes != bb.lanes or aa.is_complex != bb.is_complex:
        raise ValueError("batches are not compatible")
    n = aa.pack_size if aa.is_complex else aa.lanes
    out = np.zeros(2 * n if aa.is_complex else n)
    call("octhip_mesh_batch_dotp_vector", aa.ptr(), bb.ptr(), aa.ldf, int(rows), int(aa.is_complex), float(vol),
         as_ptr(out, C.c_double), _stream())
    return (out.view(np.complex128) if aa.is_complex else out)[: aa.nst].copy()


def mesh_batch_nrm2(vol: float, aa: Batch, rows: int) -> np.ndarray:
    """mesh_batch_nrm2 (src/grid/mesh_batch.F90:74-95)."""
    n = aa.pack_size if aa.is_complex else aa.lanes
    out = np.zeros(n)
    call("octhip_mesh_batch_nrm2", aa.ptr(), aa.ldf, int(rows), int(aa.is_complex), float(vol),
         as_ptr(out, C.c_double), _stream())
    return out[: aa.nst].copy()


def mesh_batch_orthogonalization(vol: float, psibs: list, phib: Batch, rows: int, normalize: bool = False):
    """zmesh_batch_orthogonalization, classical Gram-Schmidt (src/grid/mesh_batch_inc.F90:1168-1274).
    Returns (overlap (len(psibs), nst) complex, norm (nst,))."""
    n = len(psibs)
    ptrs = (C.c_void_p * max(n, 1))(*[b.ptr() for b in psibs])
    ov = np.zeros(2 * n * phib.nst)
    nrm = np.zeros(phib.nst)
    call("octhip_mesh_batch_orthogonalization", n, ptrs, phib.ptr(), phib.ldf, phib.nst, int(rows), float(vol),
         int(bool(normalize)), as_ptr(ov, C.c_double), as_ptr(nrm, C.c_double), _stream())
    return ov.view(np.complex128).reshape(phib.nst, n).T.copy(), nrm


def density_calc_accumulate(psib: Batch, rows: int, weights, density) -> None:
    """density_calc_accumulate on a device-packed batch (src/states/density.F90:279-424):
    density (torch float64 on the device) += sum_st weights[st] |psi_st|^2 over rows points."""
    torch = _torch()
    n = psib.pack_size if psib.is_complex else psib.lanes
    w = torch.from_numpy(_weights(weights, n)).to("cuda")
    call("octhip_density_accumulate", psib.ptr(), psib.ldf, int(rows), int(psib.is_complex),
         C.c_void_p(w.data_ptr()), C.c_void_p(density.data_ptr()), _stream())
    torch.cuda.current_stream().synchronize()  # w is freed on return


def _driver_args(batches, weights):
    ldf = batches[0].ldf
    for b in batches:
        if not b.is_complex or b.ldf != ldf:
            raise ValueError("propagator batches must be complex with one ldf")
    ptrs = (C.c_void_p * len(batches))(*[b.ptr() for b in batches])
    w = None
    if weights is not None:
        w = np.concatenate([_weights(wb, batches[0].pack_size) for wb in weights])
    return ldf, ptrs, w


def propagator_expmid(te: Exponential, hm: HamiltonianDevice, batches: list, dt: float, field_mid=None,
                      weights=None, density=None) -> None:
    """exponential_midpoint for independent particles (src/td/propagator_expmid.F90:52-100);
    weights: one list of state weights per batch; density (np,) on the device is set."""
    ldf, ptrs, w = _driver_args(batches, weights)
    f = None if field_mid is None else np.ascontiguousarray(field_mid, dtype=np.float64)
    call("octhip_propagator_expmid", hm.handle, ptrs, len(batches), ldf, float(dt), te.exp_order,
         as_ptr(f, C.c_double), as_ptr(w, C.c_double),
         None if density is None else C.c_void_p(density.data_ptr()), _stream())


def propagator_etrs(te: Exponential, hm: HamiltonianDevice, batches: list, dt: float, field_old=None,
                    field_new=None, weights=None, density=None) -> None:
    """td_etrs for independent particles (src/td/propagator_etrs.F90:68-150)."""
    ldf, ptrs, w = _driver_args(batches, weights)
    fo = None if field_old is None else np.ascontiguousarray(field_old, dtype=np.float64)
    fn = None if field_new is None else np.ascontiguousarray(field_new, dtype=np.float64)
    call("octhip_propagator_etrs", hm.handle, ptrs, len(batches), ldf, float(dt), te.exp_order,
         as_ptr(fo, C.c_double), as_ptr(fn, C.c_double), as_ptr(w, C.c_double),
         None if density is None else C.c_void_p(density.data_ptr()), _stream())


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
                            imag_time: bool = False, **optional) -> None:
    te.apply_batch(hm, psib, deltat, imag_time=imag_time, **optional)


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
