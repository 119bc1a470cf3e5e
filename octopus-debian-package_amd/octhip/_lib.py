"""ctypes binding of liboctopus_hip.so (the C ABI declared in include/octhip.h).

Loading never falls back to anything: if the library is missing the import raises,
so no caller can silently run a CPU path in its place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "liboctopus_hip.so"

SUCCESS = 0
ERR_INVALID, ERR_HIP, ERR_ALLOC, ERR_UNSUPPORTED, ERR_STATE = -1, -2, -3, -4, -5
OP_INNER, OP_OUTER, OP_ALL = 1, 2, 3
TERM_KINETIC, TERM_LOCAL_POTENTIAL, TERM_NON_LOCAL_POTENTIAL = 1, 2, 4
TERM_ALL = 0x7FFFFFFF
TUNE = dict(kernel=1, xcd_chunk=2, gather_tile=3, zseg=4, lane_chunk=5, lanes_per_thread=6, bra_mode=16)

# (name, restype-is-int, argtypes) for every symbol of include/octhip.h
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_S = C.c_size_t
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "octhip_init": [_I],
    "octhip_finalize": [],
    "octhip_get_last_error": [C.c_char_p, _S],
    "octhip_version": [_IP, _IP],
    "octhip_malloc": [C.POINTER(_P), _S],
    "octhip_free": [_P],
    "octhip_memcpy_htod": [_P, _P, _S, _P],
    "octhip_memcpy_dtoh": [_P, _P, _S, _P],
    "octhip_memcpy_dtod": [_P, _P, _S, _P],
    "octhip_stream_synchronize": [_P],
    "octhip_batch_pack": [_P, _S, _I, _I, _I, _P, _I, _I, _P],
    "octhip_batch_unpack": [_P, _I, _I, _I, _I, _I, _P, _S, _P],
    "octhip_nl_operator_create": [_I, _DP, _IP, _I, _IP, _IP, _I, _I, _IP, C.POINTER(_P)],
    "octhip_nl_operator_destroy": [_P],
    "octhip_nl_operator_uses_lattice": [_P, _IP],
    "octhip_nl_operator_operate_batch": [_P, _P, _P, _I, _D, _I, _P],
    "octhip_operate_ri_vec": [_I, _DP, _I, _P, _P, _P, _P, _I, _P, _P],
    "octhip_hamiltonian_create": [_P, _D, _DP, _I, _IP, _DP, _IP, _DP, C.POINTER(_P)],
    "octhip_hamiltonian_destroy": [_P],
    "octhip_hamiltonian_release_stream": [_P, _P],
    "octhip_hamiltonian_set_potential": [_P, _DP, _P],
    "octhip_hamiltonian_set_dipole_field": [_P, _DP, _DP, _P],
    "octhip_hamiltonian_apply_batch": [_P, _P, _P, _I, _I, _I, _I, _P],
    "octhip_hamiltonian_set_periodic": [_P, _IP],
    "octhip_hamiltonian_set_phases": [_P, _DP, _DP, _DP],
    "octhip_hamiltonian_apply_batch_phase": [_P, _P, _P, _I, _I, _I, _I, _P],
    "octhip_hamiltonian_base_phase": [_P, _P, _P, _I, _I, _I, _P],
    "octhip_exponential_taylor_batch": [_P, _P, _I, _D, _I, _I, _P],
    "octhip_exponential_apply_batch": [_P, _P, _I, _D, _I, _I, _P, _D, _P, _DP, _P, _P],
    "octhip_density_accumulate": [_P, _I, _I, _I, _P, _P, _P],
    "octhip_mesh_batch_dotp_vector": [_P, _P, _I, _I, _I, _D, _DP, _P],
    "octhip_mesh_batch_dotp_vector_noalloc": [_P, _P, _I, _I, _I, _D, _P, _P],
    "octhip_mesh_batch_nrm2": [_P, _I, _I, _I, _D, _DP, _P],
    "octhip_mesh_batch_orthogonalization": [_I, _P, _P, _I, _I, _I, _D, _I, _DP, _DP, _P],
    "octhip_exponential_lanczos_batch": [_P, _P, _I, _I, _D, _I, _I, _D, _P, _D, _IP, _DP, _P],
    "octhip_exponential_cheby_batch": [_P, _P, _I, _D, _I, _D, _D, _P],
    "octhip_propagator_expmid": [_P, _P, _I, _I, _D, _I, _DP, _DP, _P, _P],
    "octhip_propagator_etrs": [_P, _P, _I, _I, _D, _I, _DP, _DP, _DP, _P, _P],
    "octhip_taylor_graph_create": [_P, _P, _I, _I, _D, _I, _I, _P, _P],
    "octhip_taylor_graph_launch": [_P, _P],
    "octhip_taylor_graph_destroy": [_P],
    "octhip_adjoint_create": [_P, _P, _P, _I, _I, _D, _I, _I, _D, _P],
    "octhip_adjoint_forward": [_P, _P, _P, _P],
    "octhip_adjoint_backward": [_P, _P, _P, _P, _P],
    "octhip_adjoint_destroy": [_P],
    "octhip_batched_create": [_I, _P, _P, _P, _P, _D, _I, _P],
    "octhip_batched_apply": [_P, _P],
    "octhip_batched_taylor_step": [_P, _P],
    "octhip_batched_destroy": [_P],
    "octhip_batched_graph_create": [_P, _P, _P],
    "octhip_hamiltonian_autotune": [_P, _I, _I, _I, _P, _P],
    "octhip_hamiltonian_autotune_pack": [_P, _I, _I, _I, _P, _IP, _DP],
    "octhip_hamiltonian_memory": [_P, C.POINTER(C.c_size_t)],
    "octhip_hamiltonian_trim": [_P],
    "octhip_inject_alloc_failure": [C.c_long],
    "octhip_tuning_save": [C.c_char_p],
    "octhip_tuning_load": [C.c_char_p],
    "octhip_profiling_select": [C.c_char_p],
    "octhip_profiling_start": [],
    "octhip_profiling_stop": [],
    "octhip_profiling_get": [C.c_char_p, _DP, _IP],
    "octhip_nl_operator_set_tuning": [_P, _I, _I],
    "octhip_nl_operator_get_tuning": [_P, _I, _IP],
    "octhip_hamiltonian_set_tuning": [_P, _I, _I],
}


class OcthipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"octhip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OSError(f"{LIB_PATH} is missing: build it with "
                          "python octopus-debian-package_amd/build.py")
        l = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, args in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = C.c_int
            f.argtypes = args
        _lib = l
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().octhip_get_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != SUCCESS:
        raise OcthipError(rc, last_error())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def exported_symbols() -> list[str]:
    return list(SIGNATURES)


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        call("octhip_init", device)
        _initialised = True


def as_ptr(a, ctype):
    """Host numpy array -> ctypes pointer (array must stay alive during the call)."""
    import numpy as np
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def profiling_start() -> None:
    call("octhip_profiling_start")


def profiling_select(region: str | None) -> None:
    """Record only `region` (None: every region)."""
    call("octhip_profiling_select", (region or "").encode())


def profiling_stop() -> None:
    call("octhip_profiling_stop")


def profiling_get(region: str) -> tuple[float, int]:
    """(summed device milliseconds, launches) of one kernel region since profiling_start."""
    ms, n = C.c_double(), C.c_int()
    call("octhip_profiling_get", region.encode(), C.byref(ms), C.byref(n))
    return ms.value, n.value
