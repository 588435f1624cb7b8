"""ctypes binding of include/lqcd.h (liblqcd_hip.so, built for gfx950).

There is no fallback: if the library is missing or no GPU is visible, the
calls raise.  Arrays are numpy float64/uint64 in the reference's layouts.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblqcd_hip.so")

# lqcd_status
OK = 0
ERR_ZERO_DIMENSION = 1
ERR_NON_POSITIVE_SIZE = 2
ERR_DIM_TOO_SMALL = 3
ERR_INCOMPATIBLE_SIZE = 4
ERR_ZERO_INTEGRATION = 5
ERR_INVALID_ARGUMENT = 6
ERR_UNSUPPORTED = 7
ERR_DEVICE = 8
ERR_OUT_OF_MEMORY = 9
ERR_GAUSS_PROJECTION = 10
ERR_IO = 11
FORMAT_BINCODE, FORMAT_JSON = 0, 1

F64, F32 = 0, 1
DRIFT_ADDITIVE, DRIFT_EXP = 0, 1
SCHEME_REFERENCE, SCHEME_LEAPFROG = 0, 1

KERNEL_KICK_DRIFT = 0
KERNEL_KICK = 1
KERNEL_DRIFT = 2
KERNEL_PLAQUETTE = 3
KERNEL_EFIELD_ENERGY = 4
KERNEL_REUNIT = 5
KERNEL_METROPOLIS = 6
KERNEL_MOMENTA = 7
KERNEL_FORCE = 8
KERNEL_FIELDS = 9
KERNEL_STATS = 10
KERNEL_HEAT_BATH = 11
KERNEL_OVERRELAX = 12
STATS_SEQUENTIAL, STATS_PARALLEL = 0, 1
FIELD_CLOVER, FIELD_F_MU_NU, FIELD_MAGNETIC = 0, 1, 2


class HmcStats(C.Structure):
    _fields_ = [
        ("delta_h", C.c_double),
        ("h_old", C.c_double),
        ("h_new", C.c_double),
        ("probability", C.c_double),
        ("accepted", C.c_int),
    ]


class MhStats(C.Structure):
    _fields_ = [("prob_replace_mean", C.c_double), ("number_replace_last", C.c_size_t)]


class ChainSpec(C.Structure):
    _fields_ = [
        ("extent", C.c_size_t * 4),
        ("beta", C.c_double),
        ("size", C.c_double),
        ("seed", C.c_uint64),
    ]


class ChainStats(C.Structure):
    _fields_ = [
        ("h_old", C.c_double),
        ("h_new", C.c_double),
        ("delta_h", C.c_double),
        ("probability", C.c_double),
        ("plaquette", C.c_double),
        ("accepted", C.c_int32),
        ("reserved", C.c_int32),
    ]


# numpy view of an lqcd_chain_stats array
CHAIN_STATS_DTYPE = np.dtype([
    ("h_old", np.float64), ("h_new", np.float64), ("delta_h", np.float64),
    ("probability", np.float64), ("plaquette", np.float64),
    ("accepted", np.int32), ("reserved", np.int32),
])


_P = C.c_void_p
_D = C.c_double
_S = C.c_size_t
_I = C.c_int
_U64 = C.c_uint64

# name -> (restype, argtypes); every exported symbol of include/lqcd.h
SIGNATURES = {
    "lqcd_last_error": (C.c_char_p, []),
    "lqcd_version": (C.c_char_p, []),
    "lqcd_device_count": (_I, [C.POINTER(C.c_int)]),
    "lqcd_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "lqcd_ctx_destroy": (_I, [_P]),
    "lqcd_ctx_synchronize": (_I, [_P]),
    "lqcd_lattice_create": (_I, [_P, _I, C.POINTER(C.c_size_t), _D, _I, C.POINTER(_P)]),
    "lqcd_lattice_destroy": (_I, [_P]),
    "lqcd_lattice_info": (_I, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lqcd_links_upload": (_I, [_P, _P, _S]),
    "lqcd_links_download": (_I, [_P, _P, _S]),
    "lqcd_links_cold": (_I, [_P]),
    "lqcd_links_random": (_I, [_P, _U64]),
    "lqcd_efield_upload": (_I, [_P, _P, _S]),
    "lqcd_efield_download": (_I, [_P, _P, _S]),
    "lqcd_efield_cold": (_I, [_P]),
    "lqcd_efield_random": (_I, [_P, _U64, _U64, _D]),
    "lqcd_efield_gauss": (_I, [_P, _P, _S]),
    "lqcd_efield_gauss_sum_div": (_I, [_P, C.POINTER(_D)]),
    "lqcd_efield_project_gauss": (_I, [_P, _S, C.POINTER(C.c_size_t)]),
    "lqcd_average_trace_plaquette": (_I, [_P, _P]),
    "lqcd_hamiltonian_links": (_I, [_P, _D, C.POINTER(_D)]),
    "lqcd_hamiltonian_efield": (_I, [_P, _D, C.POINTER(_D)]),
    "lqcd_derivative_e": (_I, [_P, _P, _S]),
    "lqcd_derivative_u": (_I, [_P, _P, _S]),
    "lqcd_integrate_efield": (_I, [_P, _D]),
    "lqcd_integrate_link": (_I, [_P, _D, _I]),
    "lqcd_simulate_symplectic_n": (_I, [_P, _D, _S, _I, _I]),
    "lqcd_integrate_sync_sync": (_I, [_P, _D, _I]),
    "lqcd_integrate_sync_leap": (_I, [_P, _D]),
    "lqcd_integrate_leap_leap": (_I, [_P, _D, _I]),
    "lqcd_integrate_leap_sync": (_I, [_P, _D, _I]),
    "lqcd_simulate_sync_n": (_I, [_P, _D, _S, _I]),
    "lqcd_simulate_leap_n": (_I, [_P, _D, _S, _I]),
    "lqcd_hmc_trajectory": (_I, [_P, _D, _D, _S, _I, _I, _U64, C.POINTER(HmcStats)]),
    "lqcd_wilson_forward": (_I, [_P, _D, C.POINTER(_D)]),
    "lqcd_wilson_backward": (_I, [_P, _D, _P, _S]),
    "lqcd_wilson_backward_device": (_I, [_P, _D, _P, _S]),
    "lqcd_metropolis_sweep": (_I, [_P, _D, _S, _P, _P, C.POINTER(MhStats)]),
    "lqcd_metropolis_sweep_random": (_I, [_P, _D, _S, _D, _U64, _U64, C.POINTER(MhStats)]),
    "lqcd_heat_bath_sweep": (_I, [_P, _D, _U64, _U64, C.POINTER(C.c_size_t)]),
    "lqcd_metropolis_link_random": (_I, [_P, _D, _D, _U64, _U64, C.POINTER(_D), C.POINTER(C.c_int),
                                         C.POINTER(C.c_size_t)]),
    "lqcd_overrelaxation_sweep": (_I, [_P]),
    "lqcd_clover_field": (_I, [_P, _P, _S]),
    "lqcd_f_mu_nu_field": (_I, [_P, _P, _S]),
    "lqcd_plaquette_field": (_I, [_P, _I, _I, _I, _I, _I, _P, _S]),
    "lqcd_magnetic_field_vec": (_I, [_P, _P, _S]),
    "lqcd_field_strength_device": (_I, [_P, _I, _P, _S]),
    "lqcd_stats_mean": (_I, [_P, _P, _S, _I, _P]),
    "lqcd_stats_mean_and_variance": (_I, [_P, _P, _S, _I, _I, _P]),
    "lqcd_stats_mean_and_variance_noalloc": (_I, [_P, _P, _S, _I, _I, _P]),
    "lqcd_stats_mean_with_error": (_I, [_P, _P, _S, _I, _P]),
    "lqcd_stats_covariance": (_I, [_P, _P, _S, _P, _S, _I, _I, _P, C.POINTER(C.c_int)]),
    "lqcd_md_reconstruct_set": (_I, [_P, _I]),
    "lqcd_md_graph_set": (_I, [_P, _I]),
    "lqcd_batch_reconstruct_set": (_I, [_P, _I]),
    "lqcd_batch_reconstruct_get": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lqcd_md_reconstruct_get": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lqcd_su3_distance": (_I, [_P, C.POINTER(C.c_double)]),
    "lqcd_state_write": (_I, [_P, _D, C.c_char_p, _I]),
    "lqcd_state_ef_write": (_I, [_P, _D, _U64, C.c_char_p, _I]),
    "lqcd_state_read_header": (_I, [C.c_char_p, _I, _I, _I, C.POINTER(C.c_double), C.POINTER(C.c_size_t),
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "lqcd_state_read_host": (_I, [C.c_char_p, _I, _I, _I, _P, _S, _P]),
    "lqcd_state_read": (_I, [_P, C.c_char_p, _I, _I, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "lqcd_overrelaxation_reverse_sweep": (_I, [_P]),
    "lqcd_normalize_link_matrices": (_I, [_P]),
    "lqcd_stream_copy": (_I, [_P, _S, _I, C.POINTER(_D)]),
    "lqcd_stream_read": (_I, [_P, _S, _I, C.POINTER(_D)]),
    "lqcd_md_autotune": (_I, [_P, _I, C.POINTER(C.c_int)]),
    "lqcd_md_variant_set": (_I, [_P, _I]),
    "lqcd_md_variant_get": (_I, [_P, C.POINTER(C.c_int)]),
    "lqcd_profile_enable": (_I, [_P, C.c_uint]),
    "lqcd_profile_read": (_I, [_P, _I, C.POINTER(C.c_size_t), C.POINTER(_D)]),
    "lqcd_batch_create": (_I, [_P, _I, _I, C.POINTER(ChainSpec), _S, C.POINTER(_P)]),
    "lqcd_batch_destroy": (_I, [_P]),
    "lqcd_batch_info": (_I, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lqcd_batch_chain_info": (_I, [_P, _S, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lqcd_batch_cold_start": (_I, [_P]),
    "lqcd_batch_hot_start": (_I, [_P]),
    "lqcd_batch_links_upload": (_I, [_P, _S, _P, _S]),
    "lqcd_batch_links_download": (_I, [_P, _S, _P, _S]),
    "lqcd_batch_plaquette": (_I, [_P, _P]),
    "lqcd_batch_counter_get": (_I, [_P, C.POINTER(C.c_uint64)]),
    "lqcd_batch_counter_set": (_I, [_P, _U64]),
    "lqcd_batch_hmc": (_I, [_P, _D, _S, _I, _I, _S, _I, _P]),
    "lqcd_batch_autotune": (_I, [_P, _I, C.POINTER(C.c_int)]),
    "lqcd_batch_config_set": (_I, [_P, _I]),
    "lqcd_batch_config_get": (_I, [_P, C.POINTER(C.c_int)]),
    "lqcd_batch_graph_enable": (_I, [_P, _I]),
}


class LqcdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[lqcd status {code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load liblqcd_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python lattice-qcd-rs_amd/build_lib.py` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().lqcd_last_error().decode(errors="replace")
        raise LqcdError(rc, msg)


def ptr(a: np.ndarray):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


_hip = None


def hip_runtime():
    """The HIP runtime liblqcd_hip.so links against (loaded with it).  Device
    buffers handed to the library's device-pointer entry points must come from
    this runtime: a second HIP runtime in the process (a framework's bundled
    one) does not see the library's device state."""
    global _hip
    if _hip is None:
        lib()
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        h.hipMalloc.restype = C.c_int
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.restype = C.c_int
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceBuffer:
    """``nbytes`` of device memory on the current device, freed with the object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self._p = C.c_void_p()
        rc = hip_runtime().hipMalloc(C.byref(self._p), self.nbytes)
        if rc != 0:
            raise LqcdError(ERR_OUT_OF_MEMORY, f"hipMalloc({self.nbytes}) failed with {rc}")

    @property
    def ptr(self) -> int:
        return int(self._p.value)

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        rc = hip_runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._p, out.nbytes, 2)  # D2H
        if rc != 0:
            raise LqcdError(ERR_DEVICE, f"hipMemcpy failed with {rc}")
        return out

    def fill_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        rc = hip_runtime().hipMemcpy(self._p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1)  # H2D
        if rc != 0:
            raise LqcdError(ERR_DEVICE, f"hipMemcpy failed with {rc}")

    def __del__(self):
        if _hip is not None and self._p:
            _hip.hipFree(self._p)
            self._p = C.c_void_p()
