"""ctypes view of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg, never by the product package.

Arrays use the reference's in-memory layouts:

* links ``(nsite*D, 18)`` float64: column-major 3x3 complex, re/im interleaved
  (``src/field.rs:584-586``), link index ``site*D + mu``
  (``src/lattice.rs:923-927``);
* efield ``(nsite*D, 8)`` float64 (``src/field.rs:1025-1027``).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

DRIFT_ADDITIVE, DRIFT_EXP = 0, 1
SCHEME_REFERENCE, SCHEME_LEAPFROG = 0, 1
ORDER_LEXICOGRAPHIC, ORDER_CHECKERBOARD = 0, 1
MAX_DIM = 8

# The reference's test seed, src/test/mod.rs:19.
SEED_RNG = 0x4578_93F4_4AB0_67F0


class Lattice(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("extent", C.c_size_t * MAX_DIM),
        ("stride", C.c_size_t * MAX_DIM),
        ("nsite", C.c_size_t),
        ("size", C.c_double),
    ]


class Rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


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


class Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_D = C.c_double
_S = C.c_size_t
_I = C.c_int
_U64 = C.c_uint64


def _declare(L):
    LP = C.POINTER(Lattice)
    sig = {
        "or_lattice_init": (_I, [LP, _I, C.POINTER(C.c_size_t), _D]),
        "or_site_index": (_S, [LP, C.POINTER(C.c_size_t)]),
        "or_site_coords": (None, [LP, _S, C.POINTER(C.c_size_t)]),
        "or_neighbor": (_S, [LP, _S, _I, _I]),
        "or_site_parity": (_I, [LP, _S]),
        "or_link_canonical": (_S, [LP, _S, _I, _I]),
        "or_link_order": (None, [LP, _P, _P]),
        "or_generator": (None, [_I, _P]),
        "or_adjoint_to_matrix": (None, [_P, _P]),
        "or_trace_squared": (_D, [_P]),
        "or_adjoint_d": (Cplx, [_P]),
        "or_su3_exp_i": (None, [_P, _P]),
        "or_su3_exp_r": (None, [_P, _P]),
        "or_orthonormalize": (None, [_P, _P]),
        "or_det": (Cplx, [_P]),
        "or_matmul": (None, [_P, _P, _P]),
        "or_factorial_size_for_exp": (C.c_ulonglong, []),
        "or_rng_seed": (None, [C.POINTER(Rng), _U64]),
        "or_rng_next": (_U64, [C.POINTER(Rng)]),
        "or_rng_uniform": (_D, [C.POINTER(Rng), _D, _D]),
        "or_random_su3": (None, [C.POINTER(Rng), _P]),
        "or_random_su3_close_to_unity": (None, [C.POINTER(Rng), _D, _P]),
        "or_random_su2_close_to_unity": (None, [C.POINTER(Rng), _D, _P]),
        "or_hot_start": (None, [LP, _U64, _P]),
        "or_cold_start": (None, [LP, _P]),
        "or_random_efield": (None, [LP, _U64, _D, _P]),
        "or_mh_proposals": (None, [LP, _U64, _D, _S, _P]),
        "or_philox4x32": (None, [_P, _P, _P]),
        "or_device_momenta": (None, [LP, _U64, _U64, _D, _P]),
        "or_device_hot_start": (None, [LP, _U64, _P]),
        "or_device_mh_inputs": (None, [LP, _U64, _U64, _D, _S, _P, _P]),
        "or_su2_project_unorm": (None, [_P, _P]),
        "or_heat_bath_sweep": (_I, [LP, _P, _D, _U64, _U64, _I, C.POINTER(C.c_size_t)]),
        "or_heat_bath_x0": (_D, [_D, _U64, _U64, _U64]),
        "or_polar_unitary": (None, [_P, _P]),
        "or_overrelaxation_sweep": (_I, [LP, _P, _I]),
        "or_svd3": (None, [_P, _P, _P, _P]),
        "or_overrelaxation_reverse_sweep": (_I, [LP, _P, _I]),
        "or_bernoulli_accept": (_I, [_D, _U64]),
        "or_clamp_probability": (_D, [_D]),
        "or_link_matrix": (None, [LP, _P, _S, _I, _I, _P]),
        "or_sij": (None, [LP, _P, _S, _I, _I, _I, _I, _P]),
        "or_pij": (None, [LP, _P, _S, _I, _I, _I, _I, _P]),
        "or_clover": (None, [LP, _P, _S, _I, _I, _P]),
        "or_f_mu_nu": (None, [LP, _P, _S, _I, _I, _P]),
        "or_levi_civita": (_I, [_P, _I]),
        "or_magnetic_field": (None, [LP, _P, _S, _I, _P]),
        "or_clover_field": (None, [LP, _P, _P]),
        "or_f_mu_nu_field": (None, [LP, _P, _P]),
        "or_magnetic_field_vec": (None, [LP, _P, _P]),
        "or_average_trace_plaquette": (Cplx, [LP, _P]),
        "or_hamiltonian_links": (_D, [LP, _P, _D]),
        "or_hamiltonian_efield": (_D, [LP, _P, _D]),
        "or_derivative_e": (None, [LP, _P, _S, _P]),
        "or_derivative_e_all": (None, [LP, _P, _P]),
        "or_derivative_u": (None, [LP, _P, _P, _S, _I, _P]),
        "or_wilson_backward": (None, [LP, _P, _D, _P]),
        "or_integrate_efield": (None, [LP, _P, _P, _D, _P]),
        "or_integrate_link": (None, [LP, _P, _P, _D, _I, _P]),
        "or_integrate_symplectic": (None, [LP, _P, _P, _D, _I]),
        "or_simulate_symplectic_n": (_I, [LP, _P, _P, _D, _S, _I, _I]),
        "or_hmc_trajectory": (_I, [LP, _P, _P, _D, _D, _S, _I, _I, _U64, C.POINTER(HmcStats)]),
        "or_staple": (None, [LP, _P, _S, _I, _P]),
        "or_delta_s": (_D, [LP, _P, _S, _I, _P, _D]),
        "or_metropolis_sweep": (_I, [LP, _P, _D, _S, _P, _P, _I, C.POINTER(MhStats)]),
        "or_normalize_link_matrices": (None, [LP, _P]),
        "or_gauss": (None, [LP, _P, _P, _S, _P]),
        "or_gauss_sum_div": (_D, [LP, _P, _P]),
        "or_project_to_gauss": (_I, [LP, _P, _P, _P, _S, C.POINTER(C.c_size_t)]),
        "or_stats_mean": (None, [_P, _S, _I, _P]),
        "or_stats_mean_and_variance": (_I, [_P, _S, _I, _P]),
        "or_stats_mean_and_variance_par": (_I, [_P, _S, _I, _P]),
        "or_stats_covariance": (_I, [_P, _S, _P, _S, _I, _P]),
        "or_stats_covariance_par": (_I, [_P, _S, _P, _S, _I, _P]),
        "or_set_threads": (None, [_I]),
        "or_get_threads": (_I, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# matrix conversions: reference column-major (18 doubles) <-> numpy 3x3
# --------------------------------------------------------------------------
def to_np(m18: np.ndarray) -> np.ndarray:
    """(..., 18) column-major re/im -> (..., 3, 3) complex, [row, col]."""
    z = m18[..., 0::2] + 1j * m18[..., 1::2]
    return np.swapaxes(z.reshape(z.shape[:-1] + (3, 3)), -1, -2)


def from_np(m: np.ndarray) -> np.ndarray:
    """(..., 3, 3) complex [row, col] -> (..., 18) column-major re/im."""
    z = np.swapaxes(np.asarray(m, dtype=np.complex128), -1, -2)
    z = z.reshape(z.shape[:-2] + (9,))
    out = np.empty(z.shape[:-1] + (18,), dtype=np.float64)
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


class OracleLattice:
    """A ``LatticeCyclic`` (src/lattice.rs:44-49) with one extent per direction."""

    def __init__(self, dim: int, extent, size: float = 1.0):
        if np.isscalar(extent):
            extent = [int(extent)] * dim
        self.c = Lattice()
        arr = (C.c_size_t * MAX_DIM)(*([int(e) for e in extent] + [0] * (MAX_DIM - len(extent))))
        rc = lib().or_lattice_init(C.byref(self.c), dim, arr, float(size))
        if rc != 0:
            raise ValueError(f"or_lattice_init failed: {rc}")
        self.dim = dim
        self.extent = [int(e) for e in extent]
        self.nsite = int(self.c.nsite)
        self.nlink = self.nsite * dim
        self.size = float(size)

    @property
    def ref(self):
        return C.byref(self.c)

    # ---- geometry ----
    def site_index(self, x) -> int:
        arr = (C.c_size_t * MAX_DIM)(*([int(v) for v in x] + [0] * (MAX_DIM - len(x))))
        return int(lib().or_site_index(self.ref, arr))

    def site_coords(self, s: int):
        arr = (C.c_size_t * MAX_DIM)()
        lib().or_site_coords(self.ref, s, arr)
        return [int(arr[i]) for i in range(self.dim)]

    def neighbor(self, s: int, d: int, positive: bool) -> int:
        return int(lib().or_neighbor(self.ref, s, d, int(positive)))

    def parity(self, s: int) -> int:
        return int(lib().or_site_parity(self.ref, s))

    def link_order(self):
        sites = np.zeros(self.nlink, dtype=np.uint64)
        dirs = np.zeros(self.nlink, dtype=np.int32)
        lib().or_link_order(self.ref, _ptr(sites), _ptr(dirs))
        return sites, dirs

    # ---- inputs ----
    def hot_start(self, seed: int = SEED_RNG) -> np.ndarray:
        u = np.zeros((self.nlink, 18))
        lib().or_hot_start(self.ref, seed, _ptr(u))
        return u

    def cold_start(self) -> np.ndarray:
        u = np.zeros((self.nlink, 18))
        lib().or_cold_start(self.ref, _ptr(u))
        return u

    def random_efield(self, seed: int, sigma: float) -> np.ndarray:
        e = np.zeros((self.nlink, 8))
        lib().or_random_efield(self.ref, seed, sigma, _ptr(e))
        return e

    def device_hot_start(self, seed: int) -> np.ndarray:
        u = np.zeros((self.nlink, 18))
        lib().or_device_hot_start(self.ref, seed, _ptr(u))
        return u

    def device_momenta(self, seed: int, counter: int, sigma: float) -> np.ndarray:
        e = np.zeros((self.nlink, 8))
        lib().or_device_momenta(self.ref, seed, counter, sigma, _ptr(e))
        return e

    def device_mh_inputs(self, seed: int, counter: int, spread: float, n_update: int = 1):
        """Proposals (n_update, nlink, 18) and accept words (nlink,) of the
        device Metropolis stream."""
        p = np.zeros((n_update, self.nlink, 18))
        u = np.zeros(self.nlink, dtype=np.uint64)
        lib().or_device_mh_inputs(self.ref, seed, counter, spread, n_update, _ptr(p), _ptr(u))
        return p, u

    def mh_proposals(self, seed: int, spread: float, n_update: int = 1) -> np.ndarray:
        p = np.zeros((n_update, self.nlink, 18))
        lib().or_mh_proposals(self.ref, seed, spread, n_update, _ptr(p))
        return p

    # ---- stencils / observables ----
    def link_matrix(self, u, s, d, positive=True):
        out = np.zeros(18)
        lib().or_link_matrix(self.ref, _ptr(u), s, d, int(positive), _ptr(out))
        return out

    def pij(self, u, s, di, pi, dj, pj):
        out = np.zeros(18)
        lib().or_pij(self.ref, _ptr(u), s, di, int(pi), dj, int(pj), _ptr(out))
        return out

    def sij(self, u, s, di, pi, dj, pj):
        out = np.zeros(18)
        lib().or_sij(self.ref, _ptr(u), s, di, int(pi), dj, int(pj), _ptr(out))
        return out

    def clover(self, u, s, di, dj):
        out = np.zeros(18)
        lib().or_clover(self.ref, _ptr(u), s, di, dj, _ptr(out))
        return out

    def f_mu_nu(self, u, s, di, dj):
        out = np.zeros(18)
        lib().or_f_mu_nu(self.ref, _ptr(u), s, di, dj, _ptr(out))
        return out

    def magnetic_field(self, u, s, d):
        out = np.zeros(18)
        lib().or_magnetic_field(self.ref, _ptr(u), s, d, _ptr(out))
        return out

    def clover_field(self, u):
        """[site][i][j] clovers, (nsite, D, D, 18)"""
        out = np.zeros((self.nsite, self.dim, self.dim, 18))
        lib().or_clover_field(self.ref, _ptr(u), _ptr(out))
        return out

    def f_mu_nu_field(self, u):
        """[site][i][j] F_ij, (nsite, D, D, 18)"""
        out = np.zeros((self.nsite, self.dim, self.dim, 18))
        lib().or_f_mu_nu_field(self.ref, _ptr(u), _ptr(out))
        return out

    def magnetic_field_vec(self, u):
        """[site][dir] B_dir, (nsite, D, 18)"""
        out = np.zeros((self.nsite, self.dim, 18))
        lib().or_magnetic_field_vec(self.ref, _ptr(u), _ptr(out))
        return out

    def average_trace_plaquette(self, u) -> complex:
        c = lib().or_average_trace_plaquette(self.ref, _ptr(u))
        return complex(c.re, c.im)

    def hamiltonian_links(self, u, beta) -> float:
        return float(lib().or_hamiltonian_links(self.ref, _ptr(u), beta))

    def hamiltonian_efield(self, e, beta) -> float:
        return float(lib().or_hamiltonian_efield(self.ref, _ptr(e), beta))

    def derivative_e(self, u) -> np.ndarray:
        out = np.zeros((self.nlink, 8))
        lib().or_derivative_e_all(self.ref, _ptr(u), _ptr(out))
        return out

    def derivative_u(self, u, e, s, d) -> np.ndarray:
        out = np.zeros(18)
        lib().or_derivative_u(self.ref, _ptr(u), _ptr(e), s, d, _ptr(out))
        return out

    def wilson_backward(self, u, beta) -> np.ndarray:
        out = np.zeros((self.nlink, 8))
        lib().or_wilson_backward(self.ref, _ptr(u), beta, _ptr(out))
        return out

    def staple(self, u, s, d):
        out = np.zeros(18)
        lib().or_staple(self.ref, _ptr(u), s, d, _ptr(out))
        return out

    def delta_s(self, u, s, d, new_link, beta) -> float:
        nl = np.ascontiguousarray(new_link, dtype=np.float64)
        return float(lib().or_delta_s(self.ref, _ptr(u), s, d, _ptr(nl), beta))

    # ---- integrators ----
    def integrate_efield(self, u, e, dt):
        out = np.zeros_like(e)
        lib().or_integrate_efield(self.ref, _ptr(u), _ptr(e), dt, _ptr(out))
        return out

    def integrate_link(self, u, e, dt, mode=DRIFT_ADDITIVE):
        out = np.zeros_like(u)
        lib().or_integrate_link(self.ref, _ptr(u), _ptr(e), dt, mode, _ptr(out))
        return out

    # SymplecticEulerRayon's transitions (src/integrator/symplectic_euler_rayon.rs),
    # each composed from integrate_link / integrate_efield in the reference's order
    def integrate_sync_sync(self, u, e, dt, mode=DRIFT_ADDITIVE):
        """:127-144: links and momenta both from the state before the step."""
        return self.integrate_link(u, e, dt, mode), self.integrate_efield(u, e, dt)

    def integrate_sync_leap(self, u, e, dt):
        """:173-193: half a momentum step, links unchanged."""
        return np.array(u, copy=True), self.integrate_efield(u, e, dt / 2.0)

    def integrate_leap_leap(self, u, e, dt, mode=DRIFT_ADDITIVE):
        """:146-171: links first, then momenta from the new links."""
        u2 = self.integrate_link(u, e, dt, mode)
        return u2, self.integrate_efield(u2, e, dt)

    def integrate_leap_sync(self, u, e, dt, mode=DRIFT_ADDITIVE):
        """:195-220: links, then half a momentum step from the new links."""
        u2 = self.integrate_link(u, e, dt, mode)
        return u2, self.integrate_efield(u2, e, dt / 2.0)

    def simulate_symplectic_n(self, u, e, dt, n, mode=DRIFT_ADDITIVE, scheme=SCHEME_REFERENCE):
        u = np.array(u, copy=True)
        e = np.array(e, copy=True)
        rc = lib().or_simulate_symplectic_n(self.ref, _ptr(u), _ptr(e), dt, n, mode, scheme)
        if rc != 0:
            raise ValueError("ZeroIntegration")
        return u, e

    def hmc_trajectory(self, u, momenta, beta, dt, n, uniform, mode=DRIFT_ADDITIVE,
                       scheme=SCHEME_REFERENCE):
        u = np.array(u, copy=True)
        st = HmcStats()
        rc = lib().or_hmc_trajectory(self.ref, _ptr(u), _ptr(np.ascontiguousarray(momenta)), beta,
                                     dt, n, mode, scheme, C.c_uint64(uniform), C.byref(st))
        if rc != 0:
            raise ValueError("ZeroIntegration")
        return u, st

    def metropolis_sweep(self, u, beta, proposals, uniforms, order=ORDER_CHECKERBOARD):
        u = np.array(u, copy=True)
        st = MhStats()
        proposals = np.ascontiguousarray(proposals, dtype=np.float64)
        uniforms = np.ascontiguousarray(uniforms, dtype=np.uint64)
        rc = lib().or_metropolis_sweep(self.ref, _ptr(u), beta, proposals.shape[0],
                                       _ptr(proposals), _ptr(uniforms), order, C.byref(st))
        if rc != 0:
            raise ValueError(f"or_metropolis_sweep failed: {rc}")
        return u, st

    def heat_bath_sweep(self, u, beta, seed, counter, order=ORDER_CHECKERBOARD):
        """The device heat-bath stream's sweep; returns (links, exhausted samplers)."""
        u = np.array(u, copy=True)
        ex = C.c_size_t(0)
        rc = lib().or_heat_bath_sweep(self.ref, _ptr(u), beta, seed, counter, order, C.byref(ex))
        if rc != 0:
            raise ValueError(f"or_heat_bath_sweep failed: {rc}")
        return u, int(ex.value)

    def overrelaxation_sweep(self, u, order=ORDER_CHECKERBOARD):
        u = np.array(u, copy=True)
        rc = lib().or_overrelaxation_sweep(self.ref, _ptr(u), order)
        if rc != 0:
            raise ValueError(f"or_overrelaxation_sweep failed: {rc}")
        return u

    def overrelaxation_reverse_sweep(self, u, order=ORDER_CHECKERBOARD):
        u = np.array(u, copy=True)
        rc = lib().or_overrelaxation_reverse_sweep(self.ref, _ptr(u), order)
        if rc != 0:
            raise ValueError(f"or_overrelaxation_reverse_sweep failed: {rc}")
        return u

    def normalize_link_matrices(self, u):
        u = np.array(u, copy=True)
        lib().or_normalize_link_matrices(self.ref, _ptr(u))
        return u

    def gauss(self, u, e, s):
        out = np.zeros(18)
        lib().or_gauss(self.ref, _ptr(u), _ptr(e), s, _ptr(out))
        return out

    def gauss_sum_div(self, u, e) -> float:
        return float(lib().or_gauss_sum_div(self.ref, _ptr(u), _ptr(e)))

    def project_to_gauss(self, u, e, max_iter=1000):
        out = np.zeros_like(e)
        it = C.c_size_t(0)
        rc = lib().or_project_to_gauss(self.ref, _ptr(u), _ptr(e), _ptr(out), max_iter,
                                       C.byref(it))
        return rc, out, int(it.value)


# ---- su(3) helpers on single matrices ----
def generator(a: int) -> np.ndarray:
    out = np.zeros(18)
    lib().or_generator(a, _ptr(out))
    return out


def adjoint_to_matrix(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(18)
    lib().or_adjoint_to_matrix(_ptr(v), _ptr(out))
    return out


def su3_exp_i(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(18)
    lib().or_su3_exp_i(_ptr(v), _ptr(out))
    return out


def su3_exp_r(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(18)
    lib().or_su3_exp_r(_ptr(v), _ptr(out))
    return out


def orthonormalize(m18) -> np.ndarray:
    m = np.ascontiguousarray(m18, dtype=np.float64)
    out = np.zeros(18)
    lib().or_orthonormalize(_ptr(m), _ptr(out))
    return out


def adjoint_d(v) -> complex:
    v = np.ascontiguousarray(v, dtype=np.float64)
    c = lib().or_adjoint_d(_ptr(v))
    return complex(c.re, c.im)


def trace_squared(v) -> float:
    v = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().or_trace_squared(_ptr(v)))


def bernoulli_accept(p: float, u: int) -> bool:
    return bool(lib().or_bernoulli_accept(p, C.c_uint64(u)))


def clamp_probability(x: float) -> float:
    return float(lib().or_clamp_probability(x))


def replacement_probability(h_old: float, h_new: float) -> float:
    """(h_old - h_new).exp().min(1).max(0) of hybrid_monte_carlo.rs:584-589
    with f64 exp semantics: an overflow is +inf (so p = 1), NaN stays NaN
    (so p = 1 through Rust's min)."""
    try:
        e = math.exp(h_old - h_new)
    except OverflowError:
        e = math.inf
    return clamp_probability(e)


def factorial_size_for_exp() -> int:
    return int(lib().or_factorial_size_for_exp())


def su2_project_unorm(m2) -> np.ndarray:
    """su2::project_to_su2_unorm of a 2x2 complex matrix."""
    a = np.ascontiguousarray(np.asarray(m2, dtype=np.complex128).reshape(4))
    v = np.ascontiguousarray(np.stack([a.real, a.imag], axis=1).reshape(8))
    out = np.zeros(8)
    lib().or_su2_project_unorm(_ptr(v), _ptr(out))
    return (out[0::2] + 1j * out[1::2]).reshape(2, 2)


def heat_bath_x0(p: float, l: int, n: int, seed: int) -> float:
    return float(lib().or_heat_bath_x0(p, l, n, seed))


def polar_unitary(m18) -> np.ndarray:
    m = np.ascontiguousarray(m18, dtype=np.float64)
    out = np.zeros(18)
    lib().or_polar_unitary(_ptr(m), _ptr(out))
    return out


def levi_civita(index) -> int:
    idx = np.ascontiguousarray(index, dtype=np.int32)
    return int(lib().or_levi_civita(_ptr(idx), len(idx)))


def svd3(m18):
    """(U, singular values, V) of a 3x3 complex matrix (column-major 18)."""
    m = np.ascontiguousarray(m18, dtype=np.float64)
    u = np.zeros(18)
    v = np.zeros(18)
    sv = np.zeros(3)
    lib().or_svd3(_ptr(m), _ptr(u), _ptr(sv), _ptr(v))
    return u, sv, v


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32(_ptr(c), _ptr(k), _ptr(out))
    return out


class OracleRng:
    """xoshiro256++ stream with the reference's samplers (src/su3.rs, src/su2.rs)."""

    def __init__(self, seed: int):
        self.c = Rng()
        lib().or_rng_seed(C.byref(self.c), C.c_uint64(seed))

    def next_u64(self) -> int:
        return int(lib().or_rng_next(C.byref(self.c)))

    def uniform(self, low=-1.0, high=1.0) -> float:
        return float(lib().or_rng_uniform(C.byref(self.c), low, high))

    def random_su3(self) -> np.ndarray:
        out = np.zeros(18)
        lib().or_random_su3(C.byref(self.c), _ptr(out))
        return out

    def random_su3_close_to_unity(self, spread: float) -> np.ndarray:
        out = np.zeros(18)
        lib().or_random_su3_close_to_unity(C.byref(self.c), spread, _ptr(out))
        return out

    def random_su2_close_to_unity(self, spread: float) -> np.ndarray:
        out = np.zeros(8)
        lib().or_random_su2_close_to_unity(C.byref(self.c), spread, _ptr(out))
        z = out[0::2] + 1j * out[1::2]
        return z.reshape(2, 2)


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


def get_threads() -> int:
    return int(lib().or_get_threads())


# --------------------------------------------------------------------------
# statistics (src/statistics/mod.rs): float64 or complex128 1-D arrays
# --------------------------------------------------------------------------
def _stats_in(x):
    x = np.asarray(x)
    cplx = np.iscomplexobj(x)
    flat = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64).reshape(-1)
    return flat, int(cplx)


def _stats_out(v: np.ndarray, cplx: int):
    return complex(v[0], v[1]) if cplx else float(v[0])


def stats_mean(x):
    """statistics::mean (mod.rs:321-335)."""
    a, c = _stats_in(x)
    out = np.zeros(2)
    lib().or_stats_mean(_ptr(a), a.size, c, _ptr(out))
    return _stats_out(out, c)


def stats_mean_and_variance(x, parallel: bool = False):
    """[mean, variance]: two-pass mean_and_variance (mod.rs:375-399) or the
    one-pass mean_and_variance_par_iter_val (mod.rs:163-181)."""
    a, c = _stats_in(x)
    out = np.zeros(4)
    f = lib().or_stats_mean_and_variance_par if parallel else lib().or_stats_mean_and_variance
    if f(_ptr(a), a.size, c, _ptr(out)) != 0:
        raise ValueError("variance of an empty sequence")
    k = 2 if c else 1
    return [_stats_out(out[:k], c), _stats_out(out[k:], c)]


def stats_covariance(x, y, parallel: bool = False):
    """covariance (mod.rs:427-458) or covariance_par_iter_val (mod.rs:279-306);
    None when the lengths differ."""
    a, c = _stats_in(x)
    b, c2 = _stats_in(y)
    if c != c2:
        a = a.astype(np.complex128)
        b = b.astype(np.complex128)
        c = 1
    out = np.zeros(2)
    f = lib().or_stats_covariance_par if parallel else lib().or_stats_covariance
    if not f(_ptr(a), a.size, _ptr(b), b.size, c, _ptr(out)):
        return None
    return _stats_out(out, c)
