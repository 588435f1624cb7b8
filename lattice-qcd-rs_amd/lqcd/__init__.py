"""Host-side mirror of lattice-qcd-rs's state / integrator / Monte Carlo traits
over the gfx950 library (include/lqcd.h).

The reference's API is Rust generics over ``const D``; here the dimension is a
runtime argument and the fields live in HBM behind a lattice handle:

=====================================  ==========================================
reference (lattice-qcd-rs)              here
=====================================  ==========================================
``LatticeCyclic::<D>::new(size, dim)``  ``LatticeCyclic(size, dim, d)``
``LatticeStateDefault<D>``              ``LatticeStateDefault`` (device resident)
``LatticeStateEFSyncDefault``           ``LatticeStateEFSyncDefault``
``SimulationStateLeap``                 ``SimulationStateLeap`` (E half a step ahead)
``SymplecticEulerRayon``                ``SymplecticEulerGpu``
``HybridMonteCarloDiagnostic``          ``HybridMonteCarloDiagnostic``
``MetropolisHastingsSweep``             ``MetropolisHastingsSweep`` (checkerboard)
``MetropolisHastingsDeltaDiagnostic``   ``MetropolisHastingsDeltaDiagnostic`` (device stream)
``HeatBathSweep``                       ``HeatBathSweep`` (checkerboard, device stream)
``OverrelaxationSweepRotation``         ``OverrelaxationSweepRotation`` (checkerboard)
``OverrelaxationSweepReverse``          ``OverrelaxationSweepReverse`` (checkerboard)
``HybridMethodVec``                     ``HybridMethodVec``
``HybridMethodCouple``                  ``HybridMethodCouple``
``HybridMonteCarlo``                    ``HybridMonteCarlo``
independent chains, one after another   ``BatchedChains`` (one launch per MD step)
``statistics::*``                       ``lqcd.statistics`` (sums on the device)
=====================================  ==========================================

Host arrays use the reference's layouts: links ``(nsite*D, 18)`` float64,
column-major 3x3 complex, link index ``site*D + mu``; E ``(nsite*D, 8)``.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence, Union

import numpy as np

from . import _ffi as F
from ._ffi import (DRIFT_ADDITIVE, DRIFT_EXP, F32, F64, SCHEME_LEAPFROG, SCHEME_REFERENCE,
                   LqcdError)

__all__ = [
    "Context", "LatticeCyclic", "Direction", "LatticeLink", "LatticeLinkCanonical", "LatticeStateDefault", "LatticeStateEFSyncDefault",
    "SymplecticEulerGpu", "SimulationStateLeap", "HybridMonteCarloDiagnostic", "MetropolisHastingsSweep",
    "BatchedChains", "ChainSpec", "HeatBathSweep", "OverrelaxationSweepRotation",
    "OverrelaxationSweepReverse", "HybridMethodVec", "HybridMethodCouple", "HybridMonteCarlo",
    "MetropolisHastingsDeltaDiagnostic",
    "LatticeInitializationError", "StateInitializationError", "MultiIntegrationError",
    "LqcdError", "F32", "F64", "DRIFT_ADDITIVE", "DRIFT_EXP", "SCHEME_REFERENCE",
    "SCHEME_LEAPFROG",
]


class LatticeInitializationError(ValueError):
    """src/error.rs:219-241 (ZeroDimension, NonPositiveSize, DimTooSmall)."""


class StateInitializationError(ValueError):
    """src/error.rs:124-171 (IncompatibleSize ...)."""


class MultiIntegrationError(RuntimeError):
    """src/error.rs:93-118 (ZeroIntegration, IntegrationError)."""


def _raise(rc: int) -> None:
    if rc == F.OK:
        return
    msg = F.lib().lqcd_last_error().decode(errors="replace")
    if rc in (F.ERR_ZERO_DIMENSION, F.ERR_NON_POSITIVE_SIZE, F.ERR_DIM_TOO_SMALL):
        raise LatticeInitializationError(msg)
    if rc in (F.ERR_INCOMPATIBLE_SIZE, F.ERR_GAUSS_PROJECTION):
        raise StateInitializationError(msg)
    if rc == F.ERR_ZERO_INTEGRATION:
        raise MultiIntegrationError(msg)
    raise LqcdError(rc, msg)


class Context:
    """One HIP stream on one device (lqcd_ctx)."""

    _default = {}

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _raise(F.lib().lqcd_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def synchronize(self) -> None:
        _raise(F.lib().lqcd_ctx_synchronize(self.handle))

    def profile_enable(self, mask: int) -> None:
        _raise(F.lib().lqcd_profile_enable(self.handle, mask))

    def profile_read(self, kind: int):
        n = C.c_size_t(0)
        t = C.c_double(0)
        _raise(F.lib().lqcd_profile_read(self.handle, kind, C.byref(n), C.byref(t)))
        return int(n.value), float(t.value)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                F.lib().lqcd_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Direction:
    """``Direction<D>`` (src/lattice.rs:1110-1290): an axis index and a sign."""

    __slots__ = ("_index", "_positive")

    def __init__(self, index: int, is_positive: bool):
        self._index = int(index)
        self._positive = bool(is_positive)

    def index(self) -> int:
        return self._index

    def is_positive(self) -> bool:
        return self._positive

    def is_negative(self) -> bool:
        return not self._positive

    def to_positive(self) -> "Direction":
        return Direction(self._index, True)

    def __neg__(self) -> "Direction":
        return Direction(self._index, not self._positive)

    def __eq__(self, other) -> bool:
        return isinstance(other, Direction) and (self._index, self._positive) == (other._index, other._positive)

    def __hash__(self) -> int:
        return hash((self._index, self._positive))

    def __repr__(self) -> str:
        return f"Direction({self._index}, {self._positive})"

    @staticmethod
    def positive_directions(d: int):
        """``positive_directions`` / ``positives_vec`` (:1139-1145, 1183)."""
        return [Direction(i, True) for i in range(d)]

    @staticmethod
    def negative_directions(d: int):
        return [Direction(i, False) for i in range(d)]

    @staticmethod
    def directions_vec(d: int):
        """``directions_vec`` (:1149-1156): +0, -0, +1, -1, ..."""
        return [Direction(i, pos) for i in range(d) for pos in (True, False)]


class LatticeLink:
    """``LatticeLink<D>`` (src/lattice.rs:1060-1110): a point and any direction."""

    __slots__ = ("_pos", "_dir")

    def __init__(self, pos, direction: Direction):
        self._pos = tuple(int(x) for x in pos)
        self._dir = direction

    def pos(self):
        return self._pos

    def dir(self) -> Direction:
        return self._dir

    def is_dir_positive(self) -> bool:
        return self._dir.is_positive()

    def is_dir_negative(self) -> bool:
        return self._dir.is_negative()

    def __eq__(self, other) -> bool:
        return type(other) is type(self) and (self._pos, self._dir) == (other._pos, other._dir)

    def __hash__(self) -> int:
        return hash((type(self).__name__, self._pos, self._dir))

    def __repr__(self) -> str:
        return f"{type(self).__name__}({list(self._pos)}, {self._dir!r})"


class LatticeLinkCanonical(LatticeLink):
    """``LatticeLinkCanonical<D>`` (src/lattice.rs:942-1058): a link with a
    positive direction; ``new`` returns None for a negative one."""

    @classmethod
    def new(cls, pos, direction: Direction) -> Optional["LatticeLinkCanonical"]:
        return cls(pos, direction) if direction.is_positive() else None

    def to_index(self, lattice: "LatticeCyclic") -> int:
        """position index * D + direction (:925-929): the storage index of the link."""
        return lattice.index_of(self._pos) * lattice.dim_st() + self._dir.index()


class LatticeCyclic:
    """``LatticeCyclic::<D>::new(size, dim)`` (src/lattice.rs:190-201).

    ``dim`` is the number of points per direction; a sequence gives one extent
    per direction (the reference is hypercubic: all equal).
    """

    def __init__(self, size: float, dim: Union[int, Sequence[int]], d: int = 4):
        if d == 0:
            raise LatticeInitializationError("the lattice has zero dimension")
        if not (size > 0) or math.isinf(size):
            raise LatticeInitializationError("the lattice spacing must be positive and finite")
        ext = [int(dim)] * d if np.isscalar(dim) else [int(x) for x in dim]
        if len(ext) != d:
            raise ValueError("one extent per direction")
        if any(x < 2 for x in ext):
            raise LatticeInitializationError("every extent must be at least 2")
        self._size = float(size)
        self._extent = ext
        self._d = d

    def dim_st(self) -> int:
        return self._d

    def dim(self) -> int:
        return self._extent[0]

    def extents(self):
        return list(self._extent)

    def size(self) -> float:
        return self._size

    def number_of_points(self) -> int:
        return int(np.prod(self._extent))

    def number_of_canonical_links_space(self) -> int:
        return self.number_of_points() * self._d

    def index_of(self, point) -> int:
        """Site index sum_i (x_i mod L_i) prod_{k<i} L_k (``to_index``,
        src/lattice.rs:909-929), coordinates taken modulo the extents."""
        if len(point) != self._d:
            raise ValueError("one coordinate per direction")
        idx, stride = 0, 1
        for x, n in zip(point, self._extent):
            idx += (int(x) % n) * stride
            stride *= n
        return idx

    def point_of(self, index: int):
        """The point of site ``index`` (the inverse of ``index_of``)."""
        out = []
        for n in self._extent:
            out.append(index % n)
            index //= n
        return tuple(out)

    def add_point_direction_n(self, point, direction: Direction, shift_number: int):
        """``add_point_direction_n`` (src/lattice.rs:303-323): ``shift_number``
        steps along ``direction``, periodic."""
        p = [int(x) for x in point]
        i = direction.index()
        n = self._extent[i]
        k = int(shift_number) % n
        if direction.is_positive():
            p[i] = (p[i] + k) % n
        elif p[i] < k:
            p[i] = n - (k - p[i])
        else:
            p[i] = (p[i] - k) % n
        return tuple(p)

    def add_point_direction(self, point, direction: Direction):
        """``add_point_direction`` (src/lattice.rs:273-279)."""
        return self.add_point_direction_n(point, direction, 1)

    def link(self, pos, direction: Direction) -> LatticeLink:
        """``link`` (src/lattice.rs:119-125): the position taken modulo the extents."""
        return LatticeLink((int(x) % n for x, n in zip(pos, self._extent)), direction)

    def link_canonical(self, pos, direction: Direction) -> LatticeLinkCanonical:
        """``link_canonical`` (src/lattice.rs:84-97): a negative link becomes the
        positive one from the neighbouring point."""
        p = pos
        if direction.is_negative():
            p = self.add_point_direction(p, direction)
        return LatticeLinkCanonical(tuple(int(x) % n for x, n in zip(p, self._extent)), direction.to_positive())

    def into_canonical(self, link: LatticeLink) -> LatticeLinkCanonical:
        """``into_canonical`` (src/lattice.rs:130-132)."""
        return self.link_canonical(link.pos(), link.dir())

    def get_points(self):
        """``get_points`` (src/lattice.rs:155-165): every point in storage
        order, x_0 fastest."""
        for i in range(self.number_of_points()):
            yield self.point_of(i)

    def get_links(self):
        """``get_links`` (src/lattice.rs:175-183): every canonical link in
        storage order, the direction fastest."""
        dirs = Direction.positive_directions(self._d)
        for p in self.get_points():
            for d in dirs:
                yield LatticeLinkCanonical(p, d)

    def has_compatible_length_links(self, links) -> bool:
        """``has_compatible_length_links`` (src/lattice.rs:326-328), links as
        (n_links, 18)."""
        return len(links) == self.number_of_canonical_links_space()

    def has_compatible_length_e_field(self, e_field) -> bool:
        """``has_compatible_length_e_field`` (src/lattice.rs:331-333), E as
        (nsite, D, 8) or (nsite * D, 8)."""
        e = np.asarray(e_field)
        return e.shape[0] in (self.number_of_points(), self.number_of_canonical_links_space()) and \
            (e.size == self.number_of_canonical_links_space() * 8)

    def has_compatible_length(self, links, e_field) -> bool:
        return self.has_compatible_length_links(links) and self.has_compatible_length_e_field(e_field)


class _DeviceLattice:
    """Owner of one lqcd_lattice handle (links, E field, work buffers)."""

    def __init__(self, lattice: LatticeCyclic, precision: int, ctx: Optional[Context]):
        self.ctx = ctx or Context.default()
        self.lattice = lattice
        ext = (C.c_size_t * lattice.dim_st())(*lattice.extents())
        h = C.c_void_p()
        _raise(F.lib().lqcd_lattice_create(self.ctx.handle, lattice.dim_st(), ext, lattice.size(),
                                           precision, C.byref(h)))
        self.handle = h
        self.precision = precision
        self.nlink = lattice.number_of_canonical_links_space()

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                F.lib().lqcd_lattice_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _read_header(path, format, with_e, d):
    size, points, beta, t = C.c_double(), C.c_size_t(), C.c_double(), C.c_uint64()
    _raise(F.lib().lqcd_state_read_header(str(path).encode(), format, int(with_e), d, C.byref(size),
                                          C.byref(points), C.byref(beta), C.byref(t)))
    if with_e:
        return size.value, points.value, beta.value, t.value
    return size.value, points.value, beta.value


class LatticeStateDefault:
    """``LatticeStateDefault<D>`` (src/simulation/state.rs:653-850), resident in HBM.

    Implements the ``LatticeState`` / ``LatticeStateNew`` surface:
    ``link_matrix``, ``set_link_matrix``, ``lattice``, ``beta``, ``CA``,
    ``hamiltonian_links``, ``monte_carlo_step``, ``average_trace_plaquette``.
    """

    CA = 3.0

    def __init__(self, lattice: LatticeCyclic, beta: float, link_matrix=None,
                 precision: int = F64, ctx: Optional[Context] = None):
        self._dev = _DeviceLattice(lattice, precision, ctx)
        self._beta = float(beta)
        if link_matrix is not None:
            self.set_link_matrix(link_matrix)

    # LatticeStateDefault::new_cold, src/simulation/state.rs:662-671
    @classmethod
    def new_cold(cls, size: float, beta: float, number_of_points, d: int = 4,
                 precision: int = F64, ctx: Optional[Context] = None):
        return cls(LatticeCyclic(size, number_of_points, d), beta, None, precision, ctx)

    # LatticeStateDefault::new_determinist, src/simulation/state.rs:706-715,
    # with the random SU(3) links drawn on the device from ``seed``
    @classmethod
    def new_determinist(cls, size: float, beta: float, number_of_points, seed: int, d: int = 4,
                        precision: int = F64, ctx: Optional[Context] = None):
        st = cls(LatticeCyclic(size, number_of_points, d), beta, None, precision, ctx)
        _raise(F.lib().lqcd_links_random(st.handle, C.c_uint64(seed)))
        return st

    def set_md_graph(self, enable: bool) -> None:
        """Replay leapfrog MD launches from a captured graph (default) or launch
        them one by one (``lqcd_md_graph_set``); the results are the same."""
        _raise(F.lib().lqcd_md_graph_set(self.handle, 1 if enable else 0))

    def set_md_reconstruct(self, mode: int) -> None:
        """1 (default): exp-drift integrations read SU(3) links as two rows;
        0: always nine entries."""
        _raise(F.lib().lqcd_md_reconstruct_set(self.handle, mode))

    def md_reconstruct_last(self) -> bool:
        mode, used = C.c_int(), C.c_int()
        _raise(F.lib().lqcd_md_reconstruct_get(self.handle, C.byref(mode), C.byref(used)))
        return bool(used.value)

    def su3_distance(self) -> float:
        """max over links of max(|U U^+ - 1|_ij, |det U - 1|)"""
        out = C.c_double()
        _raise(F.lib().lqcd_su3_distance(self.handle, C.byref(out)))
        return float(out.value)

    # serde interchange (Serialize / Deserialize derives, state.rs:653-659):
    # bincode 1.x or serde_json of { lattice, beta, link_matrix }
    def write(self, path: str, format: int = F.FORMAT_BINCODE) -> None:
        _raise(F.lib().lqcd_state_write(self.handle, self._beta, str(path).encode(), format))

    @classmethod
    def read(cls, path: str, d: int = 4, format: int = F.FORMAT_BINCODE, precision: int = F64,
             ctx: Optional[Context] = None):
        size, points, beta = _read_header(path, format, False, d)
        st = cls(LatticeCyclic(size, points, d), beta, None, precision, ctx)
        b = C.c_double()
        _raise(F.lib().lqcd_state_read(st.handle, str(path).encode(), format, 0, C.byref(b), None))
        return st

    @property
    def handle(self):
        return self._dev.handle

    def lattice(self) -> LatticeCyclic:
        return self._dev.lattice

    def beta(self) -> float:
        return self._beta

    def precision(self) -> int:
        return self._dev.precision

    def link_matrix(self) -> np.ndarray:
        out = np.empty((self._dev.nlink, 18), dtype=np.float64)
        _raise(F.lib().lqcd_links_download(self.handle, F.ptr(out), self._dev.nlink))
        return out

    def set_link_matrix(self, links) -> None:
        a = np.ascontiguousarray(links, dtype=np.float64)
        if a.ndim != 2 or a.shape[1] != 18:
            raise StateInitializationError("links must have shape (nsite*D, 18)")
        _raise(F.lib().lqcd_links_upload(self.handle, F.ptr(a), a.shape[0]))

    def hamiltonian_links(self) -> float:
        out = C.c_double()
        _raise(F.lib().lqcd_hamiltonian_links(self.handle, self._beta, C.byref(out)))
        return out.value

    def average_trace_plaquette(self) -> complex:
        out = np.zeros(2)
        _raise(F.lib().lqcd_average_trace_plaquette(self.handle, F.ptr(out)))
        return complex(out[0], out[1])

    def normalize_link_matrices(self) -> None:
        _raise(F.lib().lqcd_normalize_link_matrices(self.handle))

    # clover / F_mu_nu / chromomagnetic field (LinkMatrix, src/field.rs:806-882),
    # computed on the device for every site; matrices as 18 doubles
    # (column-major 3x3 complex), like the links
    def clover_field(self) -> np.ndarray:
        """``clover(x, i, j)`` for every site and ordered pair, ``(nsite, D, D, 18)``."""
        return self._fields(F.lib().lqcd_clover_field, True)

    def f_mu_nu_field(self) -> np.ndarray:
        """``f_mu_nu(x, i, j)`` for every site and ordered pair, ``(nsite, D, D, 18)``."""
        return self._fields(F.lib().lqcd_f_mu_nu_field, True)

    def magnetic_field_vec_field(self) -> np.ndarray:
        """``magnetic_field_vec(x)`` for every site, ``(nsite, D, 18)``."""
        return self._fields(F.lib().lqcd_magnetic_field_vec, False)

    def sij_field(self, dir_i: "Direction", dir_j: "Direction") -> np.ndarray:
        """``LinkMatrix::sij`` (src/field.rs:743-758) at every site, ``(nsite, 18)``."""
        return self._plaquette_field(0, dir_i, dir_j)

    def pij_field(self, dir_i: "Direction", dir_j: "Direction") -> np.ndarray:
        """``LinkMatrix::pij`` (src/field.rs:761-771) at every site, ``(nsite, 18)``."""
        return self._plaquette_field(1, dir_i, dir_j)

    def sij(self, point, dir_i: "Direction", dir_j: "Direction") -> np.ndarray:
        """``LinkMatrix::sij`` at one point (one element of ``sij_field``)."""
        return self.sij_field(dir_i, dir_j)[self._site(point)]

    def pij(self, point, dir_i: "Direction", dir_j: "Direction") -> np.ndarray:
        """``LinkMatrix::pij`` at one point (one element of ``pij_field``)."""
        return self.pij_field(dir_i, dir_j)[self._site(point)]

    def _plaquette_field(self, which: int, dir_i: "Direction", dir_j: "Direction") -> np.ndarray:
        n = self.lattice().number_of_points()
        out = np.empty((n, 18), dtype=np.float64)
        _raise(F.lib().lqcd_plaquette_field(self.handle, which, dir_i.index(), int(dir_i.is_positive()),
                                            dir_j.index(), int(dir_j.is_positive()), F.ptr(out), out.size))
        return out

    def _fields(self, fn, pairs: bool) -> np.ndarray:
        lat = self.lattice()
        d = lat.dim_st()
        shape = (lat.number_of_points(), d, d, 18) if pairs else (lat.number_of_points(), d, 18)
        out = np.empty(shape, dtype=np.float64)
        _raise(fn(self.handle, F.ptr(out), out.size))
        return out

    def _site(self, point) -> int:
        return self.lattice().index_of([int(c) for c in point])

    def clover(self, point, dir_i: int, dir_j: int) -> np.ndarray:
        """``LinkMatrix::clover`` (src/field.rs:806-820) at one point (coordinates
        taken modulo the extents), positive directions."""
        return self.clover_field()[self._site(point), dir_i, dir_j]

    def f_mu_nu(self, point, dir_i: int, dir_j: int) -> np.ndarray:
        """``LinkMatrix::f_mu_nu`` (src/field.rs:822-834) at one point."""
        return self.f_mu_nu_field()[self._site(point), dir_i, dir_j]

    def magnetic_field_vec(self, point) -> np.ndarray:
        """``LinkMatrix::magnetic_field_vec`` (src/field.rs:836-845) at one point, ``(D, 18)``."""
        return self.magnetic_field_vec_field()[self._site(point)]

    def magnetic_field(self, point, dir: int) -> np.ndarray:
        """``LinkMatrix::magnetic_field`` (src/field.rs:848-868) at one point."""
        return self.magnetic_field_vec(point)[dir]

    def magnetic_field_link(self, point, dir: int) -> np.ndarray:
        """``LinkMatrix::magnetic_field_link`` (src/field.rs:871-878): the field at
        the link's position along its direction."""
        return self.magnetic_field(point, dir)

    # MD work decomposition (0 per link, 1 per link XCD-grouped, 2 per site,
    # 3 per site with slab order)
    def md_variant(self) -> int:
        v = C.c_int()
        _raise(F.lib().lqcd_md_variant_get(self.handle, C.byref(v)))
        return v.value

    def set_md_variant(self, variant: int) -> None:
        _raise(F.lib().lqcd_md_variant_set(self.handle, variant))

    def md_autotune(self, drift: int = DRIFT_EXP) -> int:
        v = C.c_int()
        _raise(F.lib().lqcd_md_autotune(self.handle, drift, C.byref(v)))
        return v.value

    def monte_carlo_step(self, m):
        """``LatticeState::monte_carlo_step`` (src/simulation/state.rs:103-109)."""
        return m.next_element(self)

    # Wilson action as a forward / backward pair
    def wilson_forward(self) -> float:
        out = C.c_double()
        _raise(F.lib().lqcd_wilson_forward(self.handle, self._beta, C.byref(out)))
        return out.value

    def wilson_backward(self) -> np.ndarray:
        out = np.empty((self._dev.nlink, 8), dtype=np.float64)
        _raise(F.lib().lqcd_wilson_backward(self.handle, self._beta, F.ptr(out), self._dev.nlink))
        return out

    def wilson_backward_device(self, buf: "F.DeviceBuffer") -> "F.DeviceBuffer":
        """The backward pass into a device buffer of at least nlink * 8 doubles
        (``lqcd_wilson_backward_device``): the gradient stays on the GPU."""
        if buf.nbytes < self._dev.nlink * 8 * 8:
            raise ValueError("device buffer smaller than nlink * 8 doubles")
        _raise(F.lib().lqcd_wilson_backward_device(self.handle, self._beta, C.c_void_p(buf.ptr),
                                                   self._dev.nlink))
        return buf


class LatticeStateEFSyncDefault:
    """``LatticeStateEFSyncDefault<LatticeStateDefault<D>, D>``
    (src/simulation/state.rs:1048-1449): links plus E field on the device."""

    CA = 3.0

    def __init__(self, lattice_state: LatticeStateDefault, e_field=None, t: int = 0):
        self._state = lattice_state
        self._t = t
        if e_field is not None:
            self.set_e_field(e_field)
        else:
            _raise(F.lib().lqcd_efield_cold(self.handle))

    # LatticeStateEFSyncDefault::new_e_cold, src/simulation/state.rs:1110-1120
    @classmethod
    def new_e_cold(cls, lattice_state: LatticeStateDefault):
        return cls(lattice_state, None)

    @classmethod
    def new_cold(cls, size: float, beta: float, number_of_points, d: int = 4, precision: int = F64,
                 ctx: Optional[Context] = None):
        """``new_cold`` (src/simulation/state.rs:1221-1231): identity links, zero E."""
        return cls(LatticeStateDefault.new_cold(size, beta, number_of_points, d, precision, ctx), None)

    @classmethod
    def new_determinist_cold_e_hot_link(cls, size: float, beta: float, number_of_points, seed: int,
                                        d: int = 4, precision: int = F64, ctx: Optional[Context] = None):
        """``new_determinist_cold_e_hot_link`` (src/simulation/state.rs:1197-1211):
        random SU(3) links drawn on the device from ``seed``, zero E."""
        st = LatticeStateDefault.new_determinist(size, beta, number_of_points, seed, d, precision, ctx)
        return cls(st, None)

    @classmethod
    def new_determinist(cls, size: float, beta: float, number_of_points, seed: int, distribution,
                        d: int = 4, precision: int = F64, ctx: Optional[Context] = None):
        """``new_determinist`` (src/simulation/state.rs:1175-1189): random SU(3)
        links drawn on the device from ``seed`` and E components drawn on the
        host from ``distribution`` (a callable taking the shape and returning
        float64 samples, the role of the reference's ``rand_distr`` argument);
        no Gauss projection, as in the reference."""
        st = LatticeStateDefault.new_determinist(size, beta, number_of_points, seed, d, precision, ctx)
        n = st.lattice().number_of_canonical_links_space()
        e = np.asarray(distribution((n, 8)), dtype=np.float64)
        return cls(st, e)

    @classmethod
    def new_random_e(cls, lattice: LatticeCyclic, beta: float, link_matrix, seed: int,
                     precision: int = F64, ctx: Optional[Context] = None):
        """``LatticeStateWithEFieldNew::new_random_e`` (src/simulation/state.rs:262-277):
        the given links, E ~ N(0, 0.5 / beta) drawn on the device and projected
        on Gauss's law."""
        if not (beta > 0) or math.isinf(beta):
            # Normal::new(0, 0.5 / beta) fails for these
            raise StateInitializationError("0.5 / beta is not a valid standard deviation")
        st = LatticeStateDefault(lattice, beta, link_matrix, precision, ctx)
        return cls.new_random_e_state(st, seed)

    # serde interchange (state.rs:1046-1062): { e_field, t, lattice_state }
    def write(self, path: str, format: int = F.FORMAT_BINCODE) -> None:
        _raise(F.lib().lqcd_state_ef_write(self.handle, self._state.beta(), C.c_uint64(self._t),
                                           str(path).encode(), format))

    @classmethod
    def read(cls, path: str, d: int = 4, format: int = F.FORMAT_BINCODE, precision: int = F64,
             ctx: Optional[Context] = None):
        size, points, beta, t = _read_header(path, format, True, d)
        st = LatticeStateDefault(LatticeCyclic(size, points, d), beta, None, precision, ctx)
        ef = cls(st, None, t)
        b, tt = C.c_double(), C.c_uint64()
        _raise(F.lib().lqcd_state_read(st.handle, str(path).encode(), format, 1, C.byref(b), C.byref(tt)))
        return ef

    @classmethod
    def new_random_e_state(cls, lattice_state: LatticeStateDefault, seed: int, counter: int = 0,
                           sigma: Optional[float] = None, project: bool = True,
                           max_iter: int = 1000):
        """``new_random_e_state`` (src/simulation/state.rs:1093-1108): momenta
        N(0, sigma) drawn on the device (sigma defaults to the reference's
        0.5 / beta), then ``project_to_gauss`` unless ``project`` is False."""
        st = cls(lattice_state, None)
        st.reset_e_field(seed, counter, sigma)
        if project:
            st.project_to_gauss(max_iter)
        return st

    def gauss(self) -> np.ndarray:
        """``EField::gauss`` (src/field.rs:1174-1195) at every site, ``(nsite, 18)``
        column-major 3x3 complex."""
        n = self.lattice().number_of_points()
        out = np.empty((n, 18))
        _raise(F.lib().lqcd_efield_gauss(self.handle, F.ptr(out), n))
        return out

    def gauss_sum_div(self) -> float:
        """``EField::gauss_sum_div`` (src/field.rs:1199-1220)."""
        out = C.c_double()
        _raise(F.lib().lqcd_efield_gauss_sum_div(self.handle, C.byref(out)))
        return float(out.value)

    def project_to_gauss(self, max_iter: int = 1000) -> int:
        """``EField::project_to_gauss`` (src/field.rs:1265-1337) in place; the
        reference's unbounded loop stops after ``max_iter`` checks with
        ``StateInitializationError`` (GaussProjectionError).  Returns the
        number of checks that failed before convergence."""
        it = C.c_size_t(0)
        _raise(F.lib().lqcd_efield_project_gauss(self.handle, max_iter, C.byref(it)))
        return int(it.value)

    @property
    def handle(self):
        return self._state.handle

    def state_owned(self) -> LatticeStateDefault:
        return self._state

    def lattice_state(self) -> LatticeStateDefault:
        return self._state

    def lattice(self) -> LatticeCyclic:
        return self._state.lattice()

    def beta(self) -> float:
        return self._state.beta()

    def t(self) -> int:
        return self._t

    def link_matrix(self) -> np.ndarray:
        return self._state.link_matrix()

    def set_link_matrix(self, links) -> None:
        self._state.set_link_matrix(links)

    def e_field(self) -> np.ndarray:
        n = self.lattice().number_of_canonical_links_space()
        out = np.empty((n, 8), dtype=np.float64)
        _raise(F.lib().lqcd_efield_download(self.handle, F.ptr(out), n))
        return out

    def set_e_field(self, e) -> None:
        a = np.ascontiguousarray(e, dtype=np.float64)
        if a.ndim != 2 or a.shape[1] != 8:
            raise StateInitializationError("e_field must have shape (nsite*D, 8)")
        _raise(F.lib().lqcd_efield_upload(self.handle, F.ptr(a), a.shape[0]))

    def reset_e_field(self, seed: int, counter: int = 0, sigma: Optional[float] = None) -> None:
        s = 0.5 / self.beta() if sigma is None else float(sigma)
        _raise(F.lib().lqcd_efield_random(self.handle, seed, counter, s))

    def hamiltonian_links(self) -> float:
        return self._state.hamiltonian_links()

    def hamiltonian_efield(self) -> float:
        out = C.c_double()
        _raise(F.lib().lqcd_hamiltonian_efield(self.handle, self.beta(), C.byref(out)))
        return out.value

    def hamiltonian_total(self) -> float:
        """``LatticeStateWithEField::hamiltonian_total`` (src/simulation/state.rs:229-231)."""
        return self.hamiltonian_links() + self.hamiltonian_efield()

    def average_trace_plaquette(self) -> complex:
        return self._state.average_trace_plaquette()

    def derivative_e(self) -> np.ndarray:
        """``derivative_e`` for every site, layout (nsite*D, 8)."""
        n = self.lattice().number_of_canonical_links_space()
        out = np.empty((n, 8), dtype=np.float64)
        _raise(F.lib().lqcd_derivative_e(self.handle, F.ptr(out), n))
        return out

    def derivative_u(self) -> np.ndarray:
        """``derivative_u`` (src/simulation/state.rs:1407-1417) of every link,
        layout (nsite*D, 18) like the links."""
        n = self.lattice().number_of_canonical_links_space()
        out = np.empty((n, 18), dtype=np.float64)
        _raise(F.lib().lqcd_derivative_u(self.handle, F.ptr(out), n))
        return out

    def simulate_symplectic(self, integrator: "SymplecticEulerGpu", delta_t: float):
        return integrator.integrate_symplectic(self, delta_t)

    def simulate_symplectic_n(self, integrator: "SymplecticEulerGpu", delta_t: float,
                              numbers_of_times: int):
        """``SimulationStateSynchronous::simulate_symplectic_n``
        (src/simulation/state.rs:470-492); advances the state in place."""
        if numbers_of_times == 0:
            raise MultiIntegrationError("no integration steps")
        _raise(F.lib().lqcd_simulate_symplectic_n(self.handle, delta_t, numbers_of_times,
                                                  integrator.drift, integrator.scheme))
        self._t += numbers_of_times
        return self

    def simulate_symplectic_n_auto(self, integrator: "SymplecticEulerGpu", delta_t: float,
                                   numbers_of_times: int):
        """``simulate_symplectic_n_auto`` (src/simulation/state.rs:500-511)."""
        return self.simulate_symplectic_n(integrator, delta_t, numbers_of_times)

    def simulate_sync(self, integrator: "SymplecticEulerGpu", delta_t: float):
        """``simulate_sync`` (src/simulation/state.rs:383-390): one
        ``integrate_sync_sync``; advances the state in place."""
        return integrator.integrate_sync_sync(self, delta_t)

    def simulate_sync_n(self, integrator: "SymplecticEulerGpu", delta_t: float, numbers_of_times: int):
        """``simulate_sync_n`` (src/simulation/state.rs:397-418)."""
        if numbers_of_times == 0:
            raise MultiIntegrationError("no integration steps")
        _raise(F.lib().lqcd_simulate_sync_n(self.handle, delta_t, numbers_of_times, integrator.drift))
        self._t += numbers_of_times
        return self

    def simulate_to_leapfrog(self, integrator: "SymplecticEulerGpu", delta_t: float) -> "SimulationStateLeap":
        """``simulate_to_leapfrog`` (src/simulation/state.rs:302-313): half a
        momentum step.  The returned leap state owns the device fields; this
        state is consumed."""
        return integrator.integrate_sync_leap(self, delta_t)

    def simulate_using_leapfrog_n(self, integrator: "SymplecticEulerGpu", delta_t: float,
                                  number_of_steps: int):
        """``simulate_using_leapfrog_n`` (src/simulation/state.rs:321-358):
        half step, number_of_steps - 1 leap steps, half step back, as one
        sequence of launches (``lqcd_simulate_symplectic_n`` with
        ``SCHEME_LEAPFROG``)."""
        if number_of_steps == 0:
            raise MultiIntegrationError("no integration steps")
        _raise(F.lib().lqcd_simulate_symplectic_n(self.handle, delta_t, number_of_steps,
                                                  integrator.drift, SCHEME_LEAPFROG))
        self._t += number_of_steps
        return self

    simulate_using_leapfrog_n_auto = simulate_using_leapfrog_n

    def monte_carlo_step(self, m):
        return m.next_element(self._state)


class SimulationStateLeap:
    """``SimulationStateLeap<State, D>`` (src/simulation/state.rs:856-1040):
    the state with its E half a step ahead of its links.  It shares the
    device fields of the synchronous state it was made from."""

    def __init__(self, sync_state: LatticeStateEFSyncDefault):
        self._sync = sync_state

    @property
    def handle(self):
        return self._sync.handle

    def lattice_state(self) -> LatticeStateDefault:
        return self._sync.lattice_state()

    def lattice(self) -> LatticeCyclic:
        return self._sync.lattice()

    def beta(self) -> float:
        return self._sync.beta()

    def t(self) -> int:
        return self._sync.t()

    def link_matrix(self) -> np.ndarray:
        return self._sync.link_matrix()

    def e_field(self) -> np.ndarray:
        return self._sync.e_field()

    def hamiltonian_links(self) -> float:
        return self._sync.hamiltonian_links()

    def hamiltonian_efield(self) -> float:
        return self._sync.hamiltonian_efield()

    def hamiltonian_total(self) -> float:
        return self._sync.hamiltonian_total()

    def simulate_leap(self, integrator: "SymplecticEulerGpu", delta_t: float) -> "SimulationStateLeap":
        """``simulate_leap`` (src/simulation/state.rs:577-584): one
        ``integrate_leap_leap``."""
        return integrator.integrate_leap_leap(self, delta_t)

    def simulate_leap_n(self, integrator: "SymplecticEulerGpu", delta_t: float,
                        numbers_of_times: int) -> "SimulationStateLeap":
        """``simulate_leap_n`` (src/simulation/state.rs:624-648)."""
        if numbers_of_times == 0:
            raise MultiIntegrationError("no integration steps")
        _raise(F.lib().lqcd_simulate_leap_n(self.handle, delta_t, numbers_of_times, integrator.drift))
        self._sync._t += numbers_of_times
        return self

    def simulate_to_synchronous(self, integrator: "SymplecticEulerGpu",
                                delta_t: float) -> LatticeStateEFSyncDefault:
        """``simulate_to_synchronous`` (src/simulation/state.rs:528-539): one
        ``integrate_leap_sync``."""
        return integrator.integrate_leap_sync(self, delta_t)


class SymplecticEulerGpu:
    """``SymplecticEulerRayon`` (src/integrator/symplectic_euler_rayon.rs) on the GPU.

    ``drift``: ``DRIFT_ADDITIVE`` (the reference's U + dt derivative_u) or
    ``DRIFT_EXP``; ``scheme``: ``SCHEME_REFERENCE`` (kick-drift-kick per step)
    or ``SCHEME_LEAPFROG`` (half kicks merged between steps).
    """

    def __init__(self, drift: int = DRIFT_ADDITIVE, scheme: int = SCHEME_REFERENCE):
        self.drift = drift
        self.scheme = scheme

    def integrate_symplectic(self, state: LatticeStateEFSyncDefault, delta_t: float):
        _raise(F.lib().lqcd_simulate_symplectic_n(state.handle, delta_t, 1, self.drift,
                                                  self.scheme))
        state._t += 1
        return state

    def integrate_efield(self, state, delta_t: float):
        _raise(F.lib().lqcd_integrate_efield(state.handle, delta_t))
        return state

    def integrate_link(self, state, delta_t: float):
        _raise(F.lib().lqcd_integrate_link(state.handle, delta_t, self.drift))
        return state

    # the four transitions of SymplecticIntegrator (src/integrator/mod.rs:93-208)
    # as SymplecticEulerRayon implements them (symplectic_euler_rayon.rs:127-220)
    def integrate_sync_sync(self, state: LatticeStateEFSyncDefault, delta_t: float):
        _raise(F.lib().lqcd_integrate_sync_sync(state.handle, delta_t, self.drift))
        state._t += 1
        return state

    def integrate_sync_leap(self, state: LatticeStateEFSyncDefault, delta_t: float) -> SimulationStateLeap:
        # the time counter does not advance
        _raise(F.lib().lqcd_integrate_sync_leap(state.handle, delta_t))
        return SimulationStateLeap(state)

    def integrate_leap_leap(self, state: SimulationStateLeap, delta_t: float) -> SimulationStateLeap:
        _raise(F.lib().lqcd_integrate_leap_leap(state.handle, delta_t, self.drift))
        state._sync._t += 1
        return state

    def integrate_leap_sync(self, state: SimulationStateLeap, delta_t: float) -> LatticeStateEFSyncDefault:
        _raise(F.lib().lqcd_integrate_leap_sync(state.handle, delta_t, self.drift))
        state._sync._t += 1
        return state._sync


def _splitmix64(x: int):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x, z ^ (z >> 31)


class HybridMonteCarloDiagnostic:
    """``HybridMonteCarloDiagnostic`` (src/simulation/monte_carlo/hybrid_monte_carlo.rs:316-472).

    ``next_element`` draws momenta N(0, 0.5/beta) on the device, projects them
    onto the Gauss-law surface as ``new_random_e_state`` does
    (src/simulation/state.rs:1093-1108; ``project_gauss=False`` skips it),
    integrates ``number_of_steps`` steps of ``delta_t`` and accepts with
    p = min(1, exp(H_old - H_new)).

    The reference's projection loop (``project_to_gauss``, field.rs:1265-1294)
    has no bound; here it stops with ``StateInitializationError`` after
    ``gauss_max_iter`` checks.  Its linear convergence slows with the volume
    (25-56 checks at 8^4, about 2,500 at 32^4), so the default bound,
    ``gauss_max_iter=None``, grows with the lattice: 1000 + sites / 256
    (5,096 at 32^4, 42,472 at 48^3 x 96).
    """

    def __init__(self, delta_t: float, number_of_steps: int, integrator: SymplecticEulerGpu,
                 seed: int, project_gauss: bool = True, gauss_max_iter: Optional[int] = None):
        self._dt = float(delta_t)
        self._n = int(number_of_steps)
        self._integrator = integrator
        self._project = bool(project_gauss)
        self._gauss_max_iter = None if gauss_max_iter is None else int(gauss_max_iter)
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._rng_state = self._seed
        self._counter = 0
        self._prob_replace_last = 0.0
        self._has_replace_last = False
        self.last_stats: Optional[F.HmcStats] = None

    def delta_t(self) -> float:
        return self._dt

    def gauss_max_iter(self, number_of_points: int) -> int:
        """The bound on the Gauss-projection checks for a lattice of
        ``number_of_points`` sites."""
        if self._gauss_max_iter is not None:
            return self._gauss_max_iter
        return 1000 + int(number_of_points) // 256

    def number_of_steps(self) -> int:
        return self._n

    def integrator(self) -> SymplecticEulerGpu:
        return self._integrator

    def prob_replace_last(self) -> float:
        return self._prob_replace_last

    def has_replace_last(self) -> bool:
        return self._has_replace_last

    def _next_u64(self) -> int:
        self._rng_state, v = _splitmix64(self._rng_state)
        return v

    def trajectory(self, state, uniform: int) -> F.HmcStats:
        """One trajectory from the momenta already in ``state``'s E field."""
        if self._n == 0:
            raise MultiIntegrationError("no integration steps")
        st = F.HmcStats()
        _raise(F.lib().lqcd_hmc_trajectory(state.handle, state.beta(), self._dt, self._n,
                                           self._integrator.drift, self._integrator.scheme,
                                           C.c_uint64(uniform), C.byref(st)))
        self._prob_replace_last = st.probability
        self._has_replace_last = bool(st.accepted)
        self.last_stats = st
        return st

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        sigma = 0.5 / state.beta()
        _raise(F.lib().lqcd_efield_random(state.handle, self._seed, self._counter, sigma))
        self._counter += 1
        if self._project:
            _raise(F.lib().lqcd_efield_project_gauss(
                state.handle, self.gauss_max_iter(state.lattice().number_of_points()), None))
        self.trajectory(state, self._next_u64())
        return state


class HybridMonteCarlo(HybridMonteCarloDiagnostic):
    """``HybridMonteCarlo`` (src/simulation/monte_carlo/hybrid_monte_carlo.rs:64-200):
    the trajectory of ``HybridMonteCarloDiagnostic``; the reference's plain
    type only lacks the two diagnostic getters."""


class MetropolisHastingsSweep:
    """``MetropolisHastingsSweep`` (src/simulation/monte_carlo/metropolis_hastings_sweep.rs),
    checkerboard order on the device.  ``sweep`` takes host-given proposals and
    uniforms; ``next_element`` draws them on the device from a Philox stream
    keyed by ``seed`` (the reference's ``rng`` argument), one counter per sweep."""

    def __init__(self, number_of_update: int, spread: float, seed: int = 0):
        # MetropolisHastingsSweep::new returns None for these (:77-88)
        if number_of_update == 0 or not (0.0 < spread < 1.0):
            raise ValueError("number_of_update must be > 0 and spread in (0, 1)")
        self._n_update = int(number_of_update)
        self._spread = float(spread)
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._counter = 0
        self._prob_replace_mean = 0.0
        self._number_replace_last = 0

    def counter(self) -> int:
        """Sweeps drawn so far from the device stream."""
        return self._counter

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        """One sweep with proposals and accept draws generated on the device
        (metropolis_hastings_sweep.rs:146-174)."""
        st = F.MhStats()
        _raise(F.lib().lqcd_metropolis_sweep_random(state.handle, state.beta(), self._n_update,
                                                    self._spread, self._seed, self._counter,
                                                    C.byref(st)))
        self._counter += 1
        self._prob_replace_mean = st.prob_replace_mean
        self._number_replace_last = int(st.number_replace_last)
        return state

    def number_of_update(self) -> int:
        return self._n_update

    def spread(self) -> float:
        return self._spread

    def prob_replace_mean(self) -> float:
        return self._prob_replace_mean

    def number_replace_last(self) -> int:
        return self._number_replace_last

    def sweep(self, state: LatticeStateDefault, proposals: np.ndarray,
              uniforms: np.ndarray) -> LatticeStateDefault:
        """One sweep with proposals ``(n_update, nlink, 18)`` (orthonormalized,
        column-major) and one u64 uniform per link."""
        p = np.ascontiguousarray(proposals, dtype=np.float64)
        u = np.ascontiguousarray(uniforms, dtype=np.uint64)
        nlink = state.lattice().number_of_canonical_links_space()
        if p.shape != (self._n_update, nlink, 18) or u.shape != (nlink,):
            raise StateInitializationError("proposals / uniforms do not match the lattice")
        st = F.MhStats()
        _raise(F.lib().lqcd_metropolis_sweep(state.handle, state.beta(), self._n_update,
                                             F.ptr(p), F.ptr(u), C.byref(st)))
        self._prob_replace_mean = st.prob_replace_mean
        self._number_replace_last = int(st.number_replace_last)
        return state


class ChainSpec:
    """One chain of a batch: extents, beta, lattice spacing and seed."""

    def __init__(self, extents: Sequence[int], beta: float, seed: int, size: float = 1.0):
        self.extents = [int(x) for x in extents]
        self.beta = float(beta)
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.size = float(size)


class BatchedChains:
    """Independent HMC chains (each a ``LatticeStateDefault`` with its own
    ``HybridMonteCarloDiagnostic``, hybrid_monte_carlo.rs:465-471, 584-613)
    resident together and advanced by one launch per MD step; a trajectory is
    replayed from a captured HIP graph.  All chains share D and precision.

    Trajectory n of chain c draws momenta as ``lqcd_efield_random(seed_c, n,
    0.5 / beta_c)`` and its accept word from the chain's Philox stream, so a
    chain reproduces a single lattice driven with the same draws."""

    def __init__(self, chains: Sequence[ChainSpec], d: int = 4, precision: int = F32,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        n = len(chains)
        specs = (F.ChainSpec * max(n, 1))()
        for i, ch in enumerate(chains):
            if len(ch.extents) != d:
                raise LatticeInitializationError("every chain needs D extents")
            for k, x in enumerate(ch.extents):
                specs[i].extent[k] = x
            specs[i].beta = ch.beta
            specs[i].size = ch.size
            specs[i].seed = ch.seed
        h = C.c_void_p()
        _raise(F.lib().lqcd_batch_create(self.ctx.handle, d, precision, specs, n, C.byref(h)))
        self.handle = h
        self.d = d
        self.precision = precision
        self.chains = list(chains)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                F.lib().lqcd_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.chains)

    def number_of_links(self, chain: int) -> int:
        nl = C.c_size_t(0)
        _raise(F.lib().lqcd_batch_chain_info(self.handle, chain, None, C.byref(nl)))
        return int(nl.value)

    def total_sites(self) -> int:
        ns = C.c_size_t(0)
        _raise(F.lib().lqcd_batch_info(self.handle, None, C.byref(ns)))
        return int(ns.value)

    def cold_start(self) -> None:
        _raise(F.lib().lqcd_batch_cold_start(self.handle))

    def hot_start(self) -> None:
        _raise(F.lib().lqcd_batch_hot_start(self.handle))

    def link_matrix(self, chain: int) -> np.ndarray:
        n = self.number_of_links(chain)
        out = np.empty((n, 18))
        _raise(F.lib().lqcd_batch_links_download(self.handle, chain, F.ptr(out), n))
        return out

    def set_link_matrix(self, chain: int, links) -> None:
        a = np.ascontiguousarray(links, dtype=np.float64).reshape(-1, 18)
        _raise(F.lib().lqcd_batch_links_upload(self.handle, chain, F.ptr(a), a.shape[0]))

    def plaquette(self) -> np.ndarray:
        out = np.empty(len(self.chains))
        _raise(F.lib().lqcd_batch_plaquette(self.handle, F.ptr(out)))
        return out

    @property
    def counter(self) -> int:
        c = C.c_uint64(0)
        _raise(F.lib().lqcd_batch_counter_get(self.handle, C.byref(c)))
        return int(c.value)

    @counter.setter
    def counter(self, value: int) -> None:
        _raise(F.lib().lqcd_batch_counter_set(self.handle, int(value)))

    def hmc(self, delta_t: float, number_of_steps: int, n_trajectories: int = 1,
            drift: int = DRIFT_EXP, scheme: int = SCHEME_LEAPFROG,
            reunitarize: bool = False) -> np.ndarray:
        """Runs the trajectories; returns stats ``(n_trajectories, nchain)``
        with fields h_old, h_new, delta_h, probability, plaquette, accepted."""
        stats = np.zeros((n_trajectories, len(self.chains)), dtype=F.CHAIN_STATS_DTYPE)
        _raise(F.lib().lqcd_batch_hmc(self.handle, delta_t, number_of_steps, drift, scheme,
                                      n_trajectories, int(bool(reunitarize)),
                                      stats.ctypes.data_as(C.c_void_p)))
        return stats

    def autotune(self, drift: int = DRIFT_EXP) -> int:
        c = C.c_int(0)
        _raise(F.lib().lqcd_batch_autotune(self.handle, drift, C.byref(c)))
        return int(c.value)

    @property
    def config(self) -> int:
        c = C.c_int(0)
        _raise(F.lib().lqcd_batch_config_get(self.handle, C.byref(c)))
        return int(c.value)

    @config.setter
    def config(self, value: int) -> None:
        _raise(F.lib().lqcd_batch_config_set(self.handle, int(value)))

    def set_graph(self, enable: bool) -> None:
        _raise(F.lib().lqcd_batch_graph_enable(self.handle, int(bool(enable))))

    def set_md_reconstruct(self, mode: int) -> None:
        _raise(F.lib().lqcd_batch_reconstruct_set(self.handle, mode))

    def md_reconstruct_last(self) -> bool:
        mode, used = C.c_int(), C.c_int()
        _raise(F.lib().lqcd_batch_reconstruct_get(self.handle, C.byref(mode), C.byref(used)))
        return bool(used.value)


class MetropolisHastingsDeltaDiagnostic:
    """``MetropolisHastingsDeltaDiagnostic``
    (src/simulation/monte_carlo/metropolis_hastings.rs:296-427): one link
    drawn uniformly per step, proposal orthonormalize(random_su3_close_to_unity
    (spread)) U, accepted with p = min(1, exp(-dS)).  The draws come from the
    device stream keyed by ``seed`` (the reference's ``rng``), one counter per
    step."""

    def __init__(self, spread: float, seed: int = 0):
        # ::new returns None outside (0, 1) (:341-351)
        if not (0.0 < spread < 1.0):
            raise ValueError("spread must be in (0, 1)")
        self._spread = float(spread)
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._counter = 0
        self._prob_replace_last = 0.0
        self._has_replace_last = False
        self.last_link: Optional[int] = None

    def spread(self) -> float:
        return self._spread

    def prob_replace_last(self) -> float:
        return self._prob_replace_last

    def has_replace_last(self) -> bool:
        return self._has_replace_last

    def counter(self) -> int:
        return self._counter

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        p, acc, link = C.c_double(), C.c_int(), C.c_size_t()
        _raise(F.lib().lqcd_metropolis_link_random(state.handle, state.beta(), self._spread, self._seed,
                                                   self._counter, C.byref(p), C.byref(acc), C.byref(link)))
        self._counter += 1
        self._prob_replace_last = float(p.value)
        self._has_replace_last = bool(acc.value)
        self.last_link = int(link.value)
        return state


class HeatBathSweep:
    """``HeatBathSweep`` (src/simulation/monte_carlo/heat_bath.rs:20-123) on the
    device, checkerboard order.  The reference's ``rng`` is a Philox stream
    keyed by ``seed``, one counter per sweep.  As in the reference, the SU(2)
    factors use sigma_3 = identity, so links leave SU(3): follow with
    ``normalize_link_matrices``."""

    def __init__(self, seed: int = 0):
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._counter = 0
        self._exhausted_last = 0

    def counter(self) -> int:
        return self._counter

    def exhausted_last(self) -> int:
        """Norm samplers of the last sweep that ran out of attempts."""
        return self._exhausted_last

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        ex = C.c_size_t(0)
        _raise(F.lib().lqcd_heat_bath_sweep(state.handle, state.beta(), self._seed, self._counter,
                                            C.byref(ex)))
        self._counter += 1
        self._exhausted_last = int(ex.value)
        return state


class OverrelaxationSweepRotation:
    """``OverrelaxationSweepRotation`` (src/simulation/monte_carlo/overrelaxation.rs:43-106)
    on the device, checkerboard order; the Wilson action is unchanged."""

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        _raise(F.lib().lqcd_overrelaxation_sweep(state.handle))
        return state


class OverrelaxationSweepReverse:
    """``OverrelaxationSweepReverse`` (src/simulation/monte_carlo/overrelaxation.rs:108-183)
    on the device, checkerboard order; the Wilson action is unchanged."""

    def next_element(self, state: LatticeStateDefault) -> LatticeStateDefault:
        _raise(F.lib().lqcd_overrelaxation_reverse_sweep(state.handle))
        return state


class HybridMethodCouple:
    """``HybridMethodCouple`` (src/simulation/monte_carlo/hybrid.rs:327-420):
    ``method_1`` then ``method_2``, each through ``monte_carlo_step``."""

    def __init__(self, method_1, method_2):
        self._method_1 = method_1
        self._method_2 = method_2

    def method_1(self):
        return self._method_1

    def method_2(self):
        return self._method_2

    def deconstruct(self):
        return self._method_1, self._method_2

    def next_element(self, state):
        state = state.monte_carlo_step(self._method_1)
        return state.monte_carlo_step(self._method_2)


class HybridMethodVec:
    """``HybridMethodVec`` (src/simulation/monte_carlo/hybrid.rs:156-268): the
    methods applied in order, each through ``monte_carlo_step``."""

    def __init__(self, methods=None):
        self._methods = list(methods or [])

    def push_method(self, method) -> None:
        self._methods.append(method)

    def methods(self):
        return list(self._methods)

    def next_element(self, state):
        if not self._methods:
            raise ValueError("HybridMethodVecError::NoMethod")
        for m in self._methods:
            state = state.monte_carlo_step(m)
        return state


from . import statistics  # noqa: E402  (statistics::*, needs Context above)
