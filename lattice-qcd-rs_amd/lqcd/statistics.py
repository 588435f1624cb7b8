"""``lattice_qcd_rs::statistics`` (src/statistics/mod.rs) on the device.

The reference's functions take slices (``mean``, ``variance``,
``mean_and_variance``, ``mean_with_error``, ``covariance``) or rayon parallel
iterators (the ``*_par_iter`` / ``*_par_iter_val`` forms, which compute the
variance and covariance in one pass over (x, x x) or (x, y, x y)).  Here both
families take a 1-D float64 or complex128 array, or ``DeviceValues`` already
resident in device memory; the ``_val`` forms are the same functions, as
Python has no borrowed iterators.  The sums run on the device in a fixed
order; the final arithmetic is the reference's.  Complex arithmetic has no
conjugation, as in the reference: the complex variance is the mean of
(x - m)(x - m).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Union

import numpy as np

from . import _ffi as F
from . import Context, _raise

__all__ = [
    "DeviceValues", "mean", "variance", "mean_and_variance", "mean_with_error", "covariance",
    "mean_par_iter", "mean_par_iter_val", "variance_par_iter", "variance_par_iter_val",
    "mean_and_variance_par_iter", "mean_and_variance_par_iter_val", "mean_with_error_par_iter",
    "mean_with_error_par_iter_val", "covariance_par_iter", "covariance_par_iter_val",
]

Number = Union[float, complex]


class DeviceValues:
    """``n`` doubles (or, when ``is_complex``, ``n`` (re, im) pairs) at device
    address ``ptr``; ``owner`` keeps the memory alive."""

    def __init__(self, ptr: int, n: int, is_complex: bool = False, owner=None):
        self.ptr = int(ptr)
        self.n = int(n)
        self.is_complex = bool(is_complex)
        self._owner = owner

    @classmethod
    def from_numpy(cls, a) -> "DeviceValues":
        a = np.asarray(a)
        if a.ndim != 1:
            raise ValueError("statistics take a 1-D sequence")
        cplx = np.iscomplexobj(a)
        a = np.ascontiguousarray(a, dtype=np.complex128 if cplx else np.float64)
        if a.size == 0:
            return cls(0, 0, cplx)
        buf = F.DeviceBuffer(a.nbytes)
        buf.fill_from(a)
        return cls(buf.ptr, a.size, cplx, buf)

    @classmethod
    def from_buffer(cls, buf: "F.DeviceBuffer", n: Optional[int] = None, is_complex: bool = False,
                    offset: int = 0) -> "DeviceValues":
        width = 16 if is_complex else 8
        if n is None:
            n = (buf.nbytes - offset) // width
        if offset < 0 or offset % 8 or offset + n * width > buf.nbytes:
            raise ValueError("range outside the buffer")
        return cls(buf.ptr + offset, n, is_complex, buf)


def _values(data) -> DeviceValues:
    return data if isinstance(data, DeviceValues) else DeviceValues.from_numpy(data)


def _ctx(ctx: Optional[Context]) -> Context:
    return ctx if ctx is not None else Context.default()


def _number(out: np.ndarray, cplx: bool) -> Number:
    return complex(out[0], out[1]) if cplx else float(out[0])


def mean(data, ctx: Optional[Context] = None) -> Number:
    """``statistics::mean`` (mod.rs:321-335): sum / len (NaN when empty)."""
    v = _values(data)
    out = np.zeros(2)
    _raise(F.lib().lqcd_stats_mean(_ctx(ctx).handle, C.c_void_p(v.ptr or None), v.n, int(v.is_complex),
                                   out.ctypes.data_as(C.c_void_p)))
    return _number(out, v.is_complex)


def _mean_and_variance(data, form: int, ctx) -> List[Number]:
    v = _values(data)
    out = np.zeros(4)
    _raise(F.lib().lqcd_stats_mean_and_variance(_ctx(ctx).handle, C.c_void_p(v.ptr or None), v.n,
                                                int(v.is_complex), form, out.ctypes.data_as(C.c_void_p)))
    k = 2 if v.is_complex else 1
    return [_number(out[:k], v.is_complex), _number(out[k:], v.is_complex)]


def mean_and_variance(data, ctx: Optional[Context] = None) -> List[Number]:
    """``statistics::mean_and_variance`` (mod.rs:375-399): [mean,
    sum (x - mean)^2 / (len - 1)]."""
    return _mean_and_variance(data, F.STATS_SEQUENTIAL, ctx)


def variance(data, ctx: Optional[Context] = None) -> Number:
    """``statistics::variance`` (mod.rs:346-369)."""
    return mean_and_variance(data, ctx)[1]


def _mean_with_error(data, form: int, ctx) -> List[float]:
    v = _values(data)
    if v.is_complex:
        raise TypeError("mean_with_error takes real values")
    out = np.zeros(2)
    _raise(F.lib().lqcd_stats_mean_with_error(_ctx(ctx).handle, C.c_void_p(v.ptr or None), v.n, form,
                                              out.ctypes.data_as(C.c_void_p)))
    return [float(out[0]), float(out[1])]


def mean_with_error(data, ctx: Optional[Context] = None) -> List[float]:
    """``statistics::mean_with_error`` (mod.rs:401-405): [mean,
    sqrt(variance / len)]."""
    return _mean_with_error(data, F.STATS_SEQUENTIAL, ctx)


def _covariance(data_1, data_2, form: int, ctx) -> Optional[Number]:
    a, b = _values(data_1), _values(data_2)
    if a.is_complex != b.is_complex:
        raise TypeError("covariance of a real and a complex sequence")
    out = np.zeros(2)
    has = C.c_int(0)
    _raise(F.lib().lqcd_stats_covariance(_ctx(ctx).handle, C.c_void_p(a.ptr or None), a.n,
                                         C.c_void_p(b.ptr or None), b.n, int(a.is_complex), form,
                                         out.ctypes.data_as(C.c_void_p), C.byref(has)))
    return _number(out, a.is_complex) if has.value else None


def covariance(data_1, data_2, ctx: Optional[Context] = None) -> Optional[Number]:
    """``statistics::covariance`` (mod.rs:427-458): mean(x y) - mean(x) mean(y),
    None when the lengths differ."""
    return _covariance(data_1, data_2, F.STATS_SEQUENTIAL, ctx)


def mean_par_iter_val(data, ctx: Optional[Context] = None) -> Number:
    """``mean_par_iter_val`` (mod.rs:57-64)."""
    return mean(data, ctx)


def mean_and_variance_par_iter_val(data, ctx: Optional[Context] = None) -> List[Number]:
    """``mean_and_variance_par_iter_val`` (mod.rs:163-181): one pass,
    (sum x^2 - (sum x)^2 / len) / (len - 1)."""
    return _mean_and_variance(data, F.STATS_PARALLEL, ctx)


def variance_par_iter_val(data, ctx: Optional[Context] = None) -> Number:
    """``variance_par_iter_val`` (mod.rs:113-127)."""
    return mean_and_variance_par_iter_val(data, ctx)[1]


def mean_with_error_par_iter_val(data, ctx: Optional[Context] = None) -> List[float]:
    """``mean_with_error_par_iter_val`` (mod.rs:201-207)."""
    return _mean_with_error(data, F.STATS_PARALLEL, ctx)


def covariance_par_iter_val(data_1, data_2, ctx: Optional[Context] = None) -> Optional[Number]:
    """``covariance_par_iter_val`` (mod.rs:279-306): one pass,
    (sum x y - sum x sum y / len) / len; None when the lengths differ."""
    return _covariance(data_1, data_2, F.STATS_PARALLEL, ctx)


# the by-reference forms (mod.rs:26-38, 79-94, 138-153, 189-194, 230-250)
mean_par_iter = mean_par_iter_val
variance_par_iter = variance_par_iter_val
mean_and_variance_par_iter = mean_and_variance_par_iter_val
mean_with_error_par_iter = mean_with_error_par_iter_val
covariance_par_iter = covariance_par_iter_val
