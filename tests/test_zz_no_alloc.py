"""lqcd_stats_mean_and_variance_noalloc (fixed device scratch, no allocation)
against lqcd_stats_mean_and_variance on the same device values, through the
C ABI: both forms, real and complex values, and two calls in a row with the
second on fewer values, so the second call runs over what the first left in
the fixed scratch.  The two entry points share their kernels and launch
shapes, so their results agree bit for bit.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import lqcd
from lqcd import _ffi as F
from lqcd import statistics as S

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def call(name, ctx, vals, form):
    out = np.zeros(4 if vals.is_complex else 2)
    rc = getattr(F.lib(), name)(ctx.handle, C.c_void_p(vals.ptr), vals.n, int(vals.is_complex), form,
                                out.ctypes.data_as(C.c_void_p))
    return rc, out


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("form", [F.STATS_SEQUENTIAL, F.STATS_PARALLEL])
def test_noalloc_matches_allocating_entry_point(cplx, form):
    ctx = lqcd.Context.default()
    rng = np.random.default_rng(7 + 2 * int(cplx) + form)
    # a grid of the full 4096 blocks first, then a much smaller one
    for n in (5_000_000, 1_234):
        announce(f"stats noalloc cplx={cplx} form={form} n={n}")
        x = rng.normal(1.5, 2.0, n)
        if cplx:
            x = x + 1j * rng.normal(-0.5, 1.0, n)
        vals = S.DeviceValues.from_numpy(x)
        rc_a, got = call("lqcd_stats_mean_and_variance_noalloc", ctx, vals, form)
        rc_b, ref = call("lqcd_stats_mean_and_variance", ctx, vals, form)
        assert rc_a == F.OK and rc_b == F.OK
        assert np.array_equal(got, ref), (got, ref)
        # and against the numpy moments of the same values
        m, v = (got[0] + 1j * got[1], got[2] + 1j * got[3]) if cplx else (got[0], got[1])
        assert abs(m - np.mean(x)) <= 1e-12 * (1 + abs(np.mean(x)))
        d = x - np.mean(x)
        assert abs(v - np.sum(d * d) / (n - 1)) <= 1e-10 * np.mean(np.abs(x) ** 2)


def test_noalloc_errors_as_the_allocating_entry_point():
    announce("stats noalloc error paths")
    ctx = lqcd.Context.default()
    vals = S.DeviceValues.from_numpy(np.arange(10.0))
    out = np.zeros(2)
    lib = F.lib()
    p_out = out.ctypes.data_as(C.c_void_p)
    # empty input, unknown form, more values than accepted: errors, no launch
    assert lib.lqcd_stats_mean_and_variance_noalloc(ctx.handle, C.c_void_p(vals.ptr), 0, 0, 0, p_out) \
        == F.ERR_INVALID_ARGUMENT
    assert lib.lqcd_stats_mean_and_variance_noalloc(ctx.handle, C.c_void_p(vals.ptr), 10, 0, 7, p_out) \
        == F.ERR_INVALID_ARGUMENT
    assert lib.lqcd_stats_mean_and_variance_noalloc(ctx.handle, C.c_void_p(vals.ptr), (1 << 62) + 1, 0, 0,
                                                    p_out) == F.ERR_INVALID_ARGUMENT
    # n = 1: NaN variance, as the reference's (len - 1) = 0 division
    rc, got = call("lqcd_stats_mean_and_variance_noalloc", ctx, S.DeviceValues(vals.ptr, 1, False, vals), 0)
    assert rc == F.OK and got[0] == 0.0 and np.isnan(got[1])
