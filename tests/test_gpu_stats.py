"""GPU parity of the statistics entry points (lqcd_stats_*, the functions of
src/statistics/mod.rs) against the oracle and the reference's known answers.

Tolerances: the device sums in a tree order, the oracle left to right as the
reference's sequential functions do.  A left-to-right sum of n values carries
a rounding error that grows like sqrt(n) eps (measured 4.5e-14 relative for
the mean of 2e7 uniforms), so device and oracle agree to 4 sqrt(n) eps times
the mean magnitude of the summed terms; the device's tree sums are also held
to the exactly rounded sums (math.fsum) at 32 eps (means) and 1e-13
(variances, covariances) times that magnitude.  Inputs whose sums are exact
(the reference's known answers) agree bit for bit.  Two calls on the same data give the same bits (fixed
reduction order)."""
import math
import sys

import numpy as np
import pytest

import oracle as O
import lqcd
from lqcd import statistics as S

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


EPS = 2.0 ** -52


def close(a, b, rel, abs_=0.0):
    return abs(a - b) <= rel * abs(b) + abs_


def fsum_c(z):
    z = np.asarray(z)
    if np.iscomplexobj(z):
        return complex(math.fsum(z.real), math.fsum(z.imag))
    return math.fsum(z)


def exact(x, y):
    """mean, variance and population covariance from exactly rounded sums"""
    n = len(x)
    m = fsum_c(x) / n
    v = fsum_c((x - m) * (x - m)) / (n - 1) if n > 1 else float("nan")
    c = fsum_c(x * y) / n - (fsum_c(x) / n) * (fsum_c(y) / n)
    return m, v, c


def check_against(x, y, dx, dy):
    n = len(x)
    tol = 4.0 * math.sqrt(n) * EPS
    # error scales: rounding in a sum is bounded by the sum of the magnitudes
    # of its terms (the complex variance of isotropic data is near zero while
    # its terms are not)
    scale_m = float(np.mean(np.abs(x)))
    scale_v = float(np.mean(np.abs(x) ** 2))
    scale_c = float(np.mean(np.abs(x) * np.abs(y))) + abs(np.mean(x)) * abs(np.mean(y))
    em, ev, ec = exact(x, y)
    m = S.mean(dx)
    assert close(m, O.stats_mean(x), 0.0, tol * scale_m)
    assert close(m, em, 0.0, 32 * EPS * scale_m)
    for parallel, mv, cov in ((False, S.mean_and_variance, S.covariance),
                              (True, S.mean_and_variance_par_iter, S.covariance_par_iter)):
        m, v = mv(dx)
        om, ov = O.stats_mean_and_variance(x, parallel)
        assert close(m, om, 0.0, tol * scale_m)
        if n == 1:
            assert math.isnan(v.real) and math.isnan(ov.real)
        else:
            assert close(v, ov, 0.0, tol * scale_v)
            assert close(v, ev, 0.0, 1e-13 * scale_v)
        c = cov(dx, dy)
        assert close(c, O.stats_covariance(x, y, parallel), 0.0, tol * scale_c)
        assert close(c, ec, 0.0, 1e-13 * scale_c)


def test_known_answers():
    announce("statistics known answers")
    a = np.ones(100)
    assert S.mean_and_variance(a) == [1.0, 0.0]
    assert S.mean_and_variance_par_iter(a) == [1.0, 0.0]
    assert S.mean_par_iter(a) == 1.0 and S.variance(a) == 0.0 and S.variance_par_iter(a) == 0.0
    assert S.mean_with_error(a) == [1.0, 0.0] and S.mean_with_error_par_iter(a) == [1.0, 0.0]
    a = np.array([0.0, 1.0, 0.0, 1.0])
    for f in (S.mean_and_variance, S.mean_and_variance_par_iter):
        assert f(a) == [0.5, 1.0 / 3.0]
    for f in (S.mean_with_error, S.mean_with_error_par_iter):
        assert f(a) == [0.5, math.sqrt(1.0 / 3.0 / 4.0)]
    for f in (S.covariance, S.covariance_par_iter):
        assert f([1.0], [0.0, 1.0]) is None
        assert f([], [1.0]) is None
        assert f([1.0, 2.0, 3.0, 4.0], [1.0, 2.0, 3.0, 4.0]) == 1.25
        assert f(np.full(4, 2.0), np.full(4, 2.0)) == 0.0
    assert S.mean_and_variance(np.array([1 + 2j, -7 - 9j])) == [-3 - 3.5j, -28.5 + 88j]


@pytest.mark.parametrize("n", [1, 2, 3, 255, 4096, 1000003, 4096 * 256 * 4 + 17, 20_000_000])
def test_real_against_oracle(n):
    announce(f"statistics real n={n}")
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 0.25
    y = rng.uniform(-1, 1, n) - 0.5 * x
    dx, dy = S.DeviceValues.from_numpy(x), S.DeviceValues.from_numpy(y)
    check_against(x, y, dx, dy)
    # same data, same bits (n = 1 gives a NaN variance both times)
    assert np.array_equal(S.mean_and_variance(dx), S.mean_and_variance(dx), equal_nan=True)


@pytest.mark.parametrize("n", [1, 7, 65537, 3_000_001])
def test_complex_against_oracle(n):
    announce(f"statistics complex n={n}")
    rng = np.random.default_rng(100 + n)
    z = rng.normal(size=n) + 1j * rng.normal(size=n) + (0.3 - 0.2j)
    w = rng.normal(size=n) + 1j * rng.normal(size=n)
    check_against(z, w, S.DeviceValues.from_numpy(z), S.DeviceValues.from_numpy(w))


def test_edge_cases_and_buffers():
    announce("statistics edge cases, device buffers")
    assert math.isnan(S.mean([]))
    with pytest.raises(lqcd.LqcdError):
        S.mean_and_variance([])
    assert math.isnan(S.covariance([], []))
    with pytest.raises(TypeError):
        S.mean_with_error(np.array([1j, 2j]))
    # a sub-range of a device buffer, 8-byte aligned, read as complex pairs
    a = np.arange(21, dtype=np.float64) * 0.5 - 3.0
    buf = lqcd._ffi.DeviceBuffer(a.nbytes)
    buf.fill_from(a)
    v = S.DeviceValues.from_buffer(buf, 10, True, offset=8)
    z = a[1:21].view(np.complex128)
    om, ov = O.stats_mean_and_variance(z)
    m, var = S.mean_and_variance(v)
    assert close(m, om, 1e-15, 1e-15) and close(var, ov, 1e-13)
    with pytest.raises(ValueError):
        S.DeviceValues.from_buffer(buf, 11, True, offset=8)


@pytest.mark.parametrize("n", [2, 3, 1_000_001])
def test_aligned_and_unaligned_reads_agree(n):
    # buffers aligned to 16 bytes are read two reals per lane, 8-byte aligned
    # ones a double at a time: the same sums up to the order of the additions,
    # and covariance of one of each takes the double-at-a-time path
    announce(f"statistics aligned / unaligned reads n={n}")
    rng = np.random.default_rng(300 + n)
    a = rng.normal(0.7, 1.3, n + 1)
    b = rng.normal(-0.2, 0.9, n + 1)
    bufs = []
    for arr in (a, b):
        buf = lqcd._ffi.DeviceBuffer(arr.nbytes)
        buf.fill_from(arr)
        bufs.append(buf)
    xu = S.DeviceValues.from_buffer(bufs[0], n, False, offset=8)   # a[1:], 8-byte aligned
    yu = S.DeviceValues.from_buffer(bufs[1], n, False, offset=8)
    xa, ya = S.DeviceValues.from_numpy(a[1:].copy()), S.DeviceValues.from_numpy(b[1:].copy())
    om, ov = O.stats_mean_and_variance(a[1:])
    for v in (xu, xa):
        m, var = S.mean_and_variance(v)
        assert close(m, om, 1e-15, 1e-13) and close(var, ov, 1e-13)
    c_ref = O.stats_covariance(a[1:], b[1:], False)
    for x, y in ((xa, ya), (xa, yu), (xu, yu)):
        assert close(S.covariance(x, y), c_ref, 1e-15, 1e-12)


def test_statistics_of_a_device_field():
    # the statistics of F_mu_nu components left on the device by
    # lqcd_field_strength_device, against the same field on the host
    announce("statistics of a device F_mu_nu field, 8^4")
    st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 8, 3, 4, lqcd.F64)
    host = st.f_mu_nu_field()
    n = host.size
    buf = lqcd._ffi.DeviceBuffer(n * 8)
    F = lqcd._ffi
    lqcd._raise(F.lib().lqcd_field_strength_device(st.handle, F.FIELD_F_MU_NU, buf.ptr, n))
    assert np.array_equal(buf.to_numpy(), host.reshape(-1))
    v = S.DeviceValues.from_buffer(buf, n // 2, True)
    z = host.reshape(-1).view(np.complex128)
    m, var = S.mean_and_variance(v)
    om, ov = O.stats_mean_and_variance(z)
    assert close(m, om, 1e-12, 1e-14) and close(var, ov, 1e-11)
