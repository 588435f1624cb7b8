"""The oracle's statistics (src/statistics/mod.rs) against the reference's
own known answers (the mean_var test, mod.rs:468-529, and the doctests of
covariance_par_iter / covariance_par_iter_val / covariance, mod.rs:212-229,
258-277, 409-425), numpy's mean / var(ddof=1) / population covariance, and the
one-pass against the two-pass forms (the reference's 1e-8 agreement,
mod.rs:500-526)."""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("parallel", [False, True])
def test_mean_var_known_answers(parallel):
    # mod.rs:468-490
    a = [1.0] * 100
    assert O.stats_mean_and_variance(a, parallel) == [1.0, 0.0]
    assert O.stats_mean(a) == 1.0
    a = [0.0, 1.0, 0.0, 1.0]
    assert O.stats_mean_and_variance(a, parallel) == [0.5, 1.0 / 3.0]
    m, v = O.stats_mean_and_variance(a, parallel)
    assert [m, math.sqrt(v / 4)] == [0.5, math.sqrt(1.0 / 3.0 / 4.0)]


@pytest.mark.parametrize("parallel", [False, True])
def test_covariance_known_answers(parallel):
    # mod.rs:496-498: lengths 1 and 2 give None
    assert O.stats_covariance([1.0], [0.0, 1.0], parallel) is None
    assert O.stats_covariance([], [1.0], parallel) is None
    x = [1.0, 2.0, 3.0, 4.0]
    # covariance_par_iter doctest: Some(1.25)
    assert O.stats_covariance(x, x, parallel) == 1.25
    # covariance_par_iter_val doctest: (1+1) against (1+1) is Some(0)
    assert O.stats_covariance([2.0] * 4, [2.0] * 4, parallel) == 0.0
    # covariance doctest: complex inputs give a value
    z = [1 + 2j, -7 - 9j]
    assert O.stats_covariance(z, z, parallel) is not None


def test_against_numpy_closed_forms():
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 1001)
    y = rng.uniform(-1, 1, 1001)
    for parallel in (False, True):
        m, v = O.stats_mean_and_variance(x, parallel)
        assert m == pytest.approx(np.mean(x), rel=1e-13)
        assert v == pytest.approx(np.var(x, ddof=1), rel=1e-12)
        c = O.stats_covariance(x, y, parallel)
        # population covariance (divides by len), not numpy's default ddof=1
        assert c == pytest.approx(np.cov(x, y, ddof=0)[0, 1], rel=1e-10, abs=1e-15)
        assert c != pytest.approx(np.cov(x, y)[0, 1], rel=1e-6)


def test_complex_has_no_conjugation():
    rng = np.random.default_rng(6)
    z = rng.normal(size=257) + 1j * rng.normal(size=257)
    w = rng.normal(size=257) + 1j * rng.normal(size=257)
    m = np.mean(z)
    want_var = np.sum((z - m) * (z - m)) / (z.size - 1)
    for parallel in (False, True):
        got_m, got_v = O.stats_mean_and_variance(z, parallel)
        assert abs(got_m - m) <= 1e-14
        assert abs(got_v - want_var) <= 1e-12 * abs(want_var)
        # numpy's var takes |z - m|^2, a real number: a conjugating variance fails here
        assert abs(got_v - np.var(z, ddof=1)) > 1e-3
        c = O.stats_covariance(z, w, parallel)
        want_c = np.mean(z * w) - np.mean(z) * np.mean(w)
        assert abs(c - want_c) <= 1e-12
    # the worked example: 1+2i and -7-9i
    assert O.stats_mean_and_variance([1 + 2j, -7 - 9j]) == [-3 - 3.5j, -28.5 + 88j]


def test_one_pass_matches_two_pass_like_the_reference():
    # mod.rs:500-526: 100 draws of 100 uniforms in [-1, 1), 1e-8 agreement
    rng = np.random.default_rng(0x4578_93F4_4AB0_67F0)
    for _ in range(100):
        x = rng.uniform(-1, 1, 100)
        y = rng.uniform(-1, 1, 100)
        a = O.stats_mean_and_variance(x)
        b = O.stats_mean_and_variance(x, True)
        assert abs(a[0] - b[0]) < 1e-8 and abs(a[1] - b[1]) < 1e-8
        assert abs(O.stats_covariance(x, y) - O.stats_covariance(x, y, True)) < 1e-8


def test_edge_lengths():
    assert math.isnan(O.stats_mean([]))
    with pytest.raises(ValueError):
        O.stats_mean_and_variance([])
    # one value: 0 / 0
    for parallel in (False, True):
        m, v = O.stats_mean_and_variance([3.0], parallel)
        assert m == 3.0 and math.isnan(v)
    assert math.isnan(O.stats_covariance([], []))
