"""Pins the oracle's heat bath and overrelaxation
(src/simulation/monte_carlo/heat_bath.rs, overrelaxation.rs,
src/statistics/distribution.rs) by exact distribution moments, an
independent SVD and the reference's own conservation criterion."""
import math

import numpy as np
import pytest
from scipy import integrate, special

import oracle as O


def test_su2_project_unorm_is_quaternion_form():
    # su2.rs:155-157: m - m^+ + I conj(Tr m) = [[a + conj(d), b - conj(c)], [c - conj(b), d + conj(a)]]
    rng = np.random.default_rng(1)
    for _ in range(50):
        m = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        p = O.su2_project_unorm(m)
        assert abs(p[1, 1] - np.conj(p[0, 0])) < 1e-14
        assert abs(p[1, 0] + np.conj(p[0, 1])) < 1e-14
        # proportional to SU(2): p p^+ = det(p) I with det(p) >= 0
        d = np.linalg.det(p)
        assert abs(d.imag) < 1e-12 and d.real >= 0
        assert np.allclose(p @ p.conj().T, d.real * np.eye(2), atol=1e-12)
    assert np.allclose(O.su2_project_unorm(np.eye(2)), 2 * np.eye(2))
    assert np.allclose(O.su2_project_unorm(np.array([[0, 1], [0, 0]])), [[0, 1], [-1, 0]])


@pytest.mark.parametrize("p", [0.7, 2.0, 8.0])
def test_heat_bath_norm_sampler_moments(p):
    # HeatBathDistributionNorm (distribution.rs:320-350) draws x0 with density
    # proportional to sqrt(1 - x^2) exp(p x) on [-1, 1]:
    #   E[x0] = I_2(p) / I_1(p), E[x0^2] from quadrature
    n = 20000
    x = np.array([O.heat_bath_x0(p, i, 0, 5) for i in range(n)])
    assert np.all(np.isfinite(x)) and np.all(np.abs(x) <= 1.0)
    mean = special.iv(2, p) / special.iv(1, p)
    w = lambda t: math.sqrt(1 - t * t) * math.exp(p * t)
    z = integrate.quad(w, -1, 1)[0]
    m2 = integrate.quad(lambda t: t * t * w(t), -1, 1)[0] / z
    se = math.sqrt((m2 - mean ** 2) / n)
    assert abs(x.mean() - mean) < 5 * se
    se2 = math.sqrt(np.var(x ** 2) / n)
    assert abs((x ** 2).mean() - m2) < 5 * se2


def test_heat_bath_sweep_limits_and_determinism():
    l = O.OracleLattice(4, 4)
    u0 = l.cold_start()
    # a cold start at large beta stays close to cold: every staple block is a
    # multiple of the identity and x0 -> 1
    u, ex = l.heat_bath_sweep(u0, 200.0, 3, 0)
    assert ex == 0
    assert l.average_trace_plaquette(u).real / 3 > 0.97
    uh = l.hot_start(4)
    a, ex_a = l.heat_bath_sweep(uh, 6.0, 3, 7)
    b, _ = l.heat_bath_sweep(uh, 6.0, 3, 7)
    c, _ = l.heat_bath_sweep(uh, 6.0, 3, 8)
    assert ex_a == 0
    assert np.array_equal(a, b) and not np.allclose(a, c)
    # the action goes down from a hot start at beta = 6
    assert l.hamiltonian_links(a, 6.0) < l.hamiltonian_links(uh, 6.0)
    with pytest.raises(ValueError):
        O.OracleLattice(4, [4, 4, 3, 4]).heat_bath_sweep(O.OracleLattice(4, [4, 4, 3, 4]).cold_start(),
                                                        6.0, 1, 0)


def test_polar_unitary_equals_svd_rotation():
    # overrelaxation.rs:66-78 takes U V^+ of the SVD; it is the unitary polar factor
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        uu, _, vh = np.linalg.svd(a)
        w = O.to_np(O.polar_unitary(O.from_np(a)))
        assert np.max(np.abs(w - uu @ vh)) < 1e-12


@pytest.mark.parametrize("order", [O.ORDER_LEXICOGRAPHIC, O.ORDER_CHECKERBOARD])
def test_overrelaxation_keeps_the_action(order):
    # overrelaxation.rs test_same_energy: D = 3, L = 4, beta = 1, hot start,
    # |h - h2| < eps * 100 * 4^3 * (h + h2) / 2
    l = O.OracleLattice(3, 4)
    u = l.hot_start(O.SEED_RNG)
    h = l.hamiltonian_links(u, 1.0)
    u2 = l.overrelaxation_sweep(u, order)
    h2 = l.hamiltonian_links(u2, 1.0)
    assert abs(h - h2) < np.finfo(float).eps * 100 * 4 ** 3 * (h + h2) * 0.5
    m = O.to_np(u2)
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) < 1e-12
    assert not np.allclose(u2, u)


def test_svd3_reconstructs():
    rng = np.random.default_rng(3)
    for _ in range(100):
        a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        u, sv, v = O.svd3(O.from_np(a))
        uu, vv = O.to_np(u), O.to_np(v)
        assert np.max(np.abs(uu @ np.diag(sv) @ vv.conj().T - a)) < 1e-13
        assert np.allclose(np.sort(sv), np.sort(np.linalg.svd(a, compute_uv=False)), atol=1e-12)


def test_overrelaxation_reverse_matches_lapack_svd_sweep():
    # OverrelaxationSweepReverse::get_modif (overrelaxation.rs:136-150):
    # U reverse(U^+ L V) V^+ for staple^+ = U S V^+; a checkerboard sweep rebuilt
    # with numpy's LAPACK SVD (other phase / order conventions) must agree.  (In
    # lexicographic order every update depends on the previous one and rounding
    # differences grow along the sweep; the checkerboard has 6 dependent stages.)
    l = O.OracleLattice(3, 4)
    u0 = l.hot_start(O.SEED_RNG)
    ref = u0.copy()
    for d in range(3):
        for par in range(2):
            new = {}
            for s in range(l.nsite):
                if l.parity(s) != par:
                    continue
                li = s * 3 + d
                a = O.to_np(l.staple(ref, s, d)).conj().T
                uu, _, vh = np.linalg.svd(a)
                m = uu.conj().T @ O.to_np(ref[li]) @ vh.conj().T
                m = np.where(np.eye(3, dtype=bool), m, -m)
                new[li] = O.from_np(uu @ m @ vh)
            for li, val in new.items():
                ref[li] = val
    got = l.overrelaxation_reverse_sweep(u0, O.ORDER_CHECKERBOARD)
    assert np.max(np.abs(got - ref)) < 1e-11
    # overrelaxation.rs test_same_energy criterion
    h, h2 = l.hamiltonian_links(u0, 1.0), l.hamiltonian_links(got, 1.0)
    assert abs(h - h2) < np.finfo(float).eps * 100 * 4 ** 3 * (h + h2) * 0.5
