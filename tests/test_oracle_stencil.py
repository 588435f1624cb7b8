"""Pins the oracle's plaquette, Hamiltonians, staple and force (src/field.rs:726-804,
src/simulation/state.rs:821-849,1370-1448, src/simulation/monte_carlo/mod.rs:324-362)."""
import math

import numpy as np
import pytest

import oracle as O
from _lattice_np import (T, clover_field, dag, expi, hamiltonian_links_np, links_to_field,
                         field_to_links, permutation_sign, staple_field, su3_matrix)

EPSILON = 1e-9


def mat(m18):
    return O.to_np(m18)


def test_clover_known_answers():
    # src/field.rs:1579-1711 (D=3, L=4) -- pins signed-direction sij/pij
    l = O.OracleLattice(3, 4)
    u = l.cold_start()
    x, y = 0, 1
    assert np.linalg.norm(mat(l.clover(u, 0, x, y)) - 4 * np.eye(3)) < EPSILON
    u2 = u.copy()
    u2[0] = O.from_np(1j * np.eye(3))
    assert np.linalg.norm(mat(l.clover(u2, 0, x, y)) - 2 * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(l.clover(u2, 0, y, x)) - 2 * np.eye(3)) < EPSILON
    u3 = u.copy()
    u3[l.site_index([1, 0, 0]) * 3 + y] = O.from_np(1j * np.eye(3))
    c_xy = mat(l.clover(u3, 0, x, y))
    c_yx = mat(l.clover(u3, 0, y, x))
    assert np.linalg.norm(c_xy - (3 + 1j) * np.eye(3)) < EPSILON
    assert np.linalg.norm(c_yx - (3 - 1j) * np.eye(3)) < EPSILON
    f = (c_xy - c_yx) / 8.0  # f_mu_nu, src/field.rs:823-833
    assert np.linalg.norm(f - 0.25j * np.eye(3)) < EPSILON
    # direct pij consequences: P_xy(0) = i I when U_x(0) = i I
    assert np.allclose(mat(l.pij(u2, 0, x, 1, y, 1)), 1j * np.eye(3), atol=1e-15)
    assert np.allclose(mat(l.pij(u2, 0, y, 1, x, 1)), -1j * np.eye(3), atol=1e-15)


def test_magnetic_field_known_answers():
    # the rest of src/field.rs:1579-1711 (test magnetic_field, D=3, L=4)
    l = O.OracleLattice(3, 4)
    x, y, z = 0, 1, 2
    zero = np.zeros((3, 3))
    u = l.cold_start()
    assert np.linalg.norm(mat(l.f_mu_nu(u, 0, x, y)) - zero) < EPSILON
    assert np.linalg.norm(mat(l.magnetic_field(u, 0, x)) - zero) < EPSILON
    assert np.all(np.abs(l.magnetic_field_vec(u)) < EPSILON)
    u2 = u.copy()
    u2[0] = O.from_np(1j * np.eye(3))
    assert np.linalg.norm(mat(l.f_mu_nu(u2, 0, x, y))) < EPSILON
    for d in range(3):
        assert np.linalg.norm(mat(l.magnetic_field(u2, 0, d))) < EPSILON
    u3 = u.copy()
    u3[l.site_index([1, 0, 0]) * 3 + y] = O.from_np(1j * np.eye(3))
    assert np.linalg.norm(mat(l.f_mu_nu(u3, 0, x, y)) - 0.25j * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(l.magnetic_field(u3, 0, x))) < EPSILON
    b = mat(l.magnetic_field(u3, 0, z))
    assert np.linalg.norm(b - 0.25 * np.eye(3)) < EPSILON
    bv = l.magnetic_field_vec(u3)
    s0 = l.site_index([0, 0, 0])
    for d in range(3):
        want = b if d == 2 else zero
        assert np.linalg.norm(mat(bv[s0, d]) - want) < EPSILON
    # [4, 0, 0] is [0, 0, 0] on the periodic lattice: the same site index
    assert l.site_index([0, 0, 0]) == s0


def test_levi_civita_known_values():
    # src/utils.rs:360-372
    for idx, want in [([], 1), ([1, 2], 1), ([0, 1], 1), ([1, 2, 3], 1), ([0, 1, 2], 1),
                      ([3, 1, 2], 1), ([2, 3, 1], 1), ([3, 1, 2, 4], 1), ([1, 3, 4, 2], 1),
                      ([3, 3, 1], 0), ([1, 1, 1], 0), ([1, 1], 0)]:
        assert O.levi_civita(idx) == want, idx
    for idx in [[1, 0], [0, 2, 1], [2, 1, 0], [0, 1, 3, 2], [3, 0, 1]]:
        assert O.levi_civita(idx) == permutation_sign(idx), idx


@pytest.mark.parametrize("dim,ext", [(4, [4, 3, 4, 2]), (3, [3, 4, 5]), (2, [4, 6])])
def test_clover_f_b_against_numpy_shifts(dim, ext):
    # clover (field.rs:806-820) as four closed loops from x written with array
    # shifts; F = (C_ij - C_ji) / (8 a^2) (field.rs:822-834);
    # B_d = sum_{i,j} eps(d, i, j) F_ij / (2i) (field.rs:848-868)
    a = 0.7
    l = O.OracleLattice(dim, ext, a)
    u = l.hot_start(21)
    f = links_to_field(u, ext)
    cl = l.clover_field(u)
    fm = l.f_mu_nu_field(u)
    bv = l.magnetic_field_vec(u)
    C = {}
    for i in range(dim):
        for j in range(dim):
            C[i, j] = clover_field(f, i, j).reshape(-1, 3, 3)
            assert np.max(np.abs(O.to_np(cl[:, i, j]) - C[i, j])) < 1e-12
    F = {}
    for i in range(dim):
        for j in range(dim):
            F[i, j] = (C[i, j] - C[j, i]) / (8 * a * a)
            assert np.max(np.abs(O.to_np(fm[:, i, j]) - F[i, j])) < 1e-12
    for d in range(dim):
        b = sum(permutation_sign([d, i, j]) * F[i, j] for i in range(dim) for j in range(dim)) / 2j
        assert np.max(np.abs(O.to_np(bv[:, d]) - b)) < 1e-12
    # algebra: C_ji = C_ij^+, so F is antihermitian and antisymmetric and B hermitian
    for i in range(dim):
        for j in range(dim):
            assert np.max(np.abs(C[j, i] - dag(C[i, j]))) < 1e-12
            assert np.max(np.abs(O.to_np(fm[:, i, j]) + dag(O.to_np(fm[:, i, j])))) < 1e-12
    for d in range(dim):
        b = O.to_np(bv[:, d])
        assert np.max(np.abs(b - dag(b))) < 1e-12
    if dim == 2:
        assert np.all(bv == 0.0)


def test_clover_is_gauge_covariant():
    # every leaf starts and ends at x: C_ij(x) -> g(x) C_ij(x) g(x)^+ under
    # U_mu(x) -> g(x) U_mu(x) g(x+mu)^+; a wrong leaf orientation breaks this
    ext = [4, 4, 2, 4]
    l = O.OracleLattice(4, ext)
    u = l.hot_start(5)
    g = np.array([O.to_np(m) for m in l.hot_start(6).reshape(-1, 4, 18)[:, 0]])
    f = links_to_field(u, ext).reshape(-1, 4, 3, 3)
    nbr = np.array([[l.neighbor(s, mu, True) for mu in range(4)] for s in range(l.nsite)])
    ft = np.empty_like(f)
    for mu in range(4):
        ft[:, mu] = g @ f[:, mu] @ dag(g[nbr[:, mu]])
    ut = O.from_np(ft.reshape(-1, 3, 3)).reshape(-1, 18)
    c0 = O.to_np(l.clover_field(u))
    c1 = O.to_np(l.clover_field(ut))
    for i in range(4):
        for j in range(4):
            assert np.max(np.abs(c1[:, i, j] - g @ c0[:, i, j] @ dag(g))) < 1e-12


@pytest.mark.parametrize("dim,L", [(3, 4), (4, 2), (4, 3)])
def test_cold_state(dim, L):
    # cold: <Tr P> = 3, H_U = 0, force = 0 exactly
    l = O.OracleLattice(dim, L)
    u = l.cold_start()
    assert l.average_trace_plaquette(u) == 3.0
    assert l.hamiltonian_links(u, 6.0) == 0.0
    assert np.all(l.derivative_e(u) == 0.0)


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 4]), (3, [3, 4, 5]), (4, [2, 3, 2, 4])])
def test_u1_embedded_closed_form(dim, ext):
    # diag(e^{i a}, e^{i b}, e^{-i(a+b)}) links: Tr P = sum_k exp(i Phi_k),
    # Phi = th(x,mu) + th(x+mu,nu) - th(x+nu,mu) - th(x,nu)
    rng = np.random.default_rng(10)
    l = O.OracleLattice(dim, ext)
    th = rng.uniform(-math.pi, math.pi, (l.nsite, dim, 2))
    ph = np.concatenate([th, -(th[..., :1] + th[..., 1:])], axis=-1)
    diag = np.zeros((l.nsite, dim, 3, 3), dtype=complex)
    for k in range(3):
        diag[..., k, k] = np.exp(1j * ph[..., k])
    u = O.from_np(diag).reshape(-1, 18)
    total = 0.0
    for s in range(l.nsite):
        for i in range(dim):
            for j in range(i + 1, dim):
                sp_i = l.neighbor(s, i, True)
                sp_j = l.neighbor(s, j, True)
                phi = ph[s, i] + ph[sp_i, j] - ph[sp_j, i] - ph[s, j]
                expect = np.sum(np.exp(1j * phi))
                got = np.trace(mat(l.pij(u, s, i, 1, j, 1)))
                assert abs(got - expect) < 1e-13
                total += expect
    nplaq = l.nsite * dim * (dim - 1) // 2
    assert abs(l.average_trace_plaquette(u) - total / nplaq) < 1e-13


@pytest.mark.parametrize("dim,ext,beta", [(4, [3, 3, 3, 3], 6.0), (3, [4, 3, 5], 2.0),
                                          (4, [2, 2, 2, 2], 5.7)])
def test_hamiltonian_links_vs_roll(dim, ext, beta):
    l = O.OracleLattice(dim, ext)
    u = l.hot_start(O.SEED_RNG)
    h = l.hamiltonian_links(u, beta)
    assert abs(h - hamiltonian_links_np(u, ext, beta)) <= 1e-12 * abs(h)
    # H_U = beta * nplaq * (1 - Re<Tr P>/3)
    p = l.average_trace_plaquette(u)
    nplaq = l.nsite * dim * (dim - 1) // 2
    assert abs(h - beta * nplaq * (1 - p.real / 3)) <= 1e-12 * abs(h)


def test_hamiltonian_efield_closed_form():
    l = O.OracleLattice(4, 3)
    e = l.random_efield(3, 0.7)
    assert abs(l.hamiltonian_efield(e, 6.0) - 6.0 * np.sum(e * e) / 2) < 1e-10


@pytest.mark.parametrize("dim,ext", [(4, [3, 3, 3, 3]), (3, [4, 4, 4])])
def test_staple_vs_roll_and_force_identity(dim, ext):
    # staple (monte_carlo/mod.rs:339-362) and derivative_e's signed-direction
    # sum (state.rs:1428-1437) are the same matrix
    l = O.OracleLattice(dim, ext)
    u = l.hot_start(5)
    f = links_to_field(u, ext)
    de = l.derivative_e(u).reshape(l.nsite, dim, 8)
    for mu in range(dim):
        st = staple_field(f, mu).reshape(l.nsite, 3, 3)
        for s in range(l.nsite):
            a = mat(l.staple(u, s, mu))
            assert np.allclose(a, st[s], atol=1e-13, rtol=0)
        umu = f[..., mu, :, :].reshape(l.nsite, 3, 3)
        m = umu @ st
        expect = np.stack([-math.sqrt(2 / 3) * np.trace(np.array(T[a]) @ m, axis1=-2, axis2=-1).imag
                           for a in range(8)], axis=-1)
        assert np.allclose(de[:, mu, :], expect, atol=1e-13, rtol=0)


def _perturb(u, ext, omega, eps):
    f = links_to_field(u, ext)
    return field_to_links(expi(eps * omega) @ f)


@pytest.mark.parametrize("dim,ext,beta", [(4, [3, 3, 3, 3], 5.7), (3, [4, 3, 4], 2.0)])
def test_force_central_finite_differences(dim, ext, beta):
    # d/de H_U(exp(i e Omega) U) = sum Omega^a dH/dw^a and
    # derivative_e^a = -(sqrt(6)/(beta a)) dH/dw^a  (state.rs:1420-1448)
    l = O.OracleLattice(dim, ext)
    u = l.hot_start(9)
    rng = np.random.default_rng(12)
    de = l.derivative_e(u)
    grad = l.wilson_backward(u, beta)
    def fd(omega, eps):
        hp = l.hamiltonian_links(_perturb(u, ext, omega, eps), beta)
        hm = l.hamiltonian_links(_perturb(u, ext, omega, -eps), beta)
        return (hp - hm) / (2 * eps)

    for _ in range(8):
        omega = rng.normal(size=tuple(reversed(ext)) + (dim, 8))
        om = omega.reshape(-1, 8)
        pred = -(beta / math.sqrt(6)) * np.sum(om * de)
        scale = max(1.0, abs(pred))
        # eps = 1e-4 as in the adjoint check: O(eps^2) truncation
        assert abs(fd(omega, 1e-4) - pred) <= 1e-5 * scale
        assert abs(fd(omega, 1e-4) - np.sum(om * grad)) <= 1e-5 * scale
        # Richardson extrapolation removes the eps^2 term
        rich = (4 * fd(omega, 1e-3) - fd(omega, 2e-3)) / 3
        assert abs(rich - pred) <= 1e-8 * scale


def test_force_single_component_finite_difference():
    l = O.OracleLattice(4, 3)
    ext = [3, 3, 3, 3]
    u = l.hot_start(21)
    beta = 6.0
    grad = l.wilson_backward(u, beta)
    rng = np.random.default_rng(3)
    for _ in range(12):
        link = int(rng.integers(l.nlink))
        a = int(rng.integers(8))
        omega = np.zeros((l.nlink, 8))
        omega[link, a] = 1.0
        omega = omega.reshape(tuple(reversed(ext)) + (4, 8))
        def fd(eps):
            return (l.hamiltonian_links(_perturb(u, ext, omega, eps), beta)
                    - l.hamiltonian_links(_perturb(u, ext, omega, -eps), beta)) / (2 * eps)
        rich = (4 * fd(1e-3) - fd(2e-3)) / 3
        assert abs(rich - grad[link, a]) <= 1e-8 * max(1.0, abs(grad[link, a]))


def test_derivative_u_closed_form():
    # state.rs:1407-1417: E * U * i sqrt(6) / a
    l = O.OracleLattice(4, 2, 0.5)
    u = l.hot_start(4)
    e = l.random_efield(8, 1.0)
    for k in [0, 5, 17, 63]:
        s, d = divmod(k, 4)
        got = mat(l.derivative_u(u, e, s, d))
        expect = su3_matrix(e[k]) @ mat(u[k]) * 1j * math.sqrt(6) / 0.5
        assert np.allclose(got, expect, atol=1e-14, rtol=0)


def test_delta_s_equals_delta_h():
    # metropolis_hastings.rs:479-514 (D=4, L=4, beta=2, size 1000): exp(-dS) == exp(-dH)
    l = O.OracleLattice(4, 4, 1000.0)
    u = l.hot_start(O.SEED_RNG)
    r = O.OracleRng(99)
    beta = 2.0
    for _ in range(10):
        k = int(r.next_u64() % l.nlink)
        s, d = divmod(k, 4)
        prop = O.orthonormalize(r.random_su3_close_to_unity(0.01))
        new = O.from_np(mat(prop) @ mat(u[k]))
        ds = l.delta_s(u, s, d, new, beta)
        u2 = u.copy()
        u2[k] = new
        dh = l.hamiltonian_links(u2, beta) - l.hamiltonian_links(u, beta)
        assert abs(min(1, math.exp(-ds)) - min(1, math.exp(-dh))) < 1e-8
        assert abs(ds - dh) < 1e-9
        u = u2


def test_checkerboard_class_independence():
    # links of one (mu, parity) class never enter each other's staples
    l = O.OracleLattice(4, 4)
    u = l.hot_start(2)
    rng = np.random.default_rng(0)
    for mu in range(4):
        s0 = int(rng.integers(l.nsite))
        u2 = u.copy()
        u2[s0 * 4 + mu] = O.from_np(np.eye(3) * 1j)
        for s in range(l.nsite):
            if l.parity(s) == l.parity(s0) and s != s0:
                assert np.array_equal(l.staple(u, s, mu), l.staple(u2, s, mu))
