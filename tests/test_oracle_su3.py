"""Pins the oracle's su(3)/SU(3) algebra (src/su3.rs, src/field.rs:29-205) to the
reference's closed forms and to scipy's matrix exponential."""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle as O

EPSILON = 1e-9  # src/test/mod.rs:16


def frob(a):
    return float(np.linalg.norm(a))


def assert_eq_matrix(a, b, eps=EPSILON):
    # assert_eq_matrix!, src/macro_def.rs:151-162 (Frobenius norm of the difference)
    assert frob(a - b) < eps, frob(a - b)


def assert_is_su3(m, eps=EPSILON):
    # assert_matrix_is_su_3!, src/macro_def.rs:251-267
    assert abs(np.linalg.det(m) - 1.0) < eps
    assert frob(m @ m.conj().T - np.eye(3)) < eps


GELL_MANN = [
    np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=complex),
    np.array([[0, -1j, 0], [1j, 0, 0], [0, 0, 0]], dtype=complex),
    np.array([[1, 0, 0], [0, -1, 0], [0, 0, 0]], dtype=complex),
    np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], dtype=complex),
    np.array([[0, 0, -1j], [0, 0, 0], [1j, 0, 0]], dtype=complex),
    np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]], dtype=complex),
    np.array([[0, 0, 0], [0, 0, -1j], [0, 1j, 0]], dtype=complex),
    np.diag([1, 1, -2]).astype(complex) / math.sqrt(3),
]


def test_generators_are_gell_mann_over_two():
    # src/su3.rs:24-194 and the module doc: lambda_a / 2
    for a in range(8):
        g = O.to_np(O.generator(a))
        assert np.allclose(g, GELL_MANN[a] / 2, atol=1e-16, rtol=0)


def test_generator_algebra():
    # src/test/mod.rs:231-250
    gens = [O.to_np(O.generator(a)) for a in range(8)]
    for a in range(7):
        assert np.linalg.det(gens[a]) == 0
    for g in gens:
        assert np.trace(g) == 0
        assert np.linalg.norm(g.conj().T - g) == 0
    for i in range(8):
        for j in range(8):
            tr = abs(np.trace(gens[i] @ gens[j]))
            assert math.isclose(tr, 0.5 if i == j else 0.0, rel_tol=1e-15, abs_tol=1e-15)


def test_to_matrix_and_invariants():
    # src/field.rs:1480-1510: trace_squared = |Tr m^2|, d = i det m, t = -Tr(m^2)/2
    rng = np.random.default_rng(1)
    for _ in range(100):
        v = rng.uniform(-1, 1, 8)
        m = O.to_np(O.adjoint_to_matrix(v))
        assert np.allclose(m, sum(v[a] * GELL_MANN[a] / 2 for a in range(8)), atol=1e-15)
        assert abs(O.trace_squared(v) - abs(np.trace(m @ m))) < EPSILON
        assert abs(O.adjoint_d(v) - 1j * np.linalg.det(m)) < EPSILON
        assert abs(-0.5 * O.trace_squared(v) + np.trace(m @ m) / 2) < EPSILON


def test_factorial_constant():
    # src/su3.rs:1017-1021: N == factorial_size_for_exp() + 1 == 26
    assert O.factorial_size_for_exp() + 1 == 26


@pytest.mark.parametrize("f", [1.0, 2.0, -1.254, 4.254])
def test_exp_su3_closed_forms(f):
    # src/test/mod.rs:96-146
    v1 = np.zeros(8)
    v1[0] = 2 * f
    expect1 = np.array([[math.cos(f), 1j * math.sin(f), 0], [1j * math.sin(f), math.cos(f), 0],
                        [0, 0, 1]])
    assert_eq_matrix(O.to_np(O.su3_exp_i(v1)), expect1)
    v2 = np.zeros(8)
    v2[1] = 2 * f
    expect2 = np.array([[math.cos(f), math.sin(f), 0], [-math.sin(f), math.cos(f), 0], [0, 0, 1]])
    assert_eq_matrix(O.to_np(O.su3_exp_i(v2)), expect2)


def test_exp_i_and_exp_r_vs_scipy():
    # src/test/mod.rs:148-190 (unit vectors and 100 draws in [-pi, pi)^8)
    rng = np.random.default_rng(2)
    vs = [np.eye(8)[a] for a in range(8)] + [rng.uniform(-math.pi, math.pi, 8) for _ in range(100)]
    for v in vs:
        m = O.to_np(O.adjoint_to_matrix(v))
        assert_eq_matrix(O.to_np(O.su3_exp_i(v)), scipy.linalg.expm(1j * m))
        assert_eq_matrix(O.to_np(O.su3_exp_r(v)), scipy.linalg.expm(m))


def test_exp_i_is_su3():
    # src/test/mod.rs:252-265
    rng = np.random.default_rng(3)
    for _ in range(100):
        assert_is_su3(O.to_np(O.su3_exp_i(rng.uniform(-math.pi, math.pi, 8))))


def _m(rows):
    return np.array(rows, dtype=complex)


def test_orthonormalize_exact_cases():
    # src/test/mod.rs:455-567 (exact equality, as the reference asserts)
    cases = [
        (np.zeros((3, 3)), np.zeros((3, 3))),
        (np.eye(3), np.eye(3)),
        (_m([[1, 0, 0], [0, 0, 0], [0, 0, 0]]), _m([[1, 0, 0], [0, 0, 0], [0, 0, 0]])),
        (_m([[2, 0, 0], [0, 2, 0], [0, 0, 0]]), np.eye(3)),
        (_m([[2j, 0, 0], [0, 2j, 0], [0, 0, 0]]), np.diag([1j, 1j, -1])),
        (_m([[0, 1, 0], [1, 0, 0], [0, 0, 0]]), _m([[0, 1, 0], [1, 0, 0], [0, 0, -1]])),
        (_m([[0, 1, 0], [0, 0, 0], [1, 0, 0]]), _m([[0, 1, 0], [0, 0, 1], [1, 0, 0]])),
    ]
    for m, expect in cases:
        got = O.to_np(O.orthonormalize(O.from_np(m)))
        assert np.array_equal(got, expect.astype(complex)), (m, got)


def test_orthonormalize_random_is_su3():
    # src/test/mod.rs:569-592 (random 3x3 with entries in [-10, 10))
    rng = np.random.default_rng(4)
    for _ in range(100):
        m = rng.uniform(-10, 10, (3, 3)) + 1j * rng.uniform(-10, 10, (3, 3))
        assert_is_su3(O.to_np(O.orthonormalize(O.from_np(m))))


def test_orthonormalize_is_column_gram_schmidt():
    # su3.rs:279-295: columns c0/|c0|, orthogonalised c1, conj(c0) x conj(c1)
    rng = np.random.default_rng(5)
    for _ in range(50):
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        got = O.to_np(O.orthonormalize(O.from_np(m)))
        c0 = m[:, 0] / np.linalg.norm(m[:, 0])
        c1 = m[:, 1] - c0 * np.vdot(c0, m[:, 1])
        c1 = c1 / np.linalg.norm(c1)
        c2 = np.cross(c0.conj(), c1.conj())
        assert np.allclose(got, np.stack([c0, c1, c2], axis=1), atol=1e-14, rtol=0)


def test_random_su3_is_su3():
    # src/test/mod.rs:594-600
    r = O.OracleRng(O.SEED_RNG)
    for _ in range(100):
        assert_is_su3(O.to_np(r.random_su3()))


def test_random_su3_close_to_unity():
    # su3.rs:385-398: after orthonormalize, SU(3); the raw product is r*s*t of
    # su2 blocks (get_r/s/t, su3.rs:428-530) whose x0 = +-sqrt(1 - spread^2)
    # carries a random sign (su2.rs:91-97), so it lies near diag(+-1, +-1, +-1)
    import itertools
    signs = [np.diag(s) for s in itertools.product([1, -1], repeat=3)]
    r = O.OracleRng(7)
    for spread in (0.1, 0.5):
        for _ in range(50):
            raw = O.to_np(r.random_su3_close_to_unity(spread))
            ortho = O.to_np(O.orthonormalize(O.from_np(raw)))
            assert_is_su3(ortho)
            dist = min(np.linalg.norm(raw - s) for s in signs)
            assert dist < 6 * spread + 1e-12


def test_su2_block_structure_pauli3_quirk():
    # su2.rs:80-100, :134-140 with PAULI_3 == identity (su2.rs:39-45):
    # m = [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 + i x3]], |x| = spread
    r = O.OracleRng(11)
    for _ in range(50):
        m = r.random_su2_close_to_unity(0.3)
        assert m[0, 0] == m[1, 1]
        x0, x3 = m[0, 0].real, m[0, 0].imag
        x2, x1 = m[0, 1].real, m[0, 1].imag
        assert m[1, 0] == complex(-x2, x1)
        assert math.isclose(math.sqrt(x1 * x1 + x2 * x2 + x3 * x3), 0.3, rel_tol=1e-14)
        assert math.isclose(abs(x0), math.sqrt(1 - 0.09), rel_tol=1e-14)


def test_bernoulli_rule():
    # rand 0.8.5 Bernoulli: p == 1 -> always; else u < floor(p * 2^64)
    assert O.bernoulli_accept(1.0, 2**64 - 1)
    assert not O.bernoulli_accept(0.0, 0)
    assert O.bernoulli_accept(0.5, 2**63 - 1)
    assert not O.bernoulli_accept(0.5, 2**63)
    assert O.bernoulli_accept(0.25, 2**62 - 1)
    assert not O.bernoulli_accept(0.25, 2**62)


def test_clamp_probability_nan_is_one():
    # hybrid_monte_carlo.rs:584-589 under Rust min/max: NaN -> 1
    assert O.clamp_probability(float("nan")) == 1.0
    assert O.clamp_probability(float("inf")) == 1.0
    assert O.clamp_probability(0.3) == 0.3
    assert O.clamp_probability(-0.0) == 0.0


def test_philox_known_answer():
    # Random123 known-answer vectors for philox4x32-10
    assert list(O.philox4x32([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                          0x9B00DBD8]
    assert list(O.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E,
                                                                     0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                             [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                            0x24126EA1]
