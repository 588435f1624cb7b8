"""The golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle) against the oracle of this tree and against what the
mathematics fixes independently of it: SU(3) membership, H_U from a numpy
np.roll restatement, H_E in closed form, derivative_e from the numpy staples,
the Wilson action kept by overrelaxation (overrelaxation.rs test_same_energy),
F_ji = -F_ij and B from F through the Levi-Civita symbol (src/field.rs:822-868).
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from _lattice_np import T, hamiltonian_links_np, links_to_field, staple_field

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def as_matrices(a18):
    """(..., 18) column-major re/im -> (..., 3, 3) complex."""
    z = a18[..., 0::2] + 1j * a18[..., 1::2]
    return np.swapaxes(z.reshape(a18.shape[:-1] + (3, 3)), -1, -2)


def su3_deviation(u18):
    m = as_matrices(u18)
    unit = np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)))
    det = np.max(np.abs(np.linalg.det(m) - 1.0))
    return max(unit, det)


@pytest.fixture(scope="module")
def d4():
    return load("d4_l3.npz")


@pytest.fixture(scope="module")
def d3():
    return load("d3_l4.npz")


def lattice(g):
    return O.OracleLattice(int(g["cfg_dim"]), int(g["cfg_L"]), float(g["cfg_size"]))


def test_inputs_regenerate(d4, d3):
    # the stored inputs are the oracle's seeded streams (xoshiro256++ hot
    # start with the reference's test seed, Gaussian E, proposals, uniforms)
    l4, l3 = lattice(d4), lattice(d3)
    assert np.array_equal(l4.hot_start(O.SEED_RNG), d4["u"])
    assert np.array_equal(l4.random_efield(int(d4["cfg_e_seed"]), 0.5 / float(d4["cfg_beta"])), d4["e"])
    assert np.array_equal(l3.hot_start(O.SEED_RNG), d3["u"])
    assert np.array_equal(l3.mh_proposals(int(d3["cfg_mh_seed"]), float(d3["cfg_spread"]), 1), d3["mh_proposals"])
    rr = O.OracleRng(int(d3["cfg_uni_seed"]))
    assert np.array_equal(np.array([rr.next_u64() for _ in range(l3.nlink)], dtype=np.uint64), d3["mh_uniforms"])


def test_oracle_reproduces_d4(d4):
    l = lattice(d4)
    u, e = d4["u"], d4["e"]
    beta, dt, n = float(d4["cfg_beta"]), float(d4["cfg_dt"]), int(d4["cfg_n_md"])
    p = l.average_trace_plaquette(u)
    assert abs(p.real - d4["plaquette"][0]) <= 1e-13 and abs(p.imag - d4["plaquette"][1]) <= 1e-13
    assert abs(l.hamiltonian_links(u, beta) - d4["h_links"]) <= 1e-12 * abs(d4["h_links"])
    assert abs(l.hamiltonian_efield(e, beta) - d4["h_efield"]) <= 1e-12 * abs(d4["h_efield"])
    assert np.max(np.abs(l.derivative_e(u) - d4["derivative_e"])) <= 1e-13
    g = np.stack([l.gauss(u, e, s) for s in range(l.nsite)])
    assert np.max(np.abs(g - d4["gauss"])) <= 1e-13
    assert abs(l.gauss_sum_div(u, e) - d4["gauss_sum_div"]) <= 1e-12 * abs(d4["gauss_sum_div"])
    for key, mode, scheme in (("reference", O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE),
                              ("leapfrog", O.DRIFT_EXP, O.SCHEME_LEAPFROG)):
        uu, ee = l.simulate_symplectic_n(u, e, dt, n, mode, scheme)
        assert np.max(np.abs(uu - d4["u_" + key])) <= 1e-13
        assert np.max(np.abs(ee - d4["e_" + key])) <= 1e-13
    uh, st = l.hmc_trajectory(u, e, beta, dt, int(d4["cfg_n_hmc"]), int(d4["cfg_uniform"]), O.DRIFT_EXP,
                              O.SCHEME_LEAPFROG)
    h_old, h_new, dh, prob, acc = d4["hmc"]
    tol = 1e-12 * abs(h_old)
    assert abs(st.h_old - h_old) <= tol and abs(st.h_new - h_new) <= tol and abs(st.delta_h - dh) <= tol
    assert st.accepted == int(acc)
    assert np.max(np.abs(uh - d4["u_hmc"])) <= 1e-13


def test_oracle_reproduces_d3(d3):
    l = lattice(d3)
    u, beta = d3["u"], float(d3["cfg_beta"])
    um, st = l.metropolis_sweep(u, beta, d3["mh_proposals"], d3["mh_uniforms"], O.ORDER_CHECKERBOARD)
    assert st.number_replace_last == int(d3["metropolis"][1])
    assert abs(st.prob_replace_mean - d3["metropolis"][0]) <= 1e-13
    assert np.max(np.abs(um - d3["u_metropolis"])) <= 1e-13
    uh, ex = l.heat_bath_sweep(u, beta, int(d3["cfg_hb_seed"]), 0)
    assert ex == 0 and np.max(np.abs(uh - d3["u_heat_bath"])) <= 1e-13
    assert np.max(np.abs(l.overrelaxation_sweep(u, O.ORDER_CHECKERBOARD) - d3["u_overrelax"])) <= 1e-13
    assert np.max(np.abs(l.overrelaxation_reverse_sweep(u, O.ORDER_CHECKERBOARD) - d3["u_overrelax_reverse"])) \
        <= 1e-13
    assert np.max(np.abs(l.normalize_link_matrices(d3["u_noisy"]) - d3["u_normalized"])) <= 1e-15
    assert np.max(np.abs(l.f_mu_nu_field(u) - d3["f_mu_nu"])) <= 1e-13
    assert np.max(np.abs(l.magnetic_field_vec(u) - d3["magnetic_field"])) <= 1e-13


def test_d4_fixture_against_numpy(d4):
    dim, L, a = int(d4["cfg_dim"]), int(d4["cfg_L"]), float(d4["cfg_size"])
    ext, beta = [L] * dim, float(d4["cfg_beta"])
    u, e = d4["u"], d4["e"]
    nsite = L ** dim
    for key in ("u", "u_leapfrog", "u_hmc"):   # the exp drift keeps SU(3)
        assert su3_deviation(d4[key]) <= 1e-13, key
    # the additive drift leaves SU(3) (U + dt (i sqrt6/a) E U), by O(dt^2)
    assert 1e-8 < su3_deviation(d4["u_reference"]) < 1e-2
    h = hamiltonian_links_np(u, ext, beta)
    assert abs(d4["h_links"] - h) <= 1e-12 * abs(h)
    nplaq = nsite * dim * (dim - 1) // 2
    assert abs(d4["h_links"] - beta * nplaq * (1 - d4["plaquette"][0] / 3)) <= 1e-12 * abs(h)
    assert abs(d4["h_efield"] - beta * np.sum(e * e) / 2) <= 1e-12 * abs(d4["h_efield"])
    # derivative_e = -(sqrt(2/3) / a) Im Tr(T^a U_mu Sigma_mu), Sigma the staples
    f = links_to_field(u, ext)
    de = d4["derivative_e"].reshape(nsite, dim, 8)
    for mu in range(dim):
        m = f[..., mu, :, :].reshape(nsite, 3, 3) @ staple_field(f, mu).reshape(nsite, 3, 3)
        expect = np.stack([-math.sqrt(2 / 3) / a * np.trace(np.array(T[k]) @ m, axis1=-2, axis2=-1).imag
                           for k in range(8)], axis=-1)
        assert np.max(np.abs(de[:, mu, :] - expect)) <= 1e-12 * np.max(np.abs(expect))
    # an accepted trajectory: p = min(1, exp(-dH)) above the stored uniform
    h_old, h_new, dh, prob, acc = d4["hmc"]
    assert abs(dh - (h_new - h_old)) <= 1e-12 * abs(h_old)
    assert abs(prob - min(1.0, math.exp(-dh))) <= 1e-15
    assert acc == 1.0 and prob >= int(d4["cfg_uniform"]) / 2.0 ** 64


def test_d3_fixture_against_numpy(d3):
    dim, L, beta = int(d3["cfg_dim"]), int(d3["cfg_L"]), float(d3["cfg_beta"])
    ext = [L] * dim
    u = d3["u"]
    for key in ("u", "u_metropolis", "u_normalized"):
        assert su3_deviation(d3[key]) <= 1e-13, key
    assert np.max(np.abs(d3["u_normalized"] - d3["u_noisy"])) <= 10 * float(d3["cfg_noise"])
    # both overrelaxations keep the Wilson action.  The rotation's links are
    # unitary (rot U^+ rot); the reverse form's are not: su3::reverse
    # (src/su3.rs:705-714) negates the off-diagonal entries of the unitary
    # W^+ U V, which keeps Re Tr(U A) but not unitarity
    h = hamiltonian_links_np(u, ext, beta)
    m = as_matrices(d3["u_overrelax"])
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) <= 1e-13
    m = as_matrices(d3["u_overrelax_reverse"])
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) > 1e-3
    for key in ("u_overrelax", "u_overrelax_reverse"):
        h2 = hamiltonian_links_np(d3[key], ext, beta)
        assert abs(h2 - h) <= 100 * np.finfo(float).eps * L ** dim * h
    # the Metropolis sweep replaced exactly the links it counts
    moved = np.max(np.abs(d3["u_metropolis"] - u), axis=1) > 0
    assert int(moved.sum()) == int(d3["metropolis"][1])
    # F_ji = -F_ij, zero diagonal; B_d = sum_ij eps(d, i, j) F_ij / (2i)
    fm = as_matrices(d3["f_mu_nu"])
    assert np.array_equal(fm, -np.swapaxes(fm, 1, 2))
    b = as_matrices(d3["magnetic_field"])
    eps = np.zeros((3, 3, 3))
    for (i, j, k) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        eps[i, j, k], eps[i, k, j] = 1.0, -1.0
    expect = np.einsum("dij,sijab->sdab", eps, fm) / 2j
    assert np.max(np.abs(b - expect)) <= 1e-13 * max(1.0, np.max(np.abs(expect)))
