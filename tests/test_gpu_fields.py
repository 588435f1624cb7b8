"""GPU parity of the clover, F_mu_nu and chromomagnetic field
(LinkMatrix::clover / f_mu_nu / magnetic_field_vec, src/field.rs:806-882)
against the oracle, and the reference's known answers (src/field.rs:1579-1711).

Tolerances: the device computes in double on the stored links, so both
storage types agree with the oracle on the same (stored) links to 1e-12
absolute; the known answers to the reference's own EPSILON = 1e-9."""
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu

EPSILON = 1e-9


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def mat(m18):
    return O.to_np(np.asarray(m18))


def test_known_answers():
    # test magnetic_field, src/field.rs:1579-1711 (D = 3, L = 4, a = 1)
    announce("clover / F / B known answers D=3 L=4")
    ol = O.OracleLattice(3, 4)
    lat = lqcd.LatticeCyclic(1.0, 4, 3)
    x, y, z = 0, 1, 2
    p0 = [0, 0, 0]
    zero = np.zeros((3, 3))
    u = ol.cold_start()
    st = lqcd.LatticeStateDefault(lat, 1.0, u, lqcd.F64)
    assert np.linalg.norm(mat(st.clover(p0, x, y)) - 4 * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.f_mu_nu(p0, x, y)) - zero) < EPSILON
    assert np.linalg.norm(mat(st.magnetic_field(p0, x)) - zero) < EPSILON
    assert np.all(np.abs(st.magnetic_field_vec_field()) < EPSILON)

    u2 = u.copy()
    u2[0] = O.from_np(1j * np.eye(3))
    st.set_link_matrix(u2)
    assert np.linalg.norm(mat(st.clover(p0, x, y)) - 2 * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.clover(p0, y, x)) - 2 * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.f_mu_nu(p0, x, y))) < EPSILON
    b = mat(st.magnetic_field(p0, x))
    assert np.linalg.norm(b) < EPSILON
    for m in st.magnetic_field_vec(p0):
        assert np.linalg.norm(mat(m)) < EPSILON
    assert np.linalg.norm(mat(st.magnetic_field_link(p0, x)) - b) < EPSILON

    u3 = u.copy()
    u3[ol.site_index([1, 0, 0]) * 3 + y] = O.from_np(1j * np.eye(3))
    st.set_link_matrix(u3)
    assert np.linalg.norm(mat(st.clover(p0, x, y)) - (3 + 1j) * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.clover(p0, y, x)) - (3 - 1j) * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.f_mu_nu(p0, x, y)) - 0.25j * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.magnetic_field(p0, x))) < EPSILON
    b = mat(st.magnetic_field(p0, z))
    assert np.linalg.norm(mat(st.magnetic_field_link(p0, z)) - b) < EPSILON
    assert np.linalg.norm(b - 0.25 * np.eye(3)) < EPSILON
    assert np.linalg.norm(mat(st.magnetic_field([4, 0, 0], z)) - b) < EPSILON
    for d, m in enumerate(st.magnetic_field_vec(p0)):
        assert np.linalg.norm(mat(m) - (b if d == 2 else zero)) < EPSILON


@pytest.mark.parametrize("dim,ext,precision", [
    (4, [4, 3, 4, 2], lqcd.F64), (4, [4, 4, 4, 4], lqcd.F32), (3, [3, 4, 5], lqcd.F64),
    (2, [4, 6], lqcd.F64)])
def test_fields_match_oracle(dim, ext, precision):
    announce(f"clover / F / B vs oracle D={dim} ext={ext} precision={precision}")
    a = 0.7
    ol = O.OracleLattice(dim, ext, a)
    u = ol.hot_start(21)
    lat = lqcd.LatticeCyclic(a, ext, dim)
    st = lqcd.LatticeStateDefault(lat, 6.0, u, precision)
    u = st.link_matrix()          # the stored (for f32: rounded) links
    got_c, got_f, got_b = st.clover_field(), st.f_mu_nu_field(), st.magnetic_field_vec_field()
    ref_c, ref_f, ref_b = ol.clover_field(u), ol.f_mu_nu_field(u), ol.magnetic_field_vec(u)
    assert got_c.shape == ref_c.shape and got_b.shape == ref_b.shape
    assert np.max(np.abs(got_c - ref_c)) <= 1e-12
    assert np.max(np.abs(got_f - ref_f)) <= 1e-12
    assert np.max(np.abs(got_b - ref_b)) <= 1e-12
    # diagonal of F is exactly zero, as clover(i,i) - clover(i,i) in the reference
    for i in range(dim):
        assert np.all(got_f[:, i, i] == 0.0)
    if dim == 2:
        assert np.all(got_b == 0.0)


def test_size_mismatch_is_an_error():
    announce("clover field size check")
    st = lqcd.LatticeStateDefault.new_cold(1.0, 6.0, 4, 3, lqcd.F64)
    out = np.empty(10)
    with pytest.raises(lqcd.StateInitializationError):
        lqcd._raise(lqcd.F.lib().lqcd_f_mu_nu_field(st.handle, lqcd.F.ptr(out), out.size))


def test_device_resident_entry_matches_host_entry():
    # lqcd_field_strength_device writes the same layout into a caller-owned
    # device buffer (allocated from the library's own HIP runtime)
    announce("field strength into a device buffer")
    F = lqcd.F
    ol = O.OracleLattice(4, [4, 4, 2, 4])
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 4, 2, 4], 4), 6.0, ol.hot_start(3), lqcd.F32)
    for what, host in [(F.FIELD_CLOVER, st.clover_field), (F.FIELD_F_MU_NU, st.f_mu_nu_field),
                       (F.FIELD_MAGNETIC, st.magnetic_field_vec_field)]:
        want = host()
        buf = F.DeviceBuffer(want.nbytes)
        buf.fill_from(np.full(want.size, np.nan))
        lqcd._raise(F.lib().lqcd_field_strength_device(st.handle, what, buf.ptr, want.size))
        got = buf.to_numpy().reshape(want.shape)
        assert np.array_equal(got, want)
    with pytest.raises(lqcd.LqcdError):
        lqcd._raise(F.lib().lqcd_field_strength_device(st.handle, 7, buf.ptr, want.size))


@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
def test_sij_pij_fields(precision):
    # LinkMatrix::sij / pij (src/field.rs:743-771) at every site for every pair
    # of signed directions, against the oracle on the stored links
    announce(f"sij / pij fields D=3 precision={precision}")
    ol = O.OracleLattice(3, [4, 6, 4])
    u = ol.hot_start(23)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 6, 4], 3), 6.0, u, precision)
    us = st.link_matrix()   # the stored links (rounded for f32)
    for di in range(3):
        for dj in range(3):
            for pi in (True, False):
                for pj in (True, False):
                    a, b = lqcd.Direction(di, pi), lqcd.Direction(dj, pj)
                    sf, pf = st.sij_field(a, b), st.pij_field(a, b)
                    for s in (0, 17, 95):
                        assert np.max(np.abs(sf[s] - ol.sij(us, s, di, pi, dj, pj))) <= 1e-12
                        assert np.max(np.abs(pf[s] - ol.pij(us, s, di, pi, dj, pj))) <= 1e-12
    # one point, and Tr P summed over i < j is the plaquette sum
    p0 = st.pij((1, 2, 3), lqcd.Direction(0, True), lqcd.Direction(2, True))
    assert np.max(np.abs(p0 - ol.pij(us, st.lattice().index_of((1, 2, 3)), 0, True, 2, True))) <= 1e-12
    tr = sum(O.to_np(st.pij_field(lqcd.Direction(i, True), lqcd.Direction(j, True))).trace(axis1=-2, axis2=-1).sum()
             for i in range(3) for j in range(i + 1, 3))
    avg = st.average_trace_plaquette()
    assert abs(tr / (st.lattice().number_of_points() * 3) - avg) <= 1e-12
