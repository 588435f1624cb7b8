"""Two-row link reads of exp-drift integrations (lqcd_md_reconstruct_*): used
only when the links are in SU(3) to 64 ulps of the storage type, agreeing
with the nine-entry path to rounding and with the f64 oracle to the tolerance
of each path; never used for the additive drift or links off SU(3)."""
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def run(u, e, precision, drift, mode, n=10, dt=0.01, ext=(4, 4, 4, 6)):
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, list(ext), 4), 6.0, u, precision)
    st.set_md_reconstruct(mode)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(drift, lqcd.SCHEME_LEAPFROG), dt, n)
    return st.md_reconstruct_last(), ef.link_matrix(), ef.e_field(), st


def test_su3_distance_known_values():
    announce("su3 distance")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(3)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 4, 4, 6], 4), 6.0, u, lqcd.F64)
    assert st.su3_distance() < 1e-14
    # one link scaled by 1.001: |U U^+ - 1| = 0.002001 on the diagonal, |det - 1| = 0.003003001
    u2 = u.copy()
    u2[17] *= 1.001
    st.set_link_matrix(u2)
    assert abs(st.su3_distance() - (1.001 ** 3 - 1)) < 1e-12
    # a U(3) link with det = i (off SU(3) by |i - 1|)
    u3 = u.copy()
    m = O.to_np(u3[5]) * np.exp(1j * np.pi / 6)
    u3[5] = O.from_np(m)
    st.set_link_matrix(u3)
    assert abs(st.su3_distance() - abs(1j - 1)) < 1e-12


def test_two_row_reads_match_nine_entries_and_oracle():
    announce("f32 exp drift: two-row vs nine-entry reads vs f64 oracle")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(9)
    e = ol.random_efield(8, 0.5 / 6.0)
    used, u_r12, e_r12, st = run(u, e, lqcd.F32, lqcd.DRIFT_EXP, 1)
    assert used
    used0, u_full, e_full, _ = run(u, e, lqcd.F32, lqcd.DRIFT_EXP, 0)
    assert not used0
    assert np.max(np.abs(u_r12 - u_full)) <= 2e-6
    assert np.max(np.abs(e_r12 - e_full)) <= 2e-6
    u32 = u.astype(np.float32).astype(np.float64)
    u_ref, e_ref = ol.simulate_symplectic_n(u32, e.astype(np.float32).astype(np.float64), 0.01, 10,
                                            O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    assert np.max(np.abs(u_r12 - u_ref)) <= 1e-5
    assert np.max(np.abs(e_r12 - e_ref)) <= 1e-5
    # the exp drift keeps the links in SU(3)
    assert st.su3_distance() <= 64 * 2.0 ** -24


def test_two_row_reads_f64_match_oracle():
    announce("f64 exp drift: two-row vs nine-entry reads vs oracle")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(14)
    e = ol.random_efield(15, 0.5 / 6.0)
    used, a, ea, _ = run(u, e, lqcd.F64, lqcd.DRIFT_EXP, 1)
    assert used
    used0, b, eb, _ = run(u, e, lqcd.F64, lqcd.DRIFT_EXP, 0)
    assert not used0
    assert np.max(np.abs(a - b)) <= 1e-13 and np.max(np.abs(ea - eb)) <= 1e-13
    u_ref, e_ref = ol.simulate_symplectic_n(u, e, 0.01, 10, O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    assert np.max(np.abs(a - u_ref)) <= 1e-12 and np.max(np.abs(ea - e_ref)) <= 1e-12


@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
def test_not_used_for_additive_drift(precision):
    announce(f"two-row reads off for the additive drift precision={precision}")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(10)
    e = ol.random_efield(11, 0.5 / 6.0)
    used, a, ea, _ = run(u, e, precision, lqcd.DRIFT_ADDITIVE, 1)
    assert not used
    _, b, eb, _ = run(u, e, precision, lqcd.DRIFT_ADDITIVE, 0)
    assert np.array_equal(a, b) and np.array_equal(ea, eb)


def test_not_used_off_su3():
    # links off SU(3) (one link scaled by 1.001; a heat-bath sweep, whose
    # subgroup factors follow the reference's sigma_3 = identity): the
    # integration reads all nine entries and matches that mode bit for bit
    announce("two-row reads off for links off SU(3)")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(12)
    u[7] *= 1.001
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 4, 4, 6], 4), 6.0, ol.hot_start(12), lqcd.F32)
    lqcd.HeatBathSweep(77).next_element(st)
    hb = st.link_matrix()
    e = ol.random_efield(13, 0.5 / 6.0)
    for links in (u, hb):
        st.set_link_matrix(links)
        assert st.su3_distance() > 1e-3
        used, a, ea, _ = run(links, e, lqcd.F32, lqcd.DRIFT_EXP, 1)
        assert not used
        _, b, eb, _ = run(links, e, lqcd.F32, lqcd.DRIFT_EXP, 0)
        assert np.array_equal(a, b) and np.array_equal(ea, eb)


def test_metropolis_two_row_staples():
    # device-proposal sweeps keep SU(3) links in SU(3) and read the staples
    # as two rows; nine-entry reads give the same accept decisions and links
    # to rounding.  After a heat-bath sweep (off SU(3)) they read nine entries
    # and agree with that mode to rounding.
    announce("metropolis two-row staples")
    ol = O.OracleLattice(4, [4, 4, 4, 6])
    u = ol.hot_start(16)
    runs = {}
    for mode in (1, 0):
        st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 4, 4, 6], 4), 5.5, None, lqcd.F64)
        st.set_link_matrix(u)
        st.normalize_link_matrices()
        st.set_md_reconstruct(mode)
        mh = lqcd.MetropolisHastingsSweep(1, 0.2, 99)
        counts = []
        for _ in range(3):
            mh.next_element(st)
            counts.append(mh.number_replace_last())
        runs[mode] = (st.md_reconstruct_last(), counts, st.link_matrix(), st)
    assert runs[1][0] and not runs[0][0]
    assert runs[1][1] == runs[0][1]
    assert np.max(np.abs(runs[1][2] - runs[0][2])) <= 1e-13
    out = {}
    for mode in (1, 0):
        st = runs[mode][3]
        lqcd.HeatBathSweep(3).next_element(st)
        mh = lqcd.MetropolisHastingsSweep(1, 0.2, 5)
        mh.next_element(st)
        out[mode] = (st.md_reconstruct_last(), st.link_matrix())
    assert not out[1][0]
    assert np.max(np.abs(out[1][1] - out[0][1])) <= 1e-12


@pytest.mark.parametrize("dim,ext", [(4, (32, 8, 4, 8)), (4, (16, 16, 4, 4)), (4, (8, 32, 4, 4)), (3, (16, 16, 4)),
                                     (2, (16, 16)), (4, (4, 4, 4, 6))])
def test_tile_variant_matches_site_variant(dim, ext):
    # MD variant 4 (LDS tile of two-row f32 links over whole x_0 rows and the
    # x_1 rows either side) against variant 3 (the same per-site arithmetic
    # with every link read from HBM) and the f64 oracle on f32-rounded
    # inputs; (4, 4, 4, 6) has no whole-row tile (16 sites per x_1 column):
    # variant 4 falls back to variant 3 there
    announce(f"MD tile variant ext={ext}")
    ol = O.OracleLattice(dim, list(ext))
    u = ol.hot_start(21).astype(np.float32).astype(np.float64)
    e = ol.random_efield(22, 0.5 / 6.0).astype(np.float32).astype(np.float64)
    got = {}
    for variant in (3, 4):
        st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, list(ext), dim), 6.0, u, lqcd.F32)
        st.set_md_variant(variant)
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG), 0.01, 6)
        assert st.md_reconstruct_last() and st.md_variant() == variant
        got[variant] = (ef.link_matrix(), ef.e_field())
    assert np.max(np.abs(got[4][0] - got[3][0])) <= 1e-6
    assert np.max(np.abs(got[4][1] - got[3][1])) <= 1e-6
    u_ref, e_ref = ol.simulate_symplectic_n(u, e, 0.01, 6, O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    assert np.max(np.abs(got[4][0] - u_ref)) <= 1e-5
    assert np.max(np.abs(got[4][1] - e_ref)) <= 1e-5
