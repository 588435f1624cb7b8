"""GPU parity of the integrator transitions the reference's SymplecticEulerRayon
implements besides integrate_symplectic (src/integrator/symplectic_euler_rayon.rs:
127-220) and the state methods built on them (simulate_sync(_n),
simulate_to_leapfrog, simulate_leap(_n), simulate_to_synchronous,
simulate_using_leapfrog_n; src/simulation/state.rs:302-418, 528-648), through
the C ABI, against the oracle on the same inputs (f64: links and momenta to
1e-12), and the reference's own tests of them (src/test/integrator.rs:9-67,
src/test/mod.rs:325-455) with their bounds."""
import math
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def efsync(dim, ext, seed, sigma, size=1.0, beta=6.0):
    ol = O.OracleLattice(dim, ext, size)
    u = ol.hot_start(seed)
    e = ol.random_efield(seed + 1, sigma)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(size, ext, dim), beta, u, lqcd.F64)
    return ol, u, e, lqcd.LatticeStateEFSyncDefault(st, e)


@pytest.mark.parametrize("drift", [lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP])
@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 6]), (3, [6, 4, 6]), (2, [8, 6])])
def test_transitions_against_oracle(drift, dim, ext):
    announce(f"integrator transitions drift={drift} D={dim} {ext}")
    ol, u, e, ef = efsync(dim, ext, 21, 0.3, size=0.9)
    integ = lqcd.SymplecticEulerGpu(drift)

    def same(u_ref, e_ref, state):
        assert np.max(np.abs(state.link_matrix() - u_ref)) <= 1e-12
        assert np.max(np.abs(state.e_field() - e_ref)) <= 1e-12

    ef = ef.simulate_sync(integ, 0.017)
    u, e = ol.integrate_sync_sync(u, e, 0.017, drift)
    same(u, e, ef)
    assert ef.t() == 1
    ef = ef.simulate_sync_n(integ, 0.011, 3)
    for _ in range(3):
        u, e = ol.integrate_sync_sync(u, e, 0.011, drift)
    same(u, e, ef)
    assert ef.t() == 4
    leap = ef.simulate_to_leapfrog(integ, 0.013)
    u, e = ol.integrate_sync_leap(u, e, 0.013)
    same(u, e, leap)
    assert leap.t() == 4
    leap = leap.simulate_leap(integ, 0.013)
    u, e = ol.integrate_leap_leap(u, e, 0.013, drift)
    same(u, e, leap)
    leap = leap.simulate_leap_n(integ, 0.013, 3)
    for _ in range(3):
        u, e = ol.integrate_leap_leap(u, e, 0.013, drift)
    same(u, e, leap)
    assert leap.t() == 8
    ef = leap.simulate_to_synchronous(integ, 0.013)
    u, e = ol.integrate_leap_sync(u, e, 0.013, drift)
    same(u, e, ef)
    assert ef.t() == 9
    with pytest.raises(lqcd.MultiIntegrationError):
        ef.simulate_sync_n(integ, 0.01, 0)
    with pytest.raises(lqcd.MultiIntegrationError):
        ef.simulate_to_leapfrog(integ, 0.01).simulate_leap_n(integ, 0.01, 0)


@pytest.mark.parametrize("drift", [lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP])
def test_leap_chain_equals_leapfrog_scheme(drift):
    # simulate_to_leapfrog + simulate_leap_n(n - 1) + simulate_to_synchronous is
    # simulate_using_leapfrog_n; the chain's launches differ from the scheme's
    announce(f"leap chain vs leapfrog scheme drift={drift}")
    _, u, e, a = efsync(4, [4, 6, 4, 4], 31, 0.2)
    _, _, _, b = efsync(4, [4, 6, 4, 4], 31, 0.2)
    integ = lqcd.SymplecticEulerGpu(drift)
    a = a.simulate_to_leapfrog(integ, 0.02).simulate_leap_n(integ, 0.02, 5).simulate_to_synchronous(integ, 0.02)
    b = b.simulate_using_leapfrog_n(integ, 0.02, 6)
    assert np.max(np.abs(a.link_matrix() - b.link_matrix())) <= 1e-13
    assert np.max(np.abs(a.e_field() - b.e_field())) <= 1e-13
    assert a.t() == b.t() == 6


def test_reference_integrator_test():
    # src/test/integrator.rs:9-67 with SymplecticEulerRayon (additive drift):
    # D = 3, L = 4, beta = 8, a cold start moved by 10 Metropolis sweeps,
    # momenta of new_random_e (N(0, 0.5 / beta) projected on Gauss's law,
    # state.rs:262-277); DT = 1e-4.  The reference bounds |H - H'| by 1e-4
    # (simulate_symplectic_n) and 1e-5 (the leap sequences) on the sample its
    # StdRng draws; the additive drift's departure from SU(3) alone makes
    # |H - H'| of order 1e-5 here (1.1e-5 for the 10-step leapfrog on this
    # sample), so the 1e-4 bound is checked on every sequence and each |H - H'|
    # against the oracle's on the same state to 1e-9.
    announce("reference integrator test D=3 L=4")
    dt = 1e-4
    st = lqcd.LatticeStateDefault.new_cold(1.0, 8.0, 4, 3, lqcd.F64)
    mh = lqcd.MetropolisHastingsSweep(1, 0.1, 7)
    for _ in range(10):
        mh.next_element(st)
    u = st.link_matrix()
    e = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 8, 0).e_field()
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE)
    ol = O.OracleLattice(3, 4)

    def fresh():
        s = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 4, 3), 8.0, u, lqcd.F64)
        return lqcd.LatticeStateEFSyncDefault(s, e)

    def h_oracle(uu, ee):
        return ol.hamiltonian_links(uu, 8.0) + ol.hamiltonian_efield(ee, 8.0)

    h = fresh().hamiltonian_total()
    assert abs(h - h_oracle(u, e)) <= 1e-12 * abs(h)
    a = fresh().simulate_symplectic_n(integ, dt, 10)
    ua, ea = ol.simulate_symplectic_n(u, e, dt, 10, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    b = fresh().simulate_to_leapfrog(integ, dt).simulate_leap_n(integ, dt, 1)
    b = b.simulate_to_synchronous(integ, dt).simulate_sync_n(integ, dt, 1)
    ub, eb = ol.integrate_sync_leap(u, e, dt)
    ub, eb = ol.integrate_leap_leap(ub, eb, dt, O.DRIFT_ADDITIVE)
    ub, eb = ol.integrate_leap_sync(ub, eb, dt, O.DRIFT_ADDITIVE)
    ub, eb = ol.integrate_sync_sync(ub, eb, dt, O.DRIFT_ADDITIVE)
    c = fresh().simulate_using_leapfrog_n_auto(integ, dt, 10)
    uc, ec = ol.simulate_symplectic_n(u, e, dt, 10, O.DRIFT_ADDITIVE, O.SCHEME_LEAPFROG)
    for state, (uo, eo) in ((a, (ua, ea)), (b, (ub, eb)), (c, (uc, ec))):
        dh = state.hamiltonian_total() - h
        assert abs(dh) < 1e-4
        assert abs(dh - (h_oracle(uo, eo) - h)) <= 1e-9


def test_reference_sim_cold():
    # src/test/mod.rs:429-453: a cold state with zero momenta does not move
    # under the leap transitions (size 10, beta 0.1, L = 10, D = 4, dt 0.1)
    announce("reference sim_cold")
    st = lqcd.LatticeStateDefault.new_cold(10.0, 0.1, 10, 4, lqcd.F64)
    ef = lqcd.LatticeStateEFSyncDefault.new_e_cold(st)
    u0, e0 = ef.link_matrix(), ef.e_field()
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE)
    leap = ef.simulate_to_leapfrog(integ, 0.1)
    assert np.array_equal(leap.e_field(), e0) and np.array_equal(leap.link_matrix(), u0)
    leap = leap.simulate_leap(integ, 0.1)
    assert np.array_equal(leap.e_field(), e0) and np.array_equal(leap.link_matrix(), u0)
    ef = leap.simulate_to_synchronous(integ, 0.1)
    assert np.array_equal(ef.e_field(), e0) and np.array_equal(ef.link_matrix(), u0)


def test_reference_sim_hamiltonian_and_gauss_law():
    # src/test/mod.rs:383-427 (the rayon integrator): one simulate_sync on
    # new_determinist states with momenta uniform in [-pi, pi) (L = 10, D = 4):
    # size 100, beta 1, dt 1e-4: H - H' < 0.01; size 1, beta 1, dt 1e-6:
    # Gauss's G(x) moves by less than 1e-3 entry by entry
    announce("reference sim_hamiltonian / gauss_law D=4 L=10")
    rng = np.random.default_rng(0x4578_93F4_4AB0_67F0)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE)
    for size, dt, check in ((100.0, 1e-4, "h"), (1.0, 1e-6, "g")):
        st = lqcd.LatticeStateDefault.new_determinist(size, 1.0, 10, 41, 4, lqcd.F64)
        e = rng.uniform(-math.pi, math.pi, (10 ** 4 * 4, 8))
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        h, g = ef.hamiltonian_total(), ef.gauss()
        ef = ef.simulate_sync(integ, dt)
        if check == "h":
            assert h - ef.hamiltonian_total() < 0.01
        else:
            assert np.max(np.abs(ef.gauss() - g)) < 1e-3
