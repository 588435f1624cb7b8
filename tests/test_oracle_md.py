"""Pins the oracle's integrators, HMC and Metropolis sweep (src/integrator/,
src/simulation/monte_carlo/) by the reference's own tolerances and invariants."""
import math

import numpy as np
import pytest

import oracle as O


def test_cold_state_invariant_under_md():
    # src/test/mod.rs:429-453: a cold state does not move (bitwise)
    l = O.OracleLattice(4, 3, 10.0)
    u = l.cold_start()
    e = np.zeros((l.nlink, 8))
    for mode in (O.DRIFT_ADDITIVE, O.DRIFT_EXP):
        for scheme in (O.SCHEME_REFERENCE, O.SCHEME_LEAPFROG):
            u2, e2 = l.simulate_symplectic_n(u, e, 0.1, 3, mode, scheme)
            assert np.array_equal(u2, u)
            assert np.array_equal(e2, e)


def test_zero_integration_error():
    # src/simulation/state.rs:480-482 -> MultiIntegrationError::ZeroIntegration
    l = O.OracleLattice(3, 2)
    with pytest.raises(ValueError):
        l.simulate_symplectic_n(l.cold_start(), np.zeros((l.nlink, 8)), 0.1, 0)


def test_leap_frog_bound():
    # src/test/mod.rs:678-698: D=4, L=4, beta=1, size=1000, one step dt=0.01,
    # |dH| < 1e-5 (E drawn N(0, 0.5/beta) as state.rs:1097)
    l = O.OracleLattice(4, 4, 1000.0)
    u = l.hot_start(0)
    e = l.random_efield(1, 0.5)
    beta = 1.0
    h1 = l.hamiltonian_links(u, beta) + l.hamiltonian_efield(e, beta)
    u2, e2 = l.simulate_symplectic_n(u, e, 0.01, 1)
    h2 = l.hamiltonian_links(u2, beta) + l.hamiltonian_efield(e2, beta)
    assert abs(h1 - h2) < 1e-5


def test_integrator_bound_after_mh():
    # src/test/integrator.rs:9-67: D=3, L=4, beta=8, cold + 10 MH sweeps,
    # dt = 1e-4, 10 symplectic steps: |dH| < 1e-4
    l = O.OracleLattice(3, 4)
    beta = 8.0
    u = l.cold_start()
    for k in range(10):
        props = l.mh_proposals(100 + k, 0.1)
        rr = O.OracleRng(300 + k)
        unis = np.array([rr.next_u64() for _ in range(l.nlink)], dtype=np.uint64)
        u, _ = l.metropolis_sweep(u, beta, props, unis, O.ORDER_LEXICOGRAPHIC)
    e = l.random_efield(5, 0.5 / beta)
    h = l.hamiltonian_links(u, beta) + l.hamiltonian_efield(e, beta)
    u2, e2 = l.simulate_symplectic_n(u, e, 1e-4, 10)
    h2 = l.hamiltonian_links(u2, beta) + l.hamiltonian_efield(e2, beta)
    assert abs(h - h2) < 1e-4


def _traj_dh(l, u, e, beta, dt, n, mode, scheme=O.SCHEME_REFERENCE):
    h1 = l.hamiltonian_links(u, beta) + l.hamiltonian_efield(e, beta)
    u2, e2 = l.simulate_symplectic_n(u, e, dt, n, mode, scheme)
    return l.hamiltonian_links(u2, beta) + l.hamiltonian_efield(e2, beta) - h1


def test_exp_mode_dh_scales_as_dt_squared():
    # leapfrog on the group manifold: |dH| ~ dt^2 at fixed trajectory length
    l = O.OracleLattice(4, 3)
    beta = 6.0
    u = l.hot_start(3)
    e = l.random_efield(4, 1.0 / math.sqrt(beta))
    d1 = abs(_traj_dh(l, u, e, beta, 0.04, 5, O.DRIFT_EXP))
    d2 = abs(_traj_dh(l, u, e, beta, 0.02, 10, O.DRIFT_EXP))
    d3 = abs(_traj_dh(l, u, e, beta, 0.01, 20, O.DRIFT_EXP))
    assert 3.0 < d1 / d2 < 5.0 and 3.0 < d2 / d3 < 5.0, (d1, d2, d3)


def test_exp_mode_reversible_and_unitary():
    l = O.OracleLattice(4, 3)
    u = l.hot_start(6)
    e = l.random_efield(7, 0.4)
    u2, e2 = l.simulate_symplectic_n(u, e, 0.02, 10, O.DRIFT_EXP)
    u3, e3 = l.simulate_symplectic_n(u2, -e2, 0.02, 10, O.DRIFT_EXP)
    assert np.max(np.abs(u3 - u)) < 1e-11
    assert np.max(np.abs(-e3 - e)) < 1e-11
    m = O.to_np(u2)
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) < 1e-12


def test_additive_mode_first_order_drift():
    # the reference drift U + dt D(U,E) agrees with exp(i dt sqrt6 E) U to O(dt^2)
    l = O.OracleLattice(4, 2)
    u = l.hot_start(8)
    e = l.random_efield(9, 1.0)
    for dt in (1e-2, 1e-3):
        ua = l.integrate_link(u, e, dt, O.DRIFT_ADDITIVE)
        ue = l.integrate_link(u, e, dt, O.DRIFT_EXP)
        err = np.max(np.abs(ua - ue))
        assert err < 20 * dt * dt
        assert err > 0.05 * dt * dt


def test_leapfrog_scheme_equals_reference_scheme():
    # merging the two half kicks changes only the rounding
    l = O.OracleLattice(4, 3)
    u = l.hot_start(10)
    e = l.random_efield(11, 0.3)
    for mode in (O.DRIFT_ADDITIVE, O.DRIFT_EXP):
        ua, ea = l.simulate_symplectic_n(u, e, 0.01, 5, mode, O.SCHEME_REFERENCE)
        ub, eb = l.simulate_symplectic_n(u, e, 0.01, 5, mode, O.SCHEME_LEAPFROG)
        assert np.max(np.abs(ua - ub)) < 1e-13
        assert np.max(np.abs(ea - eb)) < 1e-13


def test_hmc_accept_rule():
    # hybrid_monte_carlo.rs:584-613: p = clamp(exp(H_old - H_new)); accept on
    # the Bernoulli threshold; on reject the old links stay
    l = O.OracleLattice(4, 2)
    u = l.hot_start(12)
    e = l.random_efield(13, 0.5 / 6.0)
    u_acc, st = l.hmc_trajectory(u, e, 6.0, 0.01, 5, 0)
    assert st.accepted == 1
    assert math.isclose(st.delta_h, st.h_new - st.h_old)
    assert st.probability == min(1.0, math.exp(-st.delta_h))
    u_ref, _ = l.simulate_symplectic_n(u, e, 0.01, 5)
    assert np.array_equal(u_acc, u_ref)
    if st.probability < 1.0:
        u_rej, st2 = l.hmc_trajectory(u, e, 6.0, 0.01, 5, 2**64 - 1)
        assert st2.accepted == 0 and np.array_equal(u_rej, u)


def test_mh_identity_proposals_always_accept():
    l = O.OracleLattice(4, 2)
    u = l.hot_start(14)
    props = np.tile(O.from_np(np.eye(3)), (1, l.nlink, 1))
    unis = np.full(l.nlink, 2**64 - 1, dtype=np.uint64)
    for order in (O.ORDER_LEXICOGRAPHIC, O.ORDER_CHECKERBOARD):
        u2, st = l.metropolis_sweep(u, 6.0, props, unis, order)
        assert np.array_equal(u2, u)
        assert st.number_replace_last == l.nlink
        assert st.prob_replace_mean == 1.0


def test_mh_sweep_lowers_action_at_large_beta():
    l = O.OracleLattice(4, 4)
    u = l.hot_start(15)
    beta = 6.0
    h0 = l.hamiltonian_links(u, beta)
    rr = O.OracleRng(16)
    for k in range(3):
        props = l.mh_proposals(17 + k, 0.1)
        unis = np.array([rr.next_u64() for _ in range(l.nlink)], dtype=np.uint64)
        u, st = l.metropolis_sweep(u, beta, props, unis, O.ORDER_CHECKERBOARD)
        assert 0 < st.number_replace_last <= l.nlink
        assert 0.0 < st.prob_replace_mean <= 1.0
    assert l.hamiltonian_links(u, beta) < h0


def test_mh_checkerboard_rejects_odd_extent():
    l = O.OracleLattice(3, 3)
    with pytest.raises(ValueError):
        l.metropolis_sweep(l.cold_start(), 6.0, np.zeros((1, l.nlink, 18)),
                           np.zeros(l.nlink, dtype=np.uint64), O.ORDER_CHECKERBOARD)


def test_normalize_link_matrices():
    # state.rs:754-756: every link becomes SU(3); SU(3) links are fixed points
    l = O.OracleLattice(3, 3)
    u = l.hot_start(18)
    noisy = u + 1e-3 * np.random.default_rng(0).normal(size=u.shape)
    n = l.normalize_link_matrices(noisy)
    m = O.to_np(n)
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) < 1e-14
    assert np.max(np.abs(np.linalg.det(m) - 1)) < 1e-14
    assert np.max(np.abs(l.normalize_link_matrices(u) - u)) < 1e-15


def test_project_to_gauss_reduces_violation():
    # field.rs:1265-1294: converges below N*4*8*10*eps for a small field
    l = O.OracleLattice(3, 3)
    u = l.hot_start(19)
    e = l.random_efield(20, 0.01)
    before = l.gauss_sum_div(u, e)
    rc, e2, it = l.project_to_gauss(u, e, 2000)
    assert rc == 0
    after = l.gauss_sum_div(u, e2)
    assert after <= np.finfo(float).eps * l.nsite * 320
    assert after < before


def test_device_momenta_statistics():
    l = O.OracleLattice(4, 4)
    e = l.device_momenta(1234, 7, 0.5)
    assert abs(np.mean(e)) < 0.02
    assert abs(np.std(e) - 0.5) < 0.01
    e2 = l.device_momenta(1234, 8, 0.5)
    assert not np.allclose(e, e2)


def test_device_mh_stream_matches_reference_distribution():
    # The device Metropolis stream must draw orthonormalize(random_su3_close_to_unity)
    # (src/su3.rs:385-398) with the reference's distribution: every proposal is SU(3),
    # and the mean of each real component over 16384 proposals agrees with the
    # xoshiro-driven restatement within 5 standard errors (a dropped 2w-1 shift
    # moves one by ~40 sigma).  The Bernoulli(0.5) adjoint makes E|X_ij|^2 = E|X_ji|^2
    # exactly; dropping it breaks that by ~20 sigma at this sample size.
    l = O.OracleLattice(4, 4)
    n_update, spread = 16, 0.3
    p, w = l.device_mh_inputs(0xC0FFEE, 5, spread, n_update)
    q = l.mh_proposals(0xC0FFEE, spread, n_update)
    pm = p.reshape(-1, 18)
    qm = q.reshape(-1, 18)
    mats = O.to_np(pm)
    eye = np.eye(3)
    assert np.max(np.abs(mats @ np.conj(np.swapaxes(mats, -1, -2)) - eye)) < 1e-13
    assert np.max(np.abs(np.linalg.det(mats) - 1.0)) < 1e-13
    err = np.sqrt(np.var(pm, axis=0) / len(pm) + np.var(qm, axis=0) / len(qm)) + 1e-12
    assert np.max(np.abs(pm.mean(axis=0) - qm.mean(axis=0)) / err) < 5.0
    assert abs(np.var(pm) - np.var(qm)) < 0.05 * np.var(qm)
    a2 = np.abs(mats) ** 2
    d = (a2[:, 1, 0] + a2[:, 2, 1] + a2[:, 0, 2]) - (a2[:, 0, 1] + a2[:, 1, 2] + a2[:, 2, 0])
    assert abs(d.mean()) < 5.0 * d.std() / math.sqrt(len(d))
    # accept words uniform on [0, 2^64)
    frac = w.astype(np.float64) / 2.0**64
    assert abs(frac.mean() - 0.5) < 5.0 / math.sqrt(12 * len(frac))
    # reproducible per (seed, counter); a new counter is a new stream
    p2, w2 = l.device_mh_inputs(0xC0FFEE, 5, spread, n_update)
    assert np.array_equal(p, p2) and np.array_equal(w, w2)
    p3, w3 = l.device_mh_inputs(0xC0FFEE, 6, spread, n_update)
    assert not np.allclose(p, p3) and not np.array_equal(w, w3)
    # update j of a link does not depend on how many updates are drawn
    p1, _ = l.device_mh_inputs(0xC0FFEE, 5, spread, 1)
    assert np.array_equal(p1[0], p[0])


def test_parallel_first_call_matches_serial(tmp_path):
    # regression: the generator tables are initialised on first use, which
    # happens inside OpenMP loops; a fresh process with many threads must give
    # the single-thread result
    import os
    import subprocess
    import sys
    out = tmp_path / "par.npy"
    code = (
        "import numpy as np, oracle as O\n"
        "O.set_threads(32)\n"
        "l = O.OracleLattice(4, 4)\n"
        "u = l.hot_start(O.SEED_RNG)\n"
        "e = l.random_efield(1, 0.5 / 6.0)\n"
        "u2, e2 = l.simulate_symplectic_n(u, e, 0.01, 2)\n"
        f"np.save({str(out)!r}, np.concatenate([u2.ravel(), e2.ravel()]))\n"
    )
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([sys.executable, "-c", code], cwd=root, check=True)
    par = np.load(out)
    O.set_threads(1)
    try:
        l = O.OracleLattice(4, 4)
        u = l.hot_start(O.SEED_RNG)
        e = l.random_efield(1, 0.5 / 6.0)
        u2, e2 = l.simulate_symplectic_n(u, e, 0.01, 2)
    finally:
        O.set_threads(0)
    assert np.array_equal(par, np.concatenate([u2.ravel(), e2.ravel()]))


@pytest.mark.parametrize("mode", [O.DRIFT_ADDITIVE, O.DRIFT_EXP])
def test_leap_transitions_chain_is_the_leapfrog_scheme(mode):
    # simulate_using_leapfrog_n (state.rs:321-358) = sync_leap, n - 1 leap_leap,
    # leap_sync; the C loop of the leapfrog scheme is written independently
    ol = O.OracleLattice(3, [4, 4, 6])
    u = ol.hot_start(3)
    e = ol.random_efield(4, 0.3)
    a, b = ol.integrate_sync_leap(u, e, 0.02)
    for _ in range(4):
        a, b = ol.integrate_leap_leap(a, b, 0.02, mode)
    a, b = ol.integrate_leap_sync(a, b, 0.02, mode)
    u_ref, e_ref = ol.simulate_symplectic_n(u, e, 0.02, 5, mode, O.SCHEME_LEAPFROG)
    assert np.array_equal(a, u_ref) and np.array_equal(b, e_ref)


def test_sync_sync_is_first_order():
    # integrate_sync_sync (symplectic_euler_rayon.rs:127-144) is an explicit
    # Euler step: its energy error over a fixed time falls like dt, the
    # symplectic step's like dt^2
    ol = O.OracleLattice(3, [4, 4, 4])
    u0 = ol.hot_start(5)
    e0 = ol.random_efield(6, 0.3)
    errs = []
    for dt, n in ((0.02, 5), (0.01, 10)):
        u, e = u0, e0
        h0 = ol.hamiltonian_links(u, 6.0) + ol.hamiltonian_efield(e, 6.0)
        for _ in range(n):
            u, e = ol.integrate_sync_sync(u, e, dt, O.DRIFT_EXP)
        errs.append(abs(ol.hamiltonian_links(u, 6.0) + ol.hamiltonian_efield(e, 6.0) - h0))
    assert 1.6 < errs[0] / errs[1] < 2.6, errs
