"""The reference's outcome for a NaN in the links, on the device.

HybridMonteCarloDiagnostic (hybrid_monte_carlo.rs:584-600): the accept
probability is (H_old - H_new).exp().min(1).max(0); Rust's f64::min returns
the other operand when one is NaN, so a NaN dH gives p = 1 and the trajectory
is accepted without a draw.  The NaN spreads under the MD one neighbourhood
per force evaluation, link by link (every entry of a link the NaN reaches is
NaN).

MetropolisHastingsSweep (metropolis_hastings_sweep.rs:155-164): the same
clamp makes p = 1 for every link whose staple holds the NaN, so each of them
takes its (finite) proposal; the NaN link itself takes R U = NaN.  The NaN
does not spread.

Both are compared with the oracle on the same inputs: accept decisions and
counts equal, the same links NaN, the finite links within the f64 parity
bounds.
"""
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def nan_state(dim, ext, beta, bad_link):
    ol = O.OracleLattice(dim, ext)
    u = ol.hot_start(O.SEED_RNG)
    u[bad_link] = np.nan
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, ext, dim), beta, u)
    return ol, u, st


def same_nan_links(got, ref, tol):
    bad_got = np.isnan(got).any(axis=1)
    bad_ref = np.isnan(ref).any(axis=1)
    assert np.array_equal(bad_got, bad_ref)
    # link by link: a link is either wholly NaN or wholly finite
    assert np.array_equal(np.isnan(got).all(axis=1), bad_got)
    assert np.max(np.abs(got[~bad_got] - ref[~bad_ref])) <= tol
    return int(bad_got.sum())


@pytest.mark.parametrize("drift,scheme,n", [
    (lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 1),
    (lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 2),
    (lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE, 2),
])
@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
def test_hmc_nan_link_is_accepted(drift, scheme, n, precision):
    announce(f"HMC NaN link drift={drift} scheme={scheme} n={n} precision={precision}")
    beta, dt = 6.0, 0.01
    ol, u, st = nan_state(4, 4, beta, 37)
    if precision == lqcd.F32:
        u = u.astype(np.float32).astype(np.float64)
        st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 4, 4), beta, u, lqcd.F32)
    e = ol.random_efield(7, 0.5 / beta)
    lqcd.LatticeStateEFSyncDefault(st, e)
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, n, lqcd.SymplecticEulerGpu(drift, scheme), 1)
    # accept word 0 would accept any p > 0; the largest word accepts only p = 1
    got = hmc.trajectory(st, 2**64 - 1)
    u_ref, ref = ol.hmc_trajectory(u, e.astype(np.float32).astype(np.float64) if precision == lqcd.F32 else e,
                                   beta, dt, n, 2**64 - 1, drift, scheme)
    assert ref.accepted == 1 and ref.probability == 1.0 and np.isnan(ref.delta_h)
    assert got.accepted == 1 and got.probability == 1.0 and np.isnan(got.delta_h)
    assert hmc.has_replace_last() and hmc.prob_replace_last() == 1.0
    tol = 1e-12 if precision == lqcd.F64 else 1e-5
    bad = same_nan_links(st.link_matrix(), u_ref, tol)
    assert 1 < bad < ol.nlink


@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
def test_metropolis_nan_link(precision):
    announce(f"Metropolis NaN link precision={precision}")
    beta = 6.0
    ol, u, st = nan_state(4, 4, beta, 37)
    if precision == lqcd.F32:
        u = u.astype(np.float32).astype(np.float64)
        st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 4, 4), beta, u, lqcd.F32)
    mh = lqcd.MetropolisHastingsSweep(1, 0.1)
    rr = O.OracleRng(33)
    for sweep in range(2):
        props = ol.mh_proposals(40 + sweep, 0.1, 1)
        unis = np.array([rr.next_u64() for _ in range(ol.nlink)], dtype=np.uint64)
        mh.sweep(st, props, unis)
        u, ref = ol.metropolis_sweep(u, beta, props, unis, O.ORDER_CHECKERBOARD)
        got = st.link_matrix()
        if precision == lqcd.F64:
            assert mh.number_replace_last() == ref.number_replace_last
            assert abs(mh.prob_replace_mean() - ref.prob_replace_mean) <= 1e-12
            assert same_nan_links(got, u, 1e-12) == 1
        else:
            # f64 arithmetic on f32 storage: a decision with p within rounding
            # of the accept word may differ; the oracle continues from the
            # stored links
            assert abs(mh.number_replace_last() - ref.number_replace_last) <= 2
            bad = np.isnan(got).any(axis=1)
            assert np.array_equal(bad, np.isnan(u).any(axis=1)) and bad.sum() == 1
            close = np.max(np.abs(got[~bad] - u[~bad]), axis=1) <= 1e-5
            assert close.mean() >= 0.99
            u = got
    assert np.isnan(st.link_matrix()[37]).all()
