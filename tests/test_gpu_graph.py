"""Leapfrog MD launches replayed from captured HIP graphs (lqcd_md_graph_set)
against the same launches made one by one: bit-identical links and momenta
after every replay, with eager work (reductions, reunitarisation, momentum
draws, accept/reject that moves the current buffer) between the replays, and
after parameter changes that force a new capture.  The profiling records count
one fused-kernel launch per MD step either way."""
import sys

import numpy as np
import pytest

import lqcd
from lqcd import _ffi as F

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def pair(precision, seed=3):
    out = []
    for graph in (False, True):
        st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, [8, 8, 8, 8], seed, 4, precision)
        st.set_md_graph(graph)
        out.append(st)
    return out


@pytest.mark.parametrize("precision,drift,scheme", [
    (lqcd.F64, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
    (lqcd.F32, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
    (lqcd.F64, lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE),
    (lqcd.F32, lqcd.DRIFT_EXP, lqcd.SCHEME_REFERENCE),
])
def test_graph_replays_match_eager(precision, drift, scheme):
    announce(f"MD graph vs eager precision={precision} drift={drift} scheme={scheme}")
    a, b = pair(precision)
    efa = lqcd.LatticeStateEFSyncDefault.new_random_e_state(a, 5, 0, project=False)
    efb = lqcd.LatticeStateEFSyncDefault.new_random_e_state(b, 5, 0, project=False)
    integ = lqcd.SymplecticEulerGpu(drift, scheme)
    for dt, n in ((0.02, 7), (0.02, 7), (0.02, 7), (0.015, 4), (0.02, 7)):
        efa.simulate_symplectic_n(integ, dt, n)
        efb.simulate_symplectic_n(integ, dt, n)
        assert np.array_equal(efa.link_matrix(), efb.link_matrix())
        assert np.array_equal(efa.e_field(), efb.e_field())
        assert a.average_trace_plaquette() == b.average_trace_plaquette()
        a.normalize_link_matrices()
        b.normalize_link_matrices()
    # trajectories: accepts move the current buffer, so replays start from
    # each of the three buffers
    ha = lqcd.HybridMonteCarloDiagnostic(0.01, 10, integ, 9, project_gauss=False)
    hb = lqcd.HybridMonteCarloDiagnostic(0.01, 10, integ, 9, project_gauss=False)
    accepted = 0
    for _ in range(8):
        a.monte_carlo_step(ha)
        b.monte_carlo_step(hb)
        sa, sb = ha.last_stats, hb.last_stats
        assert (sa.h_old, sa.h_new, sa.accepted) == (sb.h_old, sb.h_new, sb.accepted)
        assert np.array_equal(a.link_matrix(), b.link_matrix())
        accepted += sa.accepted
    assert accepted >= 1


def test_graph_profiling_counts_launches():
    announce("MD graph profiling counts")
    _, b = pair(lqcd.F32, seed=4)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(b, 6, 0, project=False)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    ef.simulate_symplectic_n(integ, 0.02, 6)
    ctx = b._dev.ctx
    ctx.profile_enable((1 << F.KERNEL_KICK_DRIFT) | (1 << F.KERNEL_KICK))
    for _ in range(3):
        ef.simulate_symplectic_n(integ, 0.02, 6)
    n, ms = ctx.profile_read(F.KERNEL_KICK_DRIFT)
    k, _ = ctx.profile_read(F.KERNEL_KICK)
    ctx.profile_enable(0)
    assert n == 18 and k == 3 and ms > 0.0


def test_profiling_sessions_reuse_events_with_fresh_counts():
    # a session whose replays count n launches per event pair, then a session
    # of eager kicks only: the pooled event pairs must count one launch each
    announce("profiling: second session after graph replays")
    _, b = pair(lqcd.F32, seed=5)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(b, 7, 0, project=False)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    ctx = b._dev.ctx
    ctx.profile_enable(1 << F.KERNEL_KICK_DRIFT)
    for _ in range(4):
        ef.simulate_symplectic_n(integ, 0.02, 10)
    n, _ = ctx.profile_read(F.KERNEL_KICK_DRIFT)
    assert n == 40
    ctx.profile_enable((1 << F.KERNEL_KICK_DRIFT) | (1 << F.KERNEL_KICK))
    for _ in range(5):
        integ.integrate_efield(ef, 0.01)
    k, ms = ctx.profile_read(F.KERNEL_KICK)
    n, _ = ctx.profile_read(F.KERNEL_KICK_DRIFT)
    ctx.profile_enable(0)
    assert (k, n) == (5, 0) and ms > 0.0
