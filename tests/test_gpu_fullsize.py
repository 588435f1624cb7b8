"""GPU checks at the workloads' full sizes through size-independent properties
(the oracle finishes the same sizes only in minutes):

  32^4 f64 / f32   MD reversibility: n leapfrog steps, E -> -E, n steps back
                   returns U and -E (exp drift, SU(3)-preserving); f32 against
                   f64 after the same steps; |dH| of a trajectory scales as dt^2
  48^3 x 96 f32    reversibility seen through the plaquette and H_E
  64 chains        two chains of the batched configuration against lattices of
                   their own with the same draws
  D=3 128^3 f64    Metropolis sweeps keep SU(3); a rerun with the same seed
                   gives the same accept counts and plaquette

Tolerances are stated at each assertion."""
import sys

import numpy as np
import pytest

import lqcd
from test_gpu_batch import accept_word

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def integrate(ef, dt, n):
    ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG), dt, n)


@pytest.mark.parametrize("precision,tol", [(lqcd.F64, 1e-10), (lqcd.F32, 2e-5)])
def test_md_reversibility_32x4(precision, tol):
    announce(f"32^4 MD reversibility precision={precision}")
    st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 32, 11, 4, precision)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 12, 0, project=False)
    u0, e0 = ef.link_matrix(), ef.e_field()
    integrate(ef, 0.025, 10)
    moved = np.max(np.abs(ef.link_matrix() - u0))
    assert moved > 1e-2
    ef.set_e_field(-ef.e_field())
    integrate(ef, 0.025, 10)
    # leapfrog with the exp drift is exactly reversible; rounding is what remains
    assert np.max(np.abs(ef.link_matrix() - u0)) <= tol
    assert np.max(np.abs(ef.e_field() + e0)) <= tol
    print(f"moved {moved:.3e}, back to {np.max(np.abs(ef.link_matrix() - u0)):.3e}")


def test_f32_against_f64_32x4():
    announce("32^4 f32 vs f64 after 10 steps")
    out = {}
    for precision in (lqcd.F64, lqcd.F32):
        st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 32, 13, 4, precision)
        ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 14, 0, project=False)
        integrate(ef, 0.025, 10)
        out[precision] = (st.average_trace_plaquette(), ef.hamiltonian_efield(), ef.link_matrix())
    p64, he64, u64 = out[lqcd.F64]
    p32, he32, u32 = out[lqcd.F32]
    # f32 storage and arithmetic: links to 1e-5, the averages far tighter
    assert np.max(np.abs(u32 - u64)) <= 1e-5
    assert abs(p32 - p64) <= 1e-7 * abs(p64)
    assert abs(he32 - he64) <= 1e-6 * abs(he64)


def test_dh_scales_as_dt_squared_32x4():
    announce("32^4 f64 |dH| ~ dt^2")
    dh = []
    for dt, n in ((0.01, 10), (0.005, 20)):
        st = lqcd.LatticeStateDefault.new_cold(1.0, 6.0, 32, 4, lqcd.F64)
        ef = lqcd.LatticeStateEFSyncDefault(st)
        ef.reset_e_field(16, 0, 0.5 / 6.0)
        hmc = lqcd.HybridMonteCarloDiagnostic(dt, n, lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
                                              1, project_gauss=False)
        dh.append(abs(hmc.trajectory(st, 0).delta_h))
    # second-order integrator over the same trajectory length: ratio 4 up to O(dt^2)
    assert 3.0 < dh[0] / dh[1] < 5.0, dh


def test_reversibility_48x3x96_f32():
    announce("48^3x96 f32 reversibility (plaquette, H_E)")
    # cold start (<Tr P> = 3): a hot start's average trace is ~1e-4 and no scale
    st = lqcd.LatticeStateDefault.new_cold(1.0, 6.2, [48, 48, 48, 96], 4, lqcd.F32)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 18, 0, project=False)
    p0, h0 = st.average_trace_plaquette(), ef.hamiltonian_efield()
    integrate(ef, 0.02, 8)
    p1 = st.average_trace_plaquette()
    assert abs(p1 - p0) > 1e-4 * abs(p0)
    # E -> -E on the device: one E-field round trip through the host
    ef.set_e_field(-ef.e_field())
    integrate(ef, 0.02, 8)
    assert abs(st.average_trace_plaquette() - p0) <= 1e-7 * abs(p0)
    assert abs(ef.hamiltonian_efield() - h0) <= 1e-6 * abs(h0)


def test_batched_chains_full_configuration():
    # the batched workload's 64 chains (16 each of 8^4, 12^4, 16^4, 20^4);
    # the first 8^4 and the last 20^4 chain against lattices of their own
    announce("64 batched chains vs single lattices")
    sizes = [8, 12, 16, 20]
    specs = [lqcd.ChainSpec([sizes[k // 16]] * 4, 5.6 + 0.6 * k / 63, 1000 + k) for k in range(64)]
    b = lqcd.BatchedChains(specs, 4, lqcd.F32)
    b.hot_start()
    b.config = 11   # per site, 256 threads, persistent: the arithmetic of MD variant 2
    dt, n_steps, n_traj = 0.01, 10, 2
    stats = b.hmc(dt, n_steps, n_traj, reunitarize=True)
    for i in (0, 63):
        ch = specs[i]
        st = lqcd.LatticeStateDefault.new_determinist(ch.size, ch.beta, ch.extents, ch.seed, 4, lqcd.F32)
        st.set_md_variant(2)
        ef = lqcd.LatticeStateEFSyncDefault(st)
        hmc = lqcd.HybridMonteCarloDiagnostic(dt, n_steps, lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP,
                                                                                lqcd.SCHEME_LEAPFROG), 1)
        for n in range(n_traj):
            ef.reset_e_field(ch.seed, n, 0.5 / ch.beta)
            ref = hmc.trajectory(st, accept_word(ch.seed, n))
            st.normalize_link_matrices()
            assert stats[n, i]["accepted"] == ref.accepted
            # the lattice of its own sums H in its f32 MD kernels, the batch
            # with its plaquette kernel
            tol_h = 1e-8 * abs(ref.h_old) + 1e-12
            assert abs(stats[n, i]["h_old"] - ref.h_old) <= tol_h
            assert abs(stats[n, i]["h_new"] - ref.h_new) <= tol_h
        # as in the small-chain test of the batch: same per-site arithmetic
        assert np.max(np.abs(b.link_matrix(i) - st.link_matrix())) <= 2e-5


def test_metropolis_128x3_f64_keeps_su3_and_reproduces():
    announce("D=3 128^3 f64 Metropolis")
    res = []
    for _ in range(2):
        st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 128, 19, 3, lqcd.F64)
        mh = lqcd.MetropolisHastingsSweep(1, 0.1, 20)
        counts = []
        for _ in range(3):
            mh.next_element(st)
            counts.append(mh.number_replace_last())
        res.append((counts, st.average_trace_plaquette(), st.su3_distance()))
    (c0, p0, d0), (c1, p1, d1) = res
    assert c0 == c1 and p0 == p1
    nlink = 128 ** 3 * 3
    assert all(0 < c < nlink for c in c0)
    # proposals are orthonormalised: the links stay in SU(3) to rounding
    assert d0 <= 64 * 2.0 ** -53
