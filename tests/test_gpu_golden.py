"""The device path (include/lqcd.h through the lqcd mirror) against the golden
fixtures of tests/golden (the oracle's f64 results, frozen with their inputs):
no oracle call at test time.

Tolerances (f64 storage), as in the live parity tests:
  plaquette / H_U / H_E          relative 1e-13
  derivative_e, G, F, B          absolute 1e-12 * max|value|
  five MD steps                  max |dU|, |dE| <= 1e-12
  HMC trajectory                 |dH - dH_golden| <= 1e-9 |H| + 1e-10, same accept, links 1e-11
  Metropolis sweep               identical accept count, links 1e-12
  heat bath / overrelaxation     links 1e-10 / 1e-11
  reunitarisation                1e-14
FP32 storage (the leapfrog from the f64 fixture, rounded once on upload): 1e-5.
"""
import os
import sys

import numpy as np
import pytest

import lqcd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def d4():
    return np.load(os.path.join(GOLDEN, "d4_l3.npz"))


@pytest.fixture(scope="module")
def d3():
    return np.load(os.path.join(GOLDEN, "d3_l4.npz"))


def state(g, u=None, precision=lqcd.F64):
    lat = lqcd.LatticeCyclic(float(g["cfg_size"]), int(g["cfg_L"]), int(g["cfg_dim"]))
    return lqcd.LatticeStateDefault(lat, float(g["cfg_beta"]), g["u"] if u is None else u, precision)


def close(got, want, rel):
    return np.max(np.abs(got - want)) <= rel * max(1.0, np.max(np.abs(want)))


def test_golden_observables_and_force(d4):
    announce("golden D=4 3^4: plaquette, H, derivative_e, Gauss")
    st = state(d4)
    p = st.average_trace_plaquette()
    assert abs(p.real - d4["plaquette"][0]) <= 1e-13 * 3 and abs(p.imag - d4["plaquette"][1]) <= 1e-13 * 3
    assert abs(st.hamiltonian_links() - d4["h_links"]) <= 1e-13 * abs(d4["h_links"])
    ef = lqcd.LatticeStateEFSyncDefault(st, d4["e"])
    assert abs(ef.hamiltonian_efield() - d4["h_efield"]) <= 1e-13 * abs(d4["h_efield"])
    assert close(ef.derivative_e(), d4["derivative_e"], 1e-12)
    assert close(ef.gauss(), d4["gauss"], 1e-12)


@pytest.mark.parametrize("key,drift,scheme", [("reference", lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE),
                                              ("leapfrog", lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)])
def test_golden_integration(d4, key, drift, scheme):
    announce(f"golden D=4 3^4: {key} integration")
    st = state(d4)
    ef = lqcd.LatticeStateEFSyncDefault(st, d4["e"])
    ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(drift, scheme), float(d4["cfg_dt"]), int(d4["cfg_n_md"]))
    assert np.max(np.abs(ef.link_matrix() - d4["u_" + key])) <= 1e-12
    assert np.max(np.abs(ef.e_field() - d4["e_" + key])) <= 1e-12


def test_golden_integration_fp32(d4):
    announce("golden D=4 3^4: leapfrog in FP32 storage")
    st = state(d4, precision=lqcd.F32)
    ef = lqcd.LatticeStateEFSyncDefault(st, d4["e"])
    ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG), float(d4["cfg_dt"]),
                             int(d4["cfg_n_md"]))
    assert np.max(np.abs(ef.link_matrix() - d4["u_leapfrog"])) <= 1e-5
    assert np.max(np.abs(ef.e_field() - d4["e_leapfrog"])) <= 1e-5


def test_golden_hmc_trajectory(d4):
    announce("golden D=4 3^4: HMC trajectory")
    st = state(d4)
    lqcd.LatticeStateEFSyncDefault(st, d4["e"])
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc = lqcd.HybridMonteCarloDiagnostic(float(d4["cfg_dt"]), int(d4["cfg_n_hmc"]), integ, 1)
    got = hmc.trajectory(st, int(d4["cfg_uniform"]))
    h_old, h_new, dh, prob, acc = d4["hmc"]
    tol = 1e-9 * abs(h_old) + 1e-10
    assert abs(got.h_old - h_old) <= tol and abs(got.delta_h - dh) <= tol
    assert got.accepted == bool(acc)
    assert np.max(np.abs(st.link_matrix() - d4["u_hmc"])) <= 1e-11


def test_golden_metropolis_sweep(d3):
    announce("golden D=3 4^3: Metropolis sweep with stored proposals")
    st = state(d3)
    mh = lqcd.MetropolisHastingsSweep(1, float(d3["cfg_spread"]))
    mh.sweep(st, d3["mh_proposals"], d3["mh_uniforms"])
    assert mh.number_replace_last() == int(d3["metropolis"][1])
    assert abs(mh.prob_replace_mean() - d3["metropolis"][0]) <= 1e-12
    assert np.max(np.abs(st.link_matrix() - d3["u_metropolis"])) <= 1e-12


def test_golden_heat_bath_and_overrelaxation(d3):
    announce("golden D=3 4^3: heat bath, overrelaxation (both)")
    st = state(d3)
    hb = lqcd.HeatBathSweep(int(d3["cfg_hb_seed"]))
    st.monte_carlo_step(hb)
    assert hb.exhausted_last() == 0
    assert np.max(np.abs(st.link_matrix() - d3["u_heat_bath"])) <= 1e-10
    st = state(d3)
    st.monte_carlo_step(lqcd.OverrelaxationSweepRotation())
    assert np.max(np.abs(st.link_matrix() - d3["u_overrelax"])) <= 1e-11
    st = state(d3)
    st.monte_carlo_step(lqcd.OverrelaxationSweepReverse())
    assert np.max(np.abs(st.link_matrix() - d3["u_overrelax_reverse"])) <= 1e-11


def test_golden_reunitarize_and_fields(d3):
    announce("golden D=3 4^3: reunitarisation, F_mu_nu, magnetic field")
    st = state(d3, u=d3["u_noisy"])
    st.normalize_link_matrices()
    assert np.max(np.abs(st.link_matrix() - d3["u_normalized"])) <= 1e-14
    st = state(d3)
    assert close(st.f_mu_nu_field(), d3["f_mu_nu"], 1e-12)
    assert close(st.magnetic_field_vec_field(), d3["magnetic_field"], 1e-12)
