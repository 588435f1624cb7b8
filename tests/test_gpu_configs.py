"""The HMC workloads at their own sizes against the FP64 oracle.

32^4 mixed precision (D = 4, beta = 6.0, 40 leapfrog steps of dt = 0.025 with
the su(3)-exponential link update, momenta N(0, 0.5/beta)):
  One hot start of the oracle (f64, Xoshiro256++ random SU(3) per link) and
  one momentum field of the oracle are uploaded into an FP32 lattice and into
  an FP64 lattice.  The same trajectory runs on the device in both storage
  precisions and in the oracle in f64 (hybrid_monte_carlo.rs:584-613,
  state.rs:1370-1448).  Per observable the device is compared with the oracle:
  H_old, H_new, dH, the average plaquette and H_E of the proposal, the largest
  link difference and the accept decision.
    FP64 storage   dH within 1e-9 |H| + 1e-10 (the f64 parity bound), H_old
                   and <P> relative 1e-13, links 1e-10, same accept decision.
    FP32 storage   the links and momenta start rounded to f32, the force and
                   drift run in f32, H, <P> and the accept rule in f64:
                   H_old relative 1e-10, dH within 5e-6 |H_old|, <P> relative
                   1e-6, H_E relative 1e-6, links 2e-4, same accept decision.
                   (measured on MI355X: the error table in the test output)

16^4 FP64 HMC (D = 4, beta = 5.7, hot start, 50 trajectories of 20 steps of
dt = 0.01 with the reference's integrator: kick-drift-kick, additive drift):
  momenta from the device stream, which the oracle restates; the first two
  trajectories are repeated in the oracle from identical momenta and accept
  words: dH within 1e-9 |H| + 1e-10, the same accept decision, links 1e-11.
  The momenta are not projected onto the Gauss surface: the reference's
  projection stalls once the additive drift has moved the links off SU(3).
"""
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu

C3_BETA, C3_DT, C3_STEPS, C3_L = 6.0, 0.025, 40, 32
C3_SEED = O.SEED_RNG
C3_UNIFORM = 1 << 63


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c3_oracle():
    """Start state and the oracle's f64 trajectory at 32^4."""
    announce(f"32^4 oracle trajectory on {O.get_threads()} threads")
    ol = O.OracleLattice(4, C3_L)
    u = ol.hot_start(C3_SEED)
    e = ol.random_efield(C3_SEED + 1, 0.5 / C3_BETA)
    h_old = ol.hamiltonian_links(u, C3_BETA) + ol.hamiltonian_efield(e, C3_BETA)
    u_new, e_new = ol.simulate_symplectic_n(u, e, C3_DT, C3_STEPS, O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    h_u = ol.hamiltonian_links(u_new, C3_BETA)
    h_e = ol.hamiltonian_efield(e_new, C3_BETA)
    p = O.replacement_probability(h_old, h_u + h_e)
    return {
        "ol": ol, "u": u, "e": e, "u_new": u_new,
        "h_old": h_old, "h_new": h_u + h_e, "dh": h_u + h_e - h_old, "h_e": h_e,
        "plaquette": ol.average_trace_plaquette(u_new).real / 3.0,
        "accepted": O.bernoulli_accept(p, C3_UNIFORM),
    }


def c3_device(ref, precision):
    """The same trajectory on the device: observables of the proposal (an
    integration of the uploaded state) and the decision of lqcd_hmc_trajectory."""
    lat = lqcd.LatticeCyclic(1.0, C3_L, 4)
    st = lqcd.LatticeStateDefault(lat, C3_BETA, ref["u"], precision)
    ef = lqcd.LatticeStateEFSyncDefault(st, ref["e"])
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    h_old = ef.hamiltonian_total()
    ef.simulate_symplectic_n(integ, C3_DT, C3_STEPS)
    out = {"h_old": h_old, "h_new": ef.hamiltonian_total(), "h_e": ef.hamiltonian_efield(),
           "plaquette": st.average_trace_plaquette().real / 3.0, "r12": st.md_reconstruct_last()}
    out["dh"] = out["h_new"] - out["h_old"]
    out["max_du"] = float(np.max(np.abs(st.link_matrix() - ref["u_new"])))
    # the accept decision through the HMC entry point, from the same start
    st.set_link_matrix(ref["u"])
    ef.set_e_field(ref["e"])
    hmc = lqcd.HybridMonteCarloDiagnostic(C3_DT, C3_STEPS, integ, 1, project_gauss=False)
    s = hmc.trajectory(st, C3_UNIFORM)
    out["hmc"] = (s.h_old, s.h_new, s.delta_h, s.accepted)
    return out


def report(name, got, ref):
    rows = []
    for key in ("h_old", "h_new", "dh", "plaquette", "h_e"):
        err = abs(got[key] - ref[key])
        rows.append(f"{key} {got[key]:.10e} oracle {ref[key]:.10e} abs {err:.3e} rel {err / abs(ref[key]):.3e}")
    rows.append(f"max|dU| {got['max_du']:.3e}  accepted {got['hmc'][3]} oracle {int(ref['accepted'])}")
    print(f"[{name}] " + "\n  ".join(rows), flush=True)


@pytest.mark.timeout(900)
def test_c3_fp64_trajectory_vs_oracle(c3_oracle):
    announce("32^4 FP64 trajectory vs oracle")
    ref = c3_oracle
    got = c3_device(ref, lqcd.F64)
    report("32^4 f64", got, ref)
    tol = 1e-9 * abs(ref["h_old"]) + 1e-10
    assert abs(got["h_old"] - ref["h_old"]) <= 1e-13 * abs(ref["h_old"])
    assert abs(got["dh"] - ref["dh"]) <= tol
    assert abs(got["plaquette"] - ref["plaquette"]) <= 1e-13 * abs(ref["plaquette"])
    assert got["max_du"] <= 1e-10
    h_old, h_new, dh, acc = got["hmc"]
    assert abs(dh - ref["dh"]) <= tol and acc == int(ref["accepted"])


@pytest.mark.timeout(900)
def test_c3_fp32_mixed_trajectory_vs_oracle(c3_oracle):
    announce("32^4 FP32-mixed trajectory vs oracle")
    ref = c3_oracle
    got = c3_device(ref, lqcd.F32)
    report("32^4 f32-mixed", got, ref)
    assert got["r12"]
    assert abs(got["h_old"] - ref["h_old"]) <= 1e-10 * abs(ref["h_old"])
    assert abs(got["dh"] - ref["dh"]) <= 5e-6 * abs(ref["h_old"])
    assert abs(got["plaquette"] - ref["plaquette"]) <= 1e-6 * abs(ref["plaquette"])
    assert abs(got["h_e"] - ref["h_e"]) <= 1e-6 * abs(ref["h_e"])
    assert got["max_du"] <= 2e-4
    h_old, h_new, dh, acc = got["hmc"]
    assert abs(dh - ref["dh"]) <= 5e-6 * abs(ref["h_old"]) and acc == int(ref["accepted"])


@pytest.mark.timeout(300)
def test_c2_fp64_hmc_50_trajectories():
    announce("16^4 FP64 HMC, 50 trajectories, first two vs oracle")
    beta, dt, n, seed = 5.7, 0.01, 20, 0xC2C2_5EED
    ol = O.OracleLattice(4, 16)
    u = ol.hot_start(O.SEED_RNG)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 16, 4), beta, u, lqcd.F64)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE)
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, n, integ, seed, project_gauss=False)
    # the same seed gives the same accept words
    twin = lqcd.HybridMonteCarloDiagnostic(dt, n, integ, seed, project_gauss=False)
    dhs, acc = [], 0
    for k in range(50):
        if k < 2:
            e = ol.device_momenta(seed, k, 0.5 / beta)
            u, ref = ol.hmc_trajectory(u, e, beta, dt, n, twin._next_u64(), O.DRIFT_ADDITIVE,
                                       O.SCHEME_REFERENCE)
        hmc.next_element(st)
        s = hmc.last_stats
        dhs.append(s.delta_h)
        acc += s.accepted
        if k < 2:
            tol = 1e-9 * abs(ref.h_old) + 1e-10
            print(f"[16^4 trajectory {k}] dH {s.delta_h:.10e} oracle {ref.delta_h:.10e} "
                  f"accepted {s.accepted} oracle {ref.accepted}", flush=True)
            assert abs(s.h_old - ref.h_old) <= tol
            assert abs(s.delta_h - ref.delta_h) <= tol
            assert s.accepted == ref.accepted
            assert np.max(np.abs(st.link_matrix() - u)) <= 1e-11
    dhs = np.array(dhs)
    print(f"[16^4 HMC] acceptance {acc / 50:.2f}, mean |dH| {np.mean(np.abs(dhs)):.4f}, "
          f"<P> {st.average_trace_plaquette().real / 3.0:.8f}", flush=True)
    # from this hot start the reference's integrator gives dH of about +3 per
    # trajectory at dt = 0.01, so most (often all) of the 50 are rejected
    assert np.all(np.isfinite(dhs))
