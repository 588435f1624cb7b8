"""GPU parity: the gfx950 library (through include/lqcd.h) against the CPU oracle
on identical seeded inputs.

Tolerances (f64 path):
  plaquette / H          relative 1e-13
  force (derivative_e)   absolute 1e-12 * max|F|
  one MD step            max |dU|, |dE| <= 1e-12
  trajectory dH          |dH_gpu - dH_oracle| <= 1e-9 |H| + 1e-10, same accept decision
  Metropolis sweep       identical accept count, links max |dU| <= 1e-12
The f32 storage path is checked against the same f64 oracle with the bounds
stated in each test.
"""
import math
import sys

import numpy as np
import pytest

import oracle as O
import lqcd
from lqcd import _ffi as F

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def make_state(dim, ext, beta=6.0, precision=lqcd.F64, size=1.0, seed=O.SEED_RNG, cold=False):
    ol = O.OracleLattice(dim, ext, size)
    u = ol.cold_start() if cold else ol.hot_start(seed)
    lat = lqcd.LatticeCyclic(size, ext, dim)
    st = lqcd.LatticeStateDefault(lat, beta, u, precision)
    return ol, u, st


SHAPES = [(4, [4, 4, 4, 4]), (3, [6, 6, 6]), (4, [4, 6, 2, 8]), (2, [8, 6])]


@pytest.mark.parametrize("dim,ext", SHAPES)
def test_links_and_efield_roundtrip(dim, ext):
    announce(f"roundtrip D={dim} ext={ext}")
    ol, u, st = make_state(dim, ext)
    assert np.array_equal(st.link_matrix(), u)
    e = ol.random_efield(3, 0.7)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    assert np.array_equal(ef.e_field(), e)
    _, _, st32 = make_state(dim, ext, precision=lqcd.F32)
    assert np.array_equal(st32.link_matrix(), u.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("dim,ext", SHAPES + [(4, [8, 8, 8, 8])])
def test_plaquette_and_hamiltonians(dim, ext):
    announce(f"plaquette D={dim} ext={ext}")
    beta = 5.7
    ol, u, st = make_state(dim, ext, beta)
    p_ref = ol.average_trace_plaquette(u)
    p = st.average_trace_plaquette()
    assert abs(p - p_ref) <= 1e-13 * abs(p_ref)
    h_ref = ol.hamiltonian_links(u, beta)
    assert abs(st.hamiltonian_links() - h_ref) <= 1e-13 * abs(h_ref)
    assert abs(st.wilson_forward() - h_ref) <= 1e-13 * abs(h_ref)
    e = ol.random_efield(4, 0.5 / beta)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    he_ref = ol.hamiltonian_efield(e, beta)
    assert abs(ef.hamiltonian_efield() - he_ref) <= 1e-13 * abs(he_ref)


@pytest.mark.parametrize("dim,ext", SHAPES + [(4, [8, 8, 8, 8])])
def test_derivative_e_and_backward(dim, ext):
    announce(f"force D={dim} ext={ext}")
    beta = 5.7
    ol, u, st = make_state(dim, ext, beta, size=0.5)
    ref = ol.derivative_e(u)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    got = ef.derivative_e()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale
    grad = st.wilson_backward()
    gref = ol.wilson_backward(u, beta)
    assert np.max(np.abs(grad - gref)) <= 1e-12 * np.max(np.abs(gref))


@pytest.mark.parametrize("dim,ext", [(4, [4, 6, 2, 8]), (3, [6, 6, 6])])
def test_wilson_backward_into_device_buffer(dim, ext):
    # the same kernel and scale as the host entry point: bit for bit
    announce(f"wilson backward into a device buffer D={dim} ext={ext}")
    import ctypes as C
    ol, u, st = make_state(dim, ext, 5.7, size=0.5)
    grad = st.wilson_backward()
    buf = F.DeviceBuffer(grad.nbytes)
    st.wilson_backward_device(buf)
    assert np.array_equal(buf.to_numpy().reshape(grad.shape), grad)
    assert np.max(np.abs(grad - ol.wilson_backward(u, 5.7))) <= 1e-12 * np.max(np.abs(grad))
    lib = F.lib()
    assert lib.lqcd_wilson_backward_device(st.handle, 5.7, C.c_void_p(buf.ptr), grad.shape[0] - 1) \
        == F.ERR_INCOMPATIBLE_SIZE
    assert lib.lqcd_wilson_backward_device(st.handle, 5.7, None, grad.shape[0]) == F.ERR_INVALID_ARGUMENT
    assert lib.lqcd_wilson_backward_device(st.handle, float("nan"), C.c_void_p(buf.ptr), grad.shape[0]) \
        == F.ERR_INVALID_ARGUMENT


def test_cold_state_force_zero_and_invariant():
    announce("cold invariance")
    ol, u, st = make_state(4, [4, 4, 4, 4], cold=True)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    assert np.all(ef.derivative_e() == 0.0)
    assert st.average_trace_plaquette() == 3.0
    for drift in (lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP):
        ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(drift), 0.1, 3)
        assert np.array_equal(ef.link_matrix(), u)
        assert np.all(ef.e_field() == 0.0)


@pytest.mark.parametrize("drift", [lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP])
def test_single_kick_and_drift(drift):
    announce(f"kick/drift drift={drift}")
    ol, u, st = make_state(4, [4, 4, 4, 6], size=0.7)
    e = ol.random_efield(5, 0.4)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    integ = lqcd.SymplecticEulerGpu(drift)
    integ.integrate_efield(ef, 0.013)
    e_ref = ol.integrate_efield(u, e, 0.013)
    assert np.max(np.abs(ef.e_field() - e_ref)) <= 1e-13
    integ.integrate_link(ef, 0.021)
    u_ref = ol.integrate_link(u, e_ref, 0.021, drift)
    assert np.max(np.abs(ef.link_matrix() - u_ref)) <= 1e-13
    if drift == lqcd.DRIFT_EXP:
        # |X| ~ 1.5: the device series is scaled by 2^-m and squared m times;
        # the reference's 25-term series is still exact at this size
        big = ol.random_efield(9, 10.0)
        st.set_link_matrix(u)
        ef.set_e_field(big)
        integ.integrate_link(ef, 0.021)
        u_ref = ol.integrate_link(u, big, 0.021, drift)
        assert np.max(np.abs(ef.link_matrix() - u_ref)) <= 1e-13


@pytest.mark.parametrize("drift", [lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP])
@pytest.mark.parametrize("scheme", [lqcd.SCHEME_REFERENCE, lqcd.SCHEME_LEAPFROG])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_simulate_symplectic_n(drift, scheme, variant):
    announce(f"symplectic_n drift={drift} scheme={scheme} variant={variant}")
    for dim, ext in [(4, [4, 4, 4, 6]), (3, [6, 4, 6]), (2, [8, 6]), (4, [8, 4, 8, 8])]:
        ol, u, st = make_state(dim, ext, beta=6.0)
        st.set_md_variant(variant)
        e = ol.random_efield(6, 0.5 / 6.0)
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(drift, scheme), 0.01, 5)
        u_ref, e_ref = ol.simulate_symplectic_n(u, e, 0.01, 5, drift, scheme)
        assert np.max(np.abs(ef.link_matrix() - u_ref)) <= 1e-12
        assert np.max(np.abs(ef.e_field() - e_ref)) <= 1e-12
        assert st.md_variant() == variant


def test_md_autotune_picks_a_variant():
    announce("md autotune")
    ol, u, st = make_state(4, [8, 8, 8, 8])
    e = ol.random_efield(6, 0.5 / 6.0)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    v = st.md_autotune(lqcd.DRIFT_EXP)
    assert v in (0, 1, 2, 3, 4) and st.md_variant() == v
    # the tuner leaves the state untouched
    assert np.array_equal(ef.e_field(), e)
    assert np.array_equal(st.link_matrix(), u)


def test_zero_integration_error():
    announce("zero integration")
    _, _, st = make_state(3, [4, 4, 4])
    ef = lqcd.LatticeStateEFSyncDefault(st)
    with pytest.raises(lqcd.MultiIntegrationError):
        ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(), 0.01, 0)


def test_incompatible_sizes():
    announce("incompatible sizes")
    lat = lqcd.LatticeCyclic(1.0, 4, 3)
    st = lqcd.LatticeStateDefault(lat, 6.0)
    with pytest.raises(lqcd.StateInitializationError):
        st.set_link_matrix(np.zeros((10, 18)))
    with pytest.raises(lqcd.StateInitializationError):
        lqcd.LatticeStateEFSyncDefault(st, np.zeros((10, 8)))


@pytest.mark.parametrize("drift,scheme,dim,ext", [
    (lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE, 4, [8, 8, 8, 8]),
    (lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 4, [4, 4, 6, 6]),
    (lqcd.DRIFT_EXP, lqcd.SCHEME_REFERENCE, 3, [6, 6, 6]),
])
def test_hmc_trajectory(drift, scheme, dim, ext):
    announce(f"hmc drift={drift} scheme={scheme} D={dim} ext={ext}")
    beta = 6.0
    ol, u, st = make_state(dim, ext, beta)
    e = ol.random_efield(7, 0.5 / beta)
    hmc = lqcd.HybridMonteCarloDiagnostic(0.01, 10, lqcd.SymplecticEulerGpu(drift, scheme), 1)
    for uniform in (0, 2**64 - 1, 2**63):
        st.set_link_matrix(u)
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        got = hmc.trajectory(st, uniform)
        u_ref, ref = ol.hmc_trajectory(u, e, beta, 0.01, 10, uniform, drift, scheme)
        tol = 1e-9 * abs(ref.h_old) + 1e-10
        assert abs(got.h_old - ref.h_old) <= tol
        assert abs(got.delta_h - ref.delta_h) <= tol
        assert got.accepted == ref.accepted
        assert np.max(np.abs(st.link_matrix() - u_ref)) <= 1e-11
        assert hmc.has_replace_last() == bool(ref.accepted)


@pytest.mark.parametrize("dim,ext,n_update", [(4, [4, 4, 4, 4], 1), (3, [6, 4, 8], 2),
                                              (4, [8, 8, 8, 8], 1)])
def test_metropolis_sweep(dim, ext, n_update):
    announce(f"metropolis D={dim} ext={ext} n_update={n_update}")
    beta = 6.0
    ol, u, st = make_state(dim, ext, beta)
    mh = lqcd.MetropolisHastingsSweep(n_update, 0.1)
    rr = O.OracleRng(33)
    for sweep in range(3):
        props = ol.mh_proposals(40 + sweep, 0.1, n_update)
        unis = np.array([rr.next_u64() for _ in range(ol.nlink)], dtype=np.uint64)
        mh.sweep(st, props, unis)
        u, ref = ol.metropolis_sweep(u, beta, props, unis, O.ORDER_CHECKERBOARD)
        assert mh.number_replace_last() == ref.number_replace_last
        assert abs(mh.prob_replace_mean() - ref.prob_replace_mean) <= 1e-12
        assert np.max(np.abs(st.link_matrix() - u)) <= 1e-12


@pytest.mark.parametrize("dim,ext,n_update,precision", [
    (4, [4, 4, 4, 6], 1, lqcd.F64),
    (4, [4, 4, 4, 4], 3, lqcd.F64),
    (3, [6, 4, 8], 2, lqcd.F64),
    (4, [4, 4, 4, 6], 2, lqcd.F32),
])
def test_metropolis_device_stream(dim, ext, n_update, precision):
    announce(f"metropolis device stream D={dim} ext={ext} n_update={n_update} prec={precision}")
    beta, spread, seed = 5.5, 0.2, 0x5EED1234ABCD
    ol, u, st = make_state(dim, ext, beta, precision=precision)
    u = st.link_matrix()  # f32 storage: start the oracle from the stored values
    mh = lqcd.MetropolisHastingsSweep(n_update, spread, seed)
    for sweep in range(3):
        props, unis = ol.device_mh_inputs(seed, sweep, spread, n_update)
        mh.next_element(st)
        u, ref = ol.metropolis_sweep(u, beta, props, unis, O.ORDER_CHECKERBOARD)
        if precision == lqcd.F64:
            assert mh.number_replace_last() == ref.number_replace_last
            assert abs(mh.prob_replace_mean() - ref.prob_replace_mean) <= 1e-12
            assert np.max(np.abs(st.link_matrix() - u)) <= 1e-12
        else:
            # accepted links are stored f32; the oracle continues from them
            assert abs(mh.number_replace_last() - ref.number_replace_last) <= 2
            assert abs(mh.prob_replace_mean() - ref.prob_replace_mean) <= 1e-5
            got = st.link_matrix()
            same = np.max(np.abs(got - u), axis=1) <= 1e-6
            assert same.mean() >= 0.999
            u = got
    assert mh.counter() == 3
    assert 0 < mh.number_replace_last() < ol.nlink


def test_metropolis_device_stream_rejects_bad_arguments():
    announce("metropolis device stream arguments")
    _, _, st = make_state(4, [4, 4, 4, 4])
    sweep = F.lib().lqcd_metropolis_sweep_random
    assert sweep(st.handle, 6.0, 4001, 0.1, 1, 0, None) == F.ERR_INVALID_ARGUMENT
    assert sweep(st.handle, 6.0, 1, 1.5, 1, 0, None) == F.ERR_INVALID_ARGUMENT
    assert sweep(st.handle, 6.0, 0, 0.1, 1, 0, None) == F.ERR_INVALID_ARGUMENT
    assert sweep(st.handle, 6.0, 1, 0.1, 1, 0, None) == F.OK
    with pytest.raises(ValueError):
        lqcd.MetropolisHastingsSweep(1, 1.0, 3)


def test_metropolis_rejects_odd_extent():
    announce("metropolis odd extent")
    lat = lqcd.LatticeCyclic(1.0, 3, 3)
    st = lqcd.LatticeStateDefault(lat, 6.0)
    mh = lqcd.MetropolisHastingsSweep(1, 0.1)
    with pytest.raises(lqcd.LqcdError):
        mh.sweep(st, np.zeros((1, 81, 18)), np.zeros(81, dtype=np.uint64))


def test_normalize_link_matrices():
    announce("reunitarize")
    ol, u, st = make_state(4, [4, 4, 4, 4])
    noisy = u + 1e-3 * np.random.default_rng(0).normal(size=u.shape)
    st.set_link_matrix(noisy)
    st.normalize_link_matrices()
    ref = ol.normalize_link_matrices(noisy)
    assert np.max(np.abs(st.link_matrix() - ref)) <= 1e-14
    # f32 storage: reunitarised in f64, stored f32
    _, _, st32 = make_state(4, [4, 4, 4, 4], precision=lqcd.F32)
    st32.set_link_matrix(noisy)
    st32.normalize_link_matrices()
    assert np.max(np.abs(st32.link_matrix() - ref)) <= 1e-7


def test_device_hot_start_match_oracle_stream():
    announce("device hot start")
    for dim, ext in [(4, [4, 4, 4, 6]), (3, [6, 6, 6])]:
        ol = O.OracleLattice(dim, ext)
        st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, ext, 0x1234, dim)
        ref = ol.device_hot_start(0x1234)
        assert np.max(np.abs(st.link_matrix() - ref)) <= 1e-14


def test_device_momenta_match_oracle_stream():
    announce("device momenta")
    ol, u, st = make_state(4, [4, 4, 4, 6])
    ef = lqcd.LatticeStateEFSyncDefault(st)
    ef.reset_e_field(0xABCDEF, 17, 0.25)
    ref = ol.device_momenta(0xABCDEF, 17, 0.25)
    assert np.max(np.abs(ef.e_field() - ref)) <= 1e-14


def test_f32_path_errors_vs_f64_oracle():
    announce("f32 path")
    beta = 6.0
    ol, u, st = make_state(4, [8, 8, 8, 8], beta, precision=lqcd.F32)
    u32 = u.astype(np.float32).astype(np.float64)
    # plaquette of the stored (f32-rounded) links in f64 arithmetic
    p_ref = ol.average_trace_plaquette(u32)
    assert abs(st.average_trace_plaquette() - p_ref) <= 1e-13 * abs(p_ref)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    f_ref = ol.derivative_e(u32)
    assert np.max(np.abs(ef.derivative_e() - f_ref)) <= 2e-6 * np.max(np.abs(f_ref))
    e = ol.random_efield(8, 0.5 / beta)
    for drift in (lqcd.DRIFT_ADDITIVE, lqcd.DRIFT_EXP):
        st.set_link_matrix(u)
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(drift, lqcd.SCHEME_LEAPFROG), 0.01, 10)
        u_ref, e_ref = ol.simulate_symplectic_n(u32, e.astype(np.float32).astype(np.float64),
                                                0.01, 10, drift, O.SCHEME_LEAPFROG)
        assert np.max(np.abs(ef.link_matrix() - u_ref)) <= 1e-5
        assert np.max(np.abs(ef.e_field() - e_ref)) <= 1e-5


def _perturbed_action(st, ef, u, omega, eps):
    """S_W(exp(i eps Omega) U) on the device: Omega as E, one exp drift of
    dt = eps a / sqrt(6) (U' = exp(i dt sqrt(6)/a E) U)."""
    st.set_link_matrix(u)
    ef.set_e_field(omega)
    lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP).integrate_link(ef, eps * st.lattice().size() / math.sqrt(6))
    return st.wilson_forward()


@pytest.mark.parametrize("precision", [lqcd.F64])
def test_wilson_adjoint_finite_differences_16x4(precision):
    # adjoint check: D=4 16^4, hot start, beta=5.7; backward vs central finite
    # differences of the forward (eps=1e-4, 64 random directions, Gaussian
    # coefficients) and vs the oracle's derivative_e
    announce("adjoint 16^4 finite differences")
    beta = 5.7
    ol = O.OracleLattice(4, 16)
    u = ol.device_hot_start(0x5EED)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 16, 4), beta, u, precision)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    grad = st.wilson_backward()
    # backward == -(beta a / sqrt 6) derivative_e of the oracle
    ref = -(beta / math.sqrt(6)) * ol.derivative_e(u)
    assert np.max(np.abs(grad - ref)) <= 1e-12 * np.max(np.abs(ref))
    rng = np.random.default_rng(64)
    worst = 0.0
    for k in range(64):
        omega = rng.normal(size=(ol.nlink, 8))
        pred = float(np.sum(omega * grad))
        fd = (_perturbed_action(st, ef, u, omega, 1e-4)
              - _perturbed_action(st, ef, u, omega, -1e-4)) / 2e-4
        worst = max(worst, abs(fd - pred) / abs(pred))
        if k < 4:
            # Richardson extrapolation removes the eps^2 term
            f1 = (_perturbed_action(st, ef, u, omega, 1e-3)
                  - _perturbed_action(st, ef, u, omega, -1e-3)) / 2e-3
            f2 = (_perturbed_action(st, ef, u, omega, 2e-3)
                  - _perturbed_action(st, ef, u, omega, -2e-3)) / 4e-3
            assert abs((4 * f1 - f2) / 3 - pred) <= 1e-7 * abs(pred)
    print(f"adjoint: worst relative |fd - <Omega, grad>| over 64 directions = {worst:.2e}")
    assert worst <= 1e-5


@pytest.mark.parametrize("dim,ext,cold", [(4, [4, 4, 4, 4], False), (3, [4, 6, 4], False),
                                          (4, [4, 4, 2, 6], True)])
def test_gauss_law_and_projection(dim, ext, cold):
    # EField::gauss / gauss_sum_div / project_to_gauss (src/field.rs:1172-1337)
    announce(f"gauss law D={dim} ext={ext} cold={cold}")
    ol, u, st = make_state(dim, ext, cold=cold)
    e = ol.random_efield(5, 0.5 / 6.0)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    g = ef.gauss()
    ref = np.array([ol.gauss(u, e, s) for s in range(ol.nsite)])
    assert np.max(np.abs(g - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    div_ref = ol.gauss_sum_div(u, e)
    assert abs(ef.gauss_sum_div() - div_ref) <= 1e-12 * div_ref
    it = ef.project_to_gauss(2000)
    rc, e2, it_ref = ol.project_to_gauss(u, e, 2000)
    assert rc == 0
    print(f"projection: {it} checks on the device, {it_ref} in the oracle")
    assert abs(it - it_ref) <= 1
    assert np.max(np.abs(ef.e_field() - e2)) <= 1e-10
    assert ef.gauss_sum_div() <= np.finfo(float).eps * ol.nsite * 320


def test_gauss_projection_f32_storage_and_limits():
    announce("gauss projection f32 / limits")
    ol, u, _ = make_state(4, [4, 4, 4, 4])
    e = ol.random_efield(9, 0.5 / 6.0)
    rc, e2, _ = ol.project_to_gauss(u, e, 2000)
    assert rc == 0
    st32 = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, 4, 4), 6.0, u, lqcd.F32)
    ef32 = lqcd.LatticeStateEFSyncDefault(st32, e)
    ef32.project_to_gauss(2000)
    # projected in double from the f32 links, stored rounded to f32
    assert np.max(np.abs(ef32.e_field() - e2)) <= 2e-6
    # not converged within the limit: error, E left as it was
    _, _, st = make_state(4, [4, 4, 4, 4])
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    with pytest.raises(lqcd.StateInitializationError):
        ef.project_to_gauss(1)
    assert np.array_equal(ef.e_field(), e)
    bad = e.copy()
    bad[3, 2] = np.nan
    ef.set_e_field(bad)
    with pytest.raises(lqcd.StateInitializationError):
        ef.project_to_gauss(50)
    # the HMC refresh projects by default (new_random_e_state)
    hmc = lqcd.HybridMonteCarloDiagnostic(0.01, 3, lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP), 7)
    hmc.next_element(st)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 11, 0)
    assert ef.gauss_sum_div() <= np.finfo(float).eps * ol.nsite * 320


@pytest.mark.parametrize("dim,ext,precision", [(4, [4, 4, 4, 4], lqcd.F64), (3, [4, 6, 4], lqcd.F64),
                                               (4, [4, 4, 4, 6], lqcd.F32)])
def test_heat_bath_sweep_device_stream(dim, ext, precision):
    # HeatBathSweep (heat_bath.rs:73-123), checkerboard order, device stream
    announce(f"heat bath D={dim} ext={ext} prec={precision}")
    beta, seed = 5.5, 0xB0A7
    ol, u, st = make_state(dim, ext, beta, precision=precision)
    u = st.link_matrix()
    hb = lqcd.HeatBathSweep(seed)
    for sweep in range(2):
        st.monte_carlo_step(hb)
        u, ex = ol.heat_bath_sweep(u, beta, seed, sweep)
        assert hb.exhausted_last() == ex == 0
        got = st.link_matrix()
        if precision == lqcd.F64:
            assert np.max(np.abs(got - u)) <= 1e-10
        else:
            # f64 arithmetic, f32 storage; the oracle continues from the stored links
            assert np.max(np.abs(got - u)) <= 1e-4 * max(1.0, np.max(np.abs(u)))
            u = got
    assert hb.counter() == 2


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 4]), (3, [6, 4, 4])])
def test_overrelaxation_sweep(dim, ext):
    # OverrelaxationSweepRotation (overrelaxation.rs:66-91), checkerboard order
    announce(f"overrelaxation D={dim} ext={ext}")
    beta = 1.0
    ol, u, st = make_state(dim, ext, beta)
    h = st.hamiltonian_links()
    st.monte_carlo_step(lqcd.OverrelaxationSweepRotation())
    ref = ol.overrelaxation_sweep(u, O.ORDER_CHECKERBOARD)
    assert np.max(np.abs(st.link_matrix() - ref)) <= 1e-11
    h2 = st.hamiltonian_links()
    assert abs(h - h2) < np.finfo(float).eps * 100 * ol.nsite * (h + h2) * 0.5
    # composed as the reference's HybridMethodVec
    mix = lqcd.HybridMethodVec([lqcd.HeatBathSweep(3), lqcd.OverrelaxationSweepRotation()])
    st.monte_carlo_step(mix)
    with pytest.raises(lqcd.LqcdError):
        _, _, odd = make_state(3, [3, 4, 4])
        odd.monte_carlo_step(lqcd.OverrelaxationSweepRotation())


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 4]), (3, [6, 4, 4])])
def test_overrelaxation_reverse_sweep(dim, ext):
    # OverrelaxationSweepReverse (overrelaxation.rs:136-183), checkerboard order,
    # device Jacobi SVD against the oracle's
    announce(f"overrelaxation reverse D={dim} ext={ext}")
    beta = 1.0
    ol, u, st = make_state(dim, ext, beta)
    h = st.hamiltonian_links()
    for _ in range(2):
        st.monte_carlo_step(lqcd.OverrelaxationSweepReverse())
        u = ol.overrelaxation_reverse_sweep(u, O.ORDER_CHECKERBOARD)
        assert np.max(np.abs(st.link_matrix() - u)) <= 1e-11
        u = st.link_matrix()
    h2 = st.hamiltonian_links()
    assert abs(h - h2) < np.finfo(float).eps * 100 * ol.nsite * (h + h2) * 0.5
    mix = lqcd.HybridMethodVec([lqcd.HeatBathSweep(5), lqcd.OverrelaxationSweepReverse()])
    st.monte_carlo_step(mix)
    assert np.all(np.isfinite(st.link_matrix()))
    with pytest.raises(lqcd.LqcdError):
        _, _, odd = make_state(3, [3, 4, 4])
        odd.monte_carlo_step(lqcd.OverrelaxationSweepReverse())


def test_hybrid_couple_and_plain_hmc():
    # HybridMethodCouple (hybrid.rs:327-420) applies its two methods in order:
    # the same as HybridMethodVec of both, and as the two steps by hand;
    # HybridMonteCarlo (hybrid_monte_carlo.rs:64-200) moves the state as
    # HybridMonteCarloDiagnostic with the same seed
    announce("HybridMethodCouple, HybridMonteCarlo")
    states = [make_state(4, [4, 4, 4, 4], 5.5)[2] for _ in range(3)]
    states[0].monte_carlo_step(lqcd.HybridMethodCouple(lqcd.HeatBathSweep(9), lqcd.OverrelaxationSweepRotation()))
    states[1].monte_carlo_step(lqcd.HybridMethodVec([lqcd.HeatBathSweep(9), lqcd.OverrelaxationSweepRotation()]))
    states[2].monte_carlo_step(lqcd.HeatBathSweep(9)).monte_carlo_step(lqcd.OverrelaxationSweepRotation())
    assert np.array_equal(states[0].link_matrix(), states[1].link_matrix())
    assert np.array_equal(states[0].link_matrix(), states[2].link_matrix())
    a, b = (make_state(4, [4, 4, 4, 4], 6.0)[2] for _ in range(2))
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc_a = lqcd.HybridMonteCarlo(0.02, 10, integ, 13)
    hmc_b = lqcd.HybridMonteCarloDiagnostic(0.02, 10, integ, 13)
    for _ in range(3):
        a.monte_carlo_step(hmc_a)
        b.monte_carlo_step(hmc_b)
    assert np.array_equal(a.link_matrix(), b.link_matrix())


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 6]), (3, [6, 4, 4])])
def test_metropolis_delta_diagnostic(dim, ext):
    # MetropolisHastingsDeltaDiagnostic (metropolis_hastings.rs:374-419), one
    # link per step from the device stream: the oracle restates the link draw
    # (Philox block 62 of link 0) and takes the proposal and accept word of
    # that link from its restatement of the sweep's stream; dS from its staple
    announce(f"MetropolisHastingsDeltaDiagnostic D={dim} ext={ext}")
    beta, seed, spread = 5.0, 0x1234_5678_9ABC, 0.3
    ol, u, st = make_state(dim, ext, beta)
    mh = lqcd.MetropolisHastingsDeltaDiagnostic(spread, seed)
    accepted = 0
    for step in range(40):
        st.monte_carlo_step(mh)
        o = O.philox4x32([0, 62 << 16, step & 0xFFFFFFFF, step >> 32], [seed & 0xFFFFFFFF, seed >> 32])
        site = (((int(o[0]) << 32) | int(o[1])) * ol.nsite) >> 64
        mu = (int(o[2]) * dim) >> 32
        li = site * dim + mu
        assert mh.last_link == li
        props, words = ol.device_mh_inputs(seed, step, spread, 1)
        new = O.from_np(O.to_np(props[0, li]) @ O.to_np(u[li]))
        p = min(max(math.exp(-ol.delta_s(u, site, mu, new, beta)), 0.0), 1.0)
        assert abs(mh.prob_replace_last() - p) <= 1e-12 * max(p, 1e-300) + 1e-15
        acc = O.bernoulli_accept(p, int(words[li]))
        assert mh.has_replace_last() == acc
        if acc:
            u = u.copy()
            u[li] = new
            accepted += 1
    assert 0 < accepted < 40
    assert np.max(np.abs(st.link_matrix() - u)) <= 1e-12
    with pytest.raises(ValueError):
        lqcd.MetropolisHastingsDeltaDiagnostic(1.0)


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 6]), (3, [6, 4, 6])])
def test_derivative_u(dim, ext):
    # LatticeStateWithEField::derivative_u (state.rs:1407-1417) of every link
    # against the oracle's restatement, and the additive drift it defines
    # (integrate_link, integrator/mod.rs:216-233): U' = U + dt derivative_u
    announce(f"derivative_u D={dim} ext={ext}")
    ol, u, st = make_state(dim, ext, size=0.8)
    e = ol.random_efield(17, 0.4)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    du = ef.derivative_u()
    ref = np.array([ol.derivative_u(u, e, s, d) for s in range(ol.nsite) for d in range(dim)])
    assert np.max(np.abs(du - ref)) <= 1e-13
    lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE).integrate_link(ef, 0.03)
    assert np.max(np.abs(ef.link_matrix() - (u + 0.03 * du))) <= 1e-13


def test_efsync_constructors():
    # LatticeStateEFSyncDefault's constructors (state.rs:262-277, 1175-1231)
    announce("EFSync constructors")
    a = lqcd.LatticeStateEFSyncDefault.new_cold(1.0, 6.0, 4, 3)
    eye = O.from_np(np.eye(3, dtype=complex))
    assert np.all(a.e_field() == 0) and np.all(a.link_matrix() == eye)
    b = lqcd.LatticeStateEFSyncDefault.new_determinist_cold_e_hot_link(1.0, 6.0, 4, 5, 3)
    ref = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 4, 5, 3)
    assert np.all(b.e_field() == 0) and np.array_equal(b.link_matrix(), ref.link_matrix())
    # the reference's doctest (state.rs:1165-1173): the same seeds give the same state
    states = []
    for _ in range(2):
        rng = np.random.default_rng(0)
        states.append(lqcd.LatticeStateEFSyncDefault.new_determinist(
            1.0, 1.0, 4, 0, lambda shape: rng.uniform(-1.0, 1.0, shape), 4))
    assert np.array_equal(states[0].e_field(), states[1].e_field())
    assert np.array_equal(states[0].link_matrix(), states[1].link_matrix())
    assert np.all(np.abs(states[0].e_field()) < 1.0) and np.any(states[0].e_field() != 0)
    lat = lqcd.LatticeCyclic(1.0, 4, 3)
    u = ref.link_matrix()
    c = lqcd.LatticeStateEFSyncDefault.new_random_e(lat, 6.0, u, 7)
    assert np.array_equal(c.link_matrix(), u) and np.any(c.e_field() != 0)
    # projected: below project_to_gauss's stopping threshold N * 4 * 8 * 10 * eps
    assert c.gauss_sum_div() <= 4 ** 3 * 4 * 8 * 10 * 2.220446049250313e-16
    with pytest.raises(lqcd.StateInitializationError):
        lqcd.LatticeStateEFSyncDefault.new_random_e(lat, 0.0, u, 7)
