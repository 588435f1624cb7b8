"""GPU parity of the batched chains (lqcd_batch_*): every chain must evolve as
an independent lattice driven with the same draws (single-lattice path, itself
pinned against the oracle), graph replays must equal eager launches, and one
chain is checked against the oracle's HMC trajectory directly
(hybrid_monte_carlo.rs:584-613)."""
import sys

import numpy as np
import pytest

import oracle as O
import lqcd
from lqcd import _ffi as F

pytestmark = pytest.mark.gpu

ACCEPT_BLOCK = 62


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def accept_word(seed, n):
    """Accept word of trajectory n of a chain: Philox4x32-10 counter
    {0, 62 << 16, lo32(n), hi32(n)}, key seed, (x << 32) | y."""
    o = O.philox4x32([0, ACCEPT_BLOCK << 16, n & 0xFFFFFFFF, n >> 32],
                     [seed & 0xFFFFFFFF, seed >> 32])
    return (int(o[0]) << 32) | int(o[1])


def uneven_chains(d=4):
    if d == 4:
        ext = [[4, 4, 4, 4], [6, 4, 4, 8], [8, 8, 8, 8], [4, 6, 6, 4], [6, 6, 6, 6]]
    else:
        ext = [[6, 6, 6], [4, 8, 6], [10, 10, 10]]
    betas = np.linspace(5.6, 6.2, len(ext))
    return [lqcd.ChainSpec(e, float(b), 0x1000 + 77 * i) for i, (e, b) in enumerate(zip(ext, betas))]


def single_run(spec, d, precision, variant, dt, n_steps, n_traj, drift, scheme, reunit):
    """The same chain as a lattice of its own, driven with the batch's draws."""
    st = lqcd.LatticeStateDefault.new_determinist(spec.size, spec.beta, spec.extents, spec.seed, d,
                                                  precision)
    st.set_md_variant(variant)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, n_steps, lqcd.SymplecticEulerGpu(drift, scheme), 1)
    stats = []
    for n in range(n_traj):
        ef.reset_e_field(spec.seed, n, 0.5 / spec.beta)
        stats.append(hmc.trajectory(st, accept_word(spec.seed, n)))
        if reunit:
            st.normalize_link_matrices()
    return st.link_matrix(), stats


def test_batch_hot_start_matches_single_lattices():
    announce("batch hot start")
    chains = uneven_chains()
    b = lqcd.BatchedChains(chains, 4, lqcd.F64)
    b.hot_start()
    for i, ch in enumerate(chains):
        st = lqcd.LatticeStateDefault.new_determinist(1.0, ch.beta, ch.extents, ch.seed, 4)
        assert np.array_equal(b.link_matrix(i), st.link_matrix())
    ol = O.OracleLattice(4, chains[1].extents)
    assert np.max(np.abs(b.link_matrix(1) - ol.device_hot_start(chains[1].seed))) <= 1e-14
    # cold start: plaquette exactly 1
    b.cold_start()
    assert np.all(b.plaquette() == 1.0)


@pytest.mark.parametrize("precision,config,drift,scheme,d", [
    (lqcd.F64, 5, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 4),    # per site, 256, full grid
    (lqcd.F64, 6, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 4),    # per link, 64, persistent
    (lqcd.F64, 10, lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE, 4),  # per site, 128, persistent
    (lqcd.F32, 11, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 4),   # per site, 256, persistent
    (lqcd.F32, 1, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 4),    # per link, 128, full grid
    (lqcd.F64, 4, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, 3),    # D = 3
])
def test_batch_hmc_matches_single_lattices(precision, config, drift, scheme, d):
    announce(f"batch hmc prec={precision} config={config} drift={drift} scheme={scheme} D={d}")
    dt, n_steps, n_traj = 0.02, 5, 3
    chains = uneven_chains(d)
    b = lqcd.BatchedChains(chains, d, precision)
    b.hot_start()
    b.config = config
    site = (config // 3) % 2 == 1
    stats = b.hmc(dt, n_steps, n_traj, drift, scheme, reunitarize=True)
    assert b.counter == n_traj
    # exp-drift chains of SU(3) links read the links as two rows, as a
    # lattice of its own does
    assert b.md_reconstruct_last() == (drift == lqcd.DRIFT_EXP)
    tol_u = 1e-12 if precision == lqcd.F64 else 2e-5
    for i, ch in enumerate(chains):
        u_ref, ref = single_run(ch, d, precision, 2 if site else 0, dt, n_steps, n_traj, drift,
                                scheme, True)
        got = b.link_matrix(i)
        err = np.max(np.abs(got - u_ref))
        print(f"chain {i} {ch.extents}: max |U_batch - U_single| = {err:.2e}")
        assert err <= tol_u
        for t in range(n_traj):
            s = stats[t, i]
            r = ref[t]
            # a lattice of its own sums H inside its MD kernels (f32 traces
            # for f32 storage), the batch with its plaquette kernel
            tol_h = (1e-12 if precision == lqcd.F64 else 1e-8) * abs(r.h_old) + 1e-12
            assert abs(s["h_old"] - r.h_old) <= tol_h
            assert abs(s["h_new"] - r.h_new) <= tol_h
            assert s["accepted"] == r.accepted
            # p = exp(-dH) moves with dH: 1e-4 covers the f32 traces' 2e-5 in dH
            assert abs(s["probability"] - r.probability) <= (1e-9 if precision == lqcd.F64 else 1e-4)
    assert np.all(np.isfinite(stats["delta_h"]))


def test_batch_graph_replays_equal_eager_launches():
    announce("batch graph replay vs eager")
    chains = uneven_chains()
    runs = {}
    for graph in (False, True):
        b = lqcd.BatchedChains(chains, 4, lqcd.F32)
        b.config = 11
        b.set_graph(graph)
        b.hot_start()
        per_traj = []
        for t in range(4):
            # one replay per call, with eager work and synchronisation between replays
            per_traj.append(b.hmc(0.02, 6, 1, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG, reunitarize=(t % 2 == 1)))
            per_traj.append(b.plaquette())
        runs[graph] = (per_traj, [b.link_matrix(i) for i in range(len(chains))])
    eager, graphed = runs[False], runs[True]
    for a, g in zip(eager[0], graphed[0]):
        assert np.array_equal(a, g)
    for a, g in zip(eager[1], graphed[1]):
        assert np.array_equal(a, g)


def test_batch_all_launch_shapes_agree():
    announce("batch launch shapes")
    chains = uneven_chains()
    results = []
    for config in range(12):
        b = lqcd.BatchedChains(chains, 4, lqcd.F64)
        b.hot_start()
        b.config = config
        st = b.hmc(0.02, 4, 1, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
        results.append((st, [b.link_matrix(i) for i in range(len(chains))]))
    st0, u0 = results[0]
    for st, u in results[1:]:
        assert np.array_equal(st["accepted"], st0["accepted"])
        assert np.max(np.abs(st["h_new"] - st0["h_new"]) / np.abs(st0["h_new"])) <= 1e-13
        for a, c in zip(u, u0):
            assert np.max(np.abs(a - c)) <= 1e-12
    b = lqcd.BatchedChains(chains, 4, lqcd.F32)
    b.hot_start()
    assert 0 <= b.autotune() < 12


def test_batch_chain_matches_oracle_trajectory():
    announce("batch chain vs oracle")
    spec = lqcd.ChainSpec([4, 4, 4, 6], 5.8, 0xBEEF)
    other = lqcd.ChainSpec([6, 6, 4, 4], 6.1, 0xCAFE)
    b = lqcd.BatchedChains([other, spec], 4, lqcd.F64)
    b.hot_start()
    b.config = 4
    ol = O.OracleLattice(4, spec.extents)
    u = ol.device_hot_start(spec.seed)
    for n in range(2):
        st = b.hmc(0.01, 8, 1, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)[0, 1]
        e = ol.device_momenta(spec.seed, n, 0.5 / spec.beta)
        u, ref = ol.hmc_trajectory(u, e, spec.beta, 0.01, 8, accept_word(spec.seed, n),
                                   O.DRIFT_EXP, O.SCHEME_LEAPFROG)
        tol = 1e-9 * abs(ref.h_old) + 1e-10
        assert abs(st["h_old"] - ref.h_old) <= tol
        assert abs(st["delta_h"] - ref.delta_h) <= tol
        assert st["accepted"] == ref.accepted
        assert np.max(np.abs(b.link_matrix(1) - u)) <= 1e-11


def test_batch_rejects_bad_arguments():
    announce("batch arguments")
    ok = lqcd.ChainSpec([4, 4, 4, 4], 6.0, 1)
    with pytest.raises(lqcd.LqcdError):
        lqcd.BatchedChains([], 4)
    with pytest.raises(lqcd.LatticeInitializationError):
        lqcd.BatchedChains([ok, lqcd.ChainSpec([4, 1, 4, 4], 6.0, 2)], 4)
    with pytest.raises(lqcd.LatticeInitializationError):
        lqcd.BatchedChains([lqcd.ChainSpec([4, 4, 4, 4], 6.0, 2, size=0.0)], 4)
    with pytest.raises(lqcd.LqcdError):
        lqcd.BatchedChains([lqcd.ChainSpec([4, 4, 4, 4], float("nan"), 2)], 4)
    b = lqcd.BatchedChains([ok], 4)
    with pytest.raises(lqcd.LqcdError):
        b.link_matrix(1)
    with pytest.raises(lqcd.StateInitializationError):
        b.set_link_matrix(0, np.zeros((10, 18)))
    with pytest.raises(lqcd.MultiIntegrationError):
        b.hmc(0.01, 0, 1)
    with pytest.raises(lqcd.LqcdError):
        b.config = 12
    assert b.hmc(0.01, 3, 0).shape == (0, 1)
