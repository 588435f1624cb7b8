"""H_old and H_new of lqcd_hmc_trajectory summed inside the MD kernels.

With a per-site work decomposition (MD variants 2-4) the first fused launch
of a trajectory sums sum Re Tr(U Sigma) over the starting links and sum e^2
over the starting momenta, and the final kick the same over the final links
and momenta; every plaquette enters sum Re Tr(U Sigma) through its 4 links,
so H_U = beta (N_plaquettes - sum Re Tr(U Sigma) / 12) and H_E = beta sum
e^2 / 2 (state.rs:821-849, 1370-1385).  Checked against the plaquette
kernel's H of the same states (hamiltonian_total before the trajectory, and
after it on the accepted links) and against the other decompositions, which
keep the separate H kernels: relative 1e-13 in f64; in f32 the traces are
taken in the kernels' f32 arithmetic (summed in f64), relative 1e-8.
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


@pytest.mark.parametrize("precision,drift,scheme", [
    (lqcd.F64, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
    (lqcd.F64, lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE),
    (lqcd.F32, lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
    (lqcd.F32, lqcd.DRIFT_EXP, lqcd.SCHEME_REFERENCE),
])
@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_fused_hamiltonian_sums(precision, drift, scheme, variant):
    announce(f"fused H sums precision={precision} drift={drift} scheme={scheme} variant={variant}")
    ext = [16, 16, 4, 4]
    beta = 5.8
    ol = O.OracleLattice(4, ext)
    u = ol.hot_start(31)
    e = ol.random_efield(32, 0.5 / beta)
    if precision == lqcd.F32:
        u = u.astype(np.float32).astype(np.float64)
        e = e.astype(np.float32).astype(np.float64)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, ext, 4), beta, u, precision)
    st.set_md_variant(variant)
    ef = lqcd.LatticeStateEFSyncDefault(st, e)
    h_before = ef.hamiltonian_total()
    hmc = lqcd.HybridMonteCarloDiagnostic(0.01, 6, lqcd.SymplecticEulerGpu(drift, scheme), 1,
                                          project_gauss=False)
    s = hmc.trajectory(st, 0)    # accept word 0: accepted unless p == 0
    assert s.accepted == 1
    h_after = ef.hamiltonian_total()
    rel = 1e-13 if precision == lqcd.F64 else 1e-8
    print(f"h_old {s.h_old:.12e} plaquette kernel {h_before:.12e}; "
          f"h_new {s.h_new:.12e} plaquette kernel {h_after:.12e}")
    assert abs(s.h_old - h_before) <= rel * abs(h_before)
    assert abs(s.h_new - h_after) <= rel * abs(h_after)
    if drift == lqcd.DRIFT_EXP:
        # the links left behind are complete: row 2 matches rows 0, 1
        assert st.su3_distance() <= (64 * 2.0 ** -24 if precision == lqcd.F32 else 1e-13)
