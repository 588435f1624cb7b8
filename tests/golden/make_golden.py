"""Writes the golden fixtures of this directory from the CPU oracle (oracle/oracle.c).

The reference crate ships no data files: its tests assert invariants, closed forms
and known answers (src/test/mod.rs, src/field.rs:1446-1718).  These fixtures freeze
the oracle's f64 results on two small hypercubic lattices -- as the reference's
``LatticeCyclic`` is -- with every input stored beside them, so that a test can
check the device path (or a later oracle) against them without regenerating
anything:

* ``d4_l3.npz``: D = 4, 3^4 sites, lattice spacing 0.5, beta 6.0 -- plaquette,
  H_U, H_E, derivative_e of every link (src/simulation/state.rs:821-849,
  1370-1385, 1420-1448), Gauss-law G of every site and its total deviation
  (src/field.rs:1174-1220), five steps of the reference integrator (additive
  drift, kick-drift-kick, src/integrator/symplectic_euler_rayon.rs:222-252) and
  of the exp-drift leapfrog, and one HMC trajectory's H_old, H_new, dH and
  accept decision (src/simulation/monte_carlo/hybrid_monte_carlo.rs:584-613);
* ``d3_l4.npz``: D = 3, 4^3 sites, beta 5.5 -- one checkerboard Metropolis sweep
  fed stored proposals and accept words (metropolis_hastings_sweep.rs:146-174),
  one heat-bath sweep of the device's Philox stream (heat_bath.rs:73-123), one
  overrelaxation sweep of each kind (overrelaxation.rs:66-91, 136-183), the
  reunitarisation of perturbed links (src/su3.rs:279-309), and F_mu_nu and the
  magnetic field of every site (src/field.rs:822-868).

Arrays use the reference's layouts: links (nsite * D, 18) column-major 3x3
complex, link index site * D + mu; E (nsite * D, 8); fields [site][i][j][18].

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

D4 = {"dim": 4, "L": 3, "size": 0.5, "beta": 6.0, "dt": 0.01, "n_md": 5, "n_hmc": 10, "e_seed": 7,
      "uniform": 1 << 63}
D3 = {"dim": 3, "L": 4, "size": 1.0, "beta": 5.5, "spread": 0.1, "mh_seed": 40, "uni_seed": 33,
      "hb_seed": 0xB0A7, "noise_seed": 5, "noise": 1e-3}


def d4_l3():
    c = D4
    ol = O.OracleLattice(c["dim"], c["L"], c["size"])
    u = ol.hot_start(O.SEED_RNG)
    e = ol.random_efield(c["e_seed"], 0.5 / c["beta"])
    out = {"u": u, "e": e}
    p = ol.average_trace_plaquette(u)
    out["plaquette"] = np.array([p.real, p.imag])
    out["h_links"] = np.array(ol.hamiltonian_links(u, c["beta"]))
    out["h_efield"] = np.array(ol.hamiltonian_efield(e, c["beta"]))
    out["derivative_e"] = ol.derivative_e(u)
    out["gauss"] = np.stack([ol.gauss(u, e, s) for s in range(ol.nsite)])
    out["gauss_sum_div"] = np.array(ol.gauss_sum_div(u, e))
    out["u_reference"], out["e_reference"] = ol.simulate_symplectic_n(
        u, e, c["dt"], c["n_md"], O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    out["u_leapfrog"], out["e_leapfrog"] = ol.simulate_symplectic_n(
        u, e, c["dt"], c["n_md"], O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    u_hmc, st = ol.hmc_trajectory(u, e, c["beta"], c["dt"], c["n_hmc"], c["uniform"], O.DRIFT_EXP,
                                  O.SCHEME_LEAPFROG)
    out["u_hmc"] = u_hmc
    out["hmc"] = np.array([st.h_old, st.h_new, st.delta_h, st.probability, float(st.accepted)])
    return out


def d3_l4():
    c = D3
    ol = O.OracleLattice(c["dim"], c["L"], c["size"])
    u = ol.hot_start(O.SEED_RNG)
    out = {"u": u}
    props = ol.mh_proposals(c["mh_seed"], c["spread"], 1)
    rr = O.OracleRng(c["uni_seed"])
    unis = np.array([rr.next_u64() for _ in range(ol.nlink)], dtype=np.uint64)
    out["mh_proposals"], out["mh_uniforms"] = props, unis
    out["u_metropolis"], st = ol.metropolis_sweep(u, c["beta"], props, unis, O.ORDER_CHECKERBOARD)
    out["metropolis"] = np.array([st.prob_replace_mean, float(st.number_replace_last)])
    out["u_heat_bath"], ex = ol.heat_bath_sweep(u, c["beta"], c["hb_seed"], 0)
    assert ex == 0
    out["u_overrelax"] = ol.overrelaxation_sweep(u, O.ORDER_CHECKERBOARD)
    out["u_overrelax_reverse"] = ol.overrelaxation_reverse_sweep(u, O.ORDER_CHECKERBOARD)
    noisy = u + c["noise"] * np.random.default_rng(c["noise_seed"]).normal(size=u.shape)
    out["u_noisy"] = noisy
    out["u_normalized"] = ol.normalize_link_matrices(noisy)
    out["f_mu_nu"] = ol.f_mu_nu_field(u)
    out["magnetic_field"] = ol.magnetic_field_vec(u)
    return out


def write(name, arrays, config):
    meta = {f"cfg_{k}": np.array(v) for k, v in config.items()}
    np.savez_compressed(os.path.join(HERE, name), **arrays, **meta)


def main():
    write("d4_l3.npz", d4_l3(), D4)
    write("d3_l4.npz", d3_l4(), D3)


if __name__ == "__main__":
    main()
