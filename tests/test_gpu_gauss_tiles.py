"""GPU parity of the streamed-tile Gauss-law kernels (D = 4 lattices whose x_0
extent divides 64): EField::gauss, gauss_sum_div and project_to_gauss
(src/field.rs:1172-1337) against the oracle, and the streamed kernels against
the site-walk kernels of the same library (LQCD_GAUSS_TILES=0) at 32^4.

Tolerances: G relative 1e-13, div relative 1e-12, projected E 1e-10 and the
check count within one of the oracle's (as for the site walk); streamed vs
site walk on the device: G within 1e-14 relative, E within 1e-12 and the same
check count (both evaluate G and the step with the same arithmetic; only the
order of the deviation sum differs), and G of the streamed tiles at 32^4 against the
oracle on 4,000 sampled sites (1e-13).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import lqcd

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


class site_walk:
    """LQCD_GAUSS_TILES=0 for the body: the library's Gauss kernels walk sites."""

    def __enter__(self):
        self.old = os.environ.get("LQCD_GAUSS_TILES")
        os.environ["LQCD_GAUSS_TILES"] = "0"

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ["LQCD_GAUSS_TILES"]
        else:
            os.environ["LQCD_GAUSS_TILES"] = self.old


# tiles of these lattices (x_0 x a x b sites, C slices per workgroup):
#   8x16x8x4   a = 8, b = 4, 2 x 2 tiles per slice, C = 2
#   16x8x8x4   a = 4, b = 4, 2 x 2 tiles per slice, C = 2 (four x_0 rows per wavefront)
#   8x8x4x4    a = 8, b = 4, one tile per slice (halo wraps onto the tile), C = 2
@pytest.mark.parametrize("ext,cold", [([8, 16, 8, 4], False), ([16, 8, 8, 4], False), ([8, 8, 4, 4], True)])
def test_gauss_tiles_vs_oracle(ext, cold):
    announce(f"gauss tiles vs oracle ext={ext} cold={cold}")
    ol = O.OracleLattice(4, ext, 1.0)
    u = ol.cold_start() if cold else ol.hot_start(O.SEED_RNG)
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, ext, 4), 6.0, u, lqcd.F64)
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


@pytest.mark.parametrize("precision", [lqcd.F32, lqcd.F64])
def test_gauss_tiles_vs_site_walk_32x4(precision):
    # 32^4: a = 2, b = 4, 16 x 8 tiles per slice, C = 8 (512 workgroups)
    announce(f"gauss tiles vs site walk 32^4 precision={precision}")
    st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 32, 5, 4, precision)
    ef = lqcd.LatticeStateEFSyncDefault(st)
    ef.reset_e_field(7, 0, 0.5 / 6.0)
    e0 = ef.e_field()
    g_t = ef.gauss()
    d_t = ef.gauss_sum_div()
    it_t = ef.project_to_gauss(200)
    e_t = ef.e_field()
    ef.set_e_field(e0)
    with site_walk():
        g_s = ef.gauss()
        d_s = ef.gauss_sum_div()
        it_s = ef.project_to_gauss(200)
        e_s = ef.e_field()
    print(f"checks: {it_t} streamed, {it_s} site walk; div {d_t!r} / {d_s!r}")
    assert np.max(np.abs(g_t - g_s)) <= 1e-14 * np.max(np.abs(g_s))
    # G of the streamed tiles against the oracle on a sample of sites: the
    # first and last rows and slices, where the halo and ring wrap, and
    # random interior sites
    ol = O.OracleLattice(4, [32] * 4, 1.0)
    u = st.link_matrix()
    rng = np.random.default_rng(11)
    sample = np.unique(np.concatenate([np.arange(0, 1024), np.arange(ol.nsite - 1024, ol.nsite),
                                       rng.integers(0, ol.nsite, 2048)]))
    ref = np.array([ol.gauss(u, e0, int(x)) for x in sample])
    assert np.max(np.abs(g_t[sample] - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    assert abs(d_t - d_s) <= 1e-13 * d_s
    assert it_t == it_s
    assert np.max(np.abs(e_t - e_s)) <= 1e-12
