"""Host-side parameters of the HMC mirror (no device)."""
import lqcd


def test_gauss_projection_bound_grows_with_the_volume():
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc = lqcd.HybridMonteCarloDiagnostic(0.025, 40, integ, 1)
    # 25-56 checks are needed at 8^4, about 2,500 at 32^4
    assert hmc.gauss_max_iter(8 ** 4) == 1016
    assert hmc.gauss_max_iter(32 ** 4) == 5096 > 2500
    assert hmc.gauss_max_iter(48 ** 3 * 96) == 42472
    fixed = lqcd.HybridMonteCarloDiagnostic(0.025, 40, integ, 1, gauss_max_iter=7)
    assert fixed.gauss_max_iter(32 ** 4) == 7
