"""Mixed precision at a developed state (stable boundary layer, 64^3, one
simulated hour, statistics averaged over the last 30 minutes).

After the flow has become turbulent, runs that differ at rounding level are
decorrelated and their time averages differ by finite-sample variability.
The bar: the mixed run's mean wind, theta, resolved-TKE and dynamic
coefficient (Cs2, Cs2/Pr) profiles differ from the FP64 run's by no more
than twice what an FP64 run whose initial theta is perturbed at float
rounding (2^-24) differs (plus a small floor), i.e. mixed precision adds no
bias beyond that variability; and the mean theta profile agrees to 1e-3.
"""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from alfa_amd import stats  # noqa: E402


def test_mixed_coefficients_at_a_developed_state():
    r = stats.developed_state_comparison(n=64, hours=1.0, avg_min=30.0)
    print(r)
    m, p = r["mixed_vs_fp64"], r["fp64_perturbed_vs_fp64"]
    assert r["turbulent_levels"] >= 10
    for k, floor in (("mean_wind_profile", 0.005), ("tke_profile", 0.02), ("Cs2_2_profile", 0.05),
                     ("Cs2PrRatio_profile", 0.05)):
        assert m[k] <= 2.0 * p[k] + floor, (k, m[k], p[k])
    assert m["mean_theta_profile"] < 1e-3
