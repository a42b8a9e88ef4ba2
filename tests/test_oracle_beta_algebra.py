"""The beta polynomials of the scale-dependent dynamic models against the
models they come from, on the CPU.

The LASDD procedure (DynamicSGS_LASDD_SM.py:266-319, ScalarLASDD_SM.py:
241-302) chooses beta so that the coefficient estimated at the first test
filter (width TFR*Delta, model term with beta) equals the one estimated at the
second (TFR^2*Delta, model term with beta^2):

  momentum:  <L.M(b)> / <M(b).M(b)> = <Q.N(b)> / <N(b).N(b)>,
             M(b) = 2L^2 (hat(|S|S) - TFR^2 b |S^| S^),
             N(b) = 2L^2 (hathat(|S|S) - TFR^4 b^2 |S^^| S^^);
  scalar:    <LT.X(b)> / <X(b).X(b)> = <QT.Y(b)> / <Y(b).Y(b)>,
             X(b) = L^2 (hat(|S| dTH) - TFR^2 b |S^| dTH^), Y the same at the
             second filter with TFR^4 b^2.

Cross-multiplied, each is a quintic in b.  The tests form M(b), N(b), X(b),
Y(b) explicitly from the filtered fields at several b, evaluate the
cross-multiplied condition as plane means, and compare it with the oracle's
quintic (its plane sums a1..e2 and coefficients aa..ff) evaluated at the same
b: a dropped term, a wrong sign, a swapped level or TFR power fails.  Where
the beta the oracle picks is a real root, it is checked to satisfy the
condition.  The Wong-Lilly quintic's expansion is checked against its product
form.
"""
import numpy as np
import pytest

import alfa_oracle as O
from alfa_amd import cases


def _grid(optSgs=1):
    cfg, *_ = cases.initial_fields("convective", nx=16, ny=16, nz=6, optSgs=optSgs)
    return O.Grid(cfg)


def _smooth(g, seed, kmax):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((g.nx, g.ny, g.nz))
    Fh = np.fft.rfft2(F, axes=(0, 1))
    kx = np.abs(np.fft.fftfreq(g.nx, 1 / g.nx))[:, None, None]
    ky = np.fft.rfftfreq(g.ny, 1 / g.ny)[None, :, None]
    Fh *= np.exp(-((kx / kmax) ** 2 + (ky / kmax) ** 2))
    return np.fft.irfft2(Fh, s=(g.nx, g.ny), axes=(0, 1))


def _contract(A, B):
    return A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + 2 * (A[3] * B[3] + A[4] * B[4] + A[5] * B[5])


def _poly(coeffs, b):
    """coeffs = (ff, ee, dd, cc, bb, aa), each (nz,): ff b^5 + ... + aa."""
    out = np.zeros_like(coeffs[0])
    for c in coeffs:
        out = out * b + c
    return out


def _momentum(g):
    u, v, w = (_smooth(g, 10 + i, 3.0) for i in range(3))
    Sij = [0.5 * _smooth(g, 20 + i, 3.0) for i in range(6)]
    S = O._Smag(*Sij)
    return O.LASDD_SM(g, u, v, w, *Sij, S), S


def test_momentum_quintic_is_the_two_level_condition():
    g = _grid(1)
    r, _ = _momentum(g)
    p = r["parts"]
    L, T = g.L, g.TFR
    for b in (0.3, 0.8, 1.0, 1.7, 2.9):
        M = [2 * L ** 2 * (p["SSh"][n] - T ** 2 * b * r["S_hat"] * p["Sh"][n]) for n in range(6)]
        N = [2 * L ** 2 * (p["SSd"][n] - T ** 4 * b ** 2 * r["S_hatd"] * p["Sd"][n])
             for n in range(6)]
        LM = O.PlanarMean(_contract(p["Lij"], M))
        MM = O.PlanarMean(_contract(M, M))
        QN = O.PlanarMean(_contract(p["Qij"], N))
        NN = O.PlanarMean(_contract(N, N))
        cond = LM * NN - QN * MM
        got = _poly(r["beta_coeffs"], b)
        np.testing.assert_allclose(got, cond, rtol=1e-9, atol=1e-12 * np.abs(cond).max(),
                                   err_msg=f"beta={b}")


def test_momentum_beta_satisfies_the_condition():
    """Where the beta the oracle picks is a real root of its quintic, the two
    levels' coefficients agree.  (The Laguerre iteration of Roots, kept as
    the reference has it, can also stop at a real point next to a complex
    pair of roots, where the polynomial is small but not zero: such levels
    are skipped here, as their beta is not a root.)"""
    g = _grid(1)
    r, _ = _momentum(g)
    p = r["parts"]
    L, T = g.L, g.TFR
    beta = r["beta1_1D"]
    found = 0
    for k in range(g.nz):
        b = beta[k]
        real_roots = [z.real for z in np.roots([c[k] for c in r["beta_coeffs"]])
                      if abs(z.imag) < 1e-9]
        if b == 1.0 or min(abs(b - z) for z in real_roots) > 1e-6 * max(1.0, abs(b)):
            continue   # the default, or not a root (see above)
        found += 1
        M = [2 * L ** 2 * (p["SSh"][n][..., k] - T ** 2 * b * r["S_hat"][..., k] * p["Sh"][n][..., k])
             for n in range(6)]
        N = [2 * L ** 2 * (p["SSd"][n][..., k]
                           - T ** 4 * b ** 2 * r["S_hatd"][..., k] * p["Sd"][n][..., k])
             for n in range(6)]
        c1 = np.mean(_contract([x[..., k] for x in p["Lij"]], M)) / np.mean(_contract(M, M))
        c2 = np.mean(_contract([x[..., k] for x in p["Qij"]], N)) / np.mean(_contract(N, N))
        assert abs(c1 - c2) <= 1e-5 * max(abs(c1), abs(c2)), (k, b, c1, c2)
    assert found > 0


def test_scalar_quintic_is_the_two_level_condition():
    g = _grid(1)
    mom, S = _momentum(g)
    TH = _smooth(g, 40, 3.0)
    dT = [_smooth(g, 41 + i, 3.0) for i in range(3)]
    r = O.ScalarLASDD_SM(g, mom, TH, *dT, S)
    p = r["parts"]
    L, T = g.L, g.TFR
    Sh, Sd = mom["S_hat"], mom["S_hatd"]
    dot = lambda A, B: A[0] * B[0] + A[1] * B[1] + A[2] * B[2]
    for b in (0.3, 0.8, 1.0, 1.7, 2.9):
        X = [L ** 2 * (p["SdTh"][i] - T ** 2 * b * Sh * p["dh"][i]) for i in range(3)]
        Y = [L ** 2 * (p["SdTd"][i] - T ** 4 * b ** 2 * Sd * p["dd"][i]) for i in range(3)]
        LX = O.PlanarMean(dot(p["LT"], X))
        XX = O.PlanarMean(dot(X, X))
        QY = O.PlanarMean(dot(p["QT"], Y))
        YY = O.PlanarMean(dot(Y, Y))
        cond = LX * YY - QY * XX
        got = _poly(r["beta_coeffs"], b)
        np.testing.assert_allclose(got, cond, rtol=1e-9, atol=1e-12 * np.abs(cond).max(),
                                   err_msg=f"beta={b}")


def test_scalar_up_down_terms_are_the_level_one_model():
    """T_up = LT.X(beta), T_dn = X(beta).X(beta) pointwise (the ratio gives
    Cs2/Pr, ScalarLASDD_SM.py:304-312)."""
    g = _grid(1)
    mom, S = _momentum(g)
    TH = _smooth(g, 50, 3.0)
    dT = [_smooth(g, 51 + i, 3.0) for i in range(3)]
    r = O.ScalarLASDD_SM(g, mom, TH, *dT, S)
    p = r["parts"]
    b = r["beta2_1D"].reshape(1, 1, -1)
    X = [g.L ** 2 * (p["SdTh"][i] - g.TFR ** 2 * b * mom["S_hat"] * p["dh"][i]) for i in range(3)]
    up = sum(p["LT"][i] * X[i] for i in range(3))
    dn = sum(X[i] * X[i] for i in range(3))
    np.testing.assert_allclose(r["T_up"], up, rtol=1e-10, atol=1e-14 * np.abs(up).max())
    np.testing.assert_allclose(r["T_dn"], dn, rtol=1e-10, atol=1e-14 * np.abs(dn).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wl_quintic_expansion(seed):
    """The Wong-Lilly coefficients expand (a1 + a2 b^2)(a3 + a4 b + a5 b^2)
    - (a6 + a7 b)(a8 + a9 b^2 + a10 b^4) with a2, a4, a5, a7, a9, a10 the
    TFR-power multiples of the four plane means (DynamicSGS_LASDD_WL.py:
    255-274).  The multiples themselves follow the reference and are not
    checked independently."""
    g = _grid(2)
    rng = np.random.default_rng(seed)
    a1, a3, a6, a8 = (rng.standard_normal(4) for _ in range(4))
    T = g.TFR
    a2, a4, a5 = -(T ** (8 / 3)) * a1, -2 * T ** (4 / 3) * a3, T ** (8 / 3) * a3
    a7, a9, a10 = -T ** (4 / 3) * a6, -2 * T ** (8 / 3) * a8, T ** (16 / 3) * a8
    co = O._wl_quintic(g, a1, a3, a6, a8)
    for b in (0.4, 1.0, 2.2):
        prod = (a1 + a2 * b ** 2) * (a3 + a4 * b + a5 * b ** 2) - (a6 + a7 * b) * (
            a8 + a9 * b ** 2 + a10 * b ** 4)
        np.testing.assert_allclose(_poly(co, b), prod, rtol=1e-12, atol=1e-12)
