"""Independent pins of four oracle functions (CPU only).

- ``Wall`` (SGSStresses_SM.py:63-90): the MOST wall stress is the
  log-law (or log-linear, with a stability correction) drag
  tau = -u*^2 (U, V) / M along the absolute surface wind, with u* formed
  WITHOUT the 1e-3 floor that the surface-flux routines apply.
- The prescribed-temperature surface flux (SurfaceFlux.py:335-443): in the
  neutral limit the diagnosed heat flux is the bulk transfer law
  q = kappa u* (T_s - theta) / ln(z1 / zT), homogeneous (plane means) and
  heterogeneous (per column); zT is the screen height zTemperature when it
  is set, else z0T; a surface colder than the air gives q < 0 and the
  Businger-Dyer stable functions at the Obukhov length of that flux.
- ``Advection_Dealias`` (NSE_AdvectionTerms.py:65-155): with band-limited
  inputs (|k| < n/3) the 3/2-rule product is exact, so each product of two
  single Fourier modes is the two modes of the product-to-sum identity,
  and a product mode at or beyond n/2 is dropped, not folded back as the
  non-dealiased product would alias it.
- The Wong-Lilly beta quintic (DynamicSGS_LASDD_WL.py:232-275,
  DynamicSGS_ScalarLASDD_WL.py:212-243): it is the cross-multiplied
  two-level condition of the scale-dependent procedure with the WL model
  terms at each level (C(TFR Delta) = beta C(Delta) at the first test filter,
  C(TFR^2 Delta) = beta^2 C(Delta) at the second), formed here from the
  filtered tensors at several beta; it therefore vanishes where either
  level's model term vanishes.

Every expected value is written from the physics or the mathematics (the
log law, the Obukhov length, the product-to-sum identity, the model
tensors), not from the oracle's arithmetic, so a dropped term, a wrong
height, sign, TFR power or a folded mode fails a test.
"""
import math

import numpy as np
import pytest

import alfa_oracle as O
from alfa_amd import cases

KAPPA = 0.4
GRAV = 9.81


def _grid(name="convective", **over):
    cfg, *_ = cases.initial_fields(name, nx=over.pop("nx", 8), ny=over.pop("ny", 8),
                                   nz=over.pop("nz", 6))
    cfg.update(over)
    return O.Grid(cfg)


def _z1_m(g):
    """height of the first uvp level in metres"""
    return 0.5 * g.dz * g.z_scale


# ---------------------------------------------------------------------------
# Wall
# ---------------------------------------------------------------------------

def test_wall_is_the_neutral_log_law_drag_along_the_absolute_wind():
    g = _grid(Ugal=3.0)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 2, (g.nx, g.ny, g.nz))
    v = rng.uniform(-1, 1, (g.nx, g.ny, g.nz))
    U, V = u[:, :, 0] + 3.0, v[:, :, 0]
    M = np.hypot(U, V)
    z = np.zeros((g.nx, g.ny))
    txz, tyz = O.Wall(g, u, v, M, z, z)
    ustar = KAPPA * M / math.log(_z1_m(g) / g.z0m)
    # magnitude u*^2, direction opposite to the absolute surface wind
    np.testing.assert_allclose(np.hypot(txz, tyz), ustar ** 2, rtol=1e-13)
    np.testing.assert_allclose(txz * V, tyz * U, atol=1e-15 * np.max(ustar) ** 2 * np.max(M))
    assert np.all(txz * U <= 0) and np.all(tyz * V <= 0)


def test_wall_log_linear_stable_correction():
    # stable Businger-Dyer: U(z) = u*/kappa (ln(z/z0) + 5 (z - z0)/L); the
    # reference's psi carries the sign convention psi = +5 z/L here
    g = _grid()
    Lob = 37.0                                   # Obukhov length, metres
    z1, z0 = _z1_m(g), g.z0m
    rng = np.random.default_rng(2)
    u = rng.uniform(1, 3, (g.nx, g.ny, g.nz))
    v = rng.uniform(-1, 1, (g.nx, g.ny, g.nz))
    M = np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0])
    one = np.ones((g.nx, g.ny))
    txz, tyz = O.Wall(g, u, v, M, 5 * z1 / Lob * one, 5 * z0 / Lob * one)
    ustar = KAPPA * M / (math.log(z1 / z0) + 5 * (z1 - z0) / Lob)
    np.testing.assert_allclose(np.hypot(txz, tyz), ustar ** 2, rtol=1e-13)


def test_wall_has_no_ustar_floor():
    # a near-calm surface: kappa M / ln(z1/z0) is far below the 1e-3 floor of
    # the surface-flux routines, and the wall stress keeps the unfloored u*
    g = _grid()
    u = np.full((g.nx, g.ny, g.nz), 1e-5) - g.Ugal
    v = np.zeros_like(u)
    M = np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0])
    z = np.zeros((g.nx, g.ny))
    txz, tyz = O.Wall(g, u, v, M, z, z)
    ustar = KAPPA * 1e-5 / math.log(_z1_m(g) / g.z0m)
    assert ustar < 1e-5
    np.testing.assert_allclose(-txz, ustar ** 2, rtol=1e-12)
    assert np.all(tyz == 0)
    # the surface flux on the same wind does apply the floor
    sfc = O.SurfaceFlux(g, u, v, np.zeros_like(u), (z, z, z, z, z + 1, z + 1))
    assert np.all(sfc[1] == 1e-3)


# ---------------------------------------------------------------------------
# Prescribed-temperature surface flux
# ---------------------------------------------------------------------------

def _sfc_fields(g, seed, th0=0.0, th_amp=0.0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(2, 4, (g.nx, g.ny, g.nz))
    v = rng.uniform(-1, 1, (g.nx, g.ny, g.nz))
    TH = th0 + th_amp * rng.uniform(-1, 1, (g.nx, g.ny, g.nz))
    return u, v, TH


@pytest.mark.parametrize("zT", [0.0, 2.0])
def test_prescribed_temperature_neutral_bulk_law_homogeneous(zT):
    g = _grid("stable", optSurfBC=2, optSurfFlux=0, zTemperature=zT, z0T=0.05)
    u, v, TH = _sfc_fields(g, 3, th0=0.4, th_amp=0.2)
    z = np.zeros((g.nx, g.ny))
    Ts = 0.1                                  # surface anomaly, 0.3 K below the air on average
    M, us, q2, qa, iob, most = O.SurfaceFlux(g, u, v, TH, (z, z, z, z, z + 1, z + 1), Ts)
    z1 = _z1_m(g)
    zref = zT if zT > 0 else 0.05
    Mbar = np.mean(np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0]))
    thbar = np.mean(TH[:, :, 0])
    ustar = max(KAPPA * Mbar / math.log(z1 / g.z0m), 1e-3)
    q = KAPPA * ustar * (Ts - thbar) / math.log(z1 / zref)
    np.testing.assert_allclose(M, Mbar, rtol=1e-14)
    np.testing.assert_allclose(us, ustar, rtol=1e-14)
    np.testing.assert_allclose(q2, q, rtol=1e-13)
    assert abs(qa - q) <= 1e-13 * abs(q) and q < 0


def test_prescribed_temperature_neutral_bulk_law_heterogeneous():
    g = _grid("stable", optSurfBC=2, optSurfFlux=1, zTemperature=0.0, z0T=0.05)
    u, v, TH = _sfc_fields(g, 4, th0=-0.2, th_amp=0.3)
    z = np.zeros((g.nx, g.ny))
    Ts = 0.05
    M, us, q2, qa, iob, most = O.SurfaceFlux(g, u, v, TH, (z, z, z, z, z + 1, z + 1), Ts)
    z1 = _z1_m(g)
    Mloc = np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0])
    ustar = np.maximum(KAPPA * Mloc / math.log(z1 / g.z0m), 1e-3)
    q = KAPPA * ustar * (Ts - TH[:, :, 0]) / math.log(z1 / 0.05)
    np.testing.assert_allclose(M, Mloc, rtol=1e-14)
    np.testing.assert_allclose(us, ustar, rtol=1e-14)
    np.testing.assert_allclose(q2, q, rtol=1e-13)
    np.testing.assert_allclose(qa, q.mean(), rtol=1e-13)


@pytest.mark.parametrize("hetero", [0, 1])
@pytest.mark.parametrize("zT", [0.0, 2.0])
def test_prescribed_temperature_cold_surface_gives_the_stable_functions(hetero, zT):
    """A surface colder than the air: q < 0 and the updated MOST functions
    are Businger-Dyer stable, psi = 5 z/L, phi = 1 + 5 z1/L, at the Obukhov
    length L = -u*^3 theta_ref / (kappa g q) of the mean diagnosed flux, with
    theta_ref the absolute mean air temperature; psi_h0 sits at the
    temperature reference height (zTemperature, else z0T)."""
    g = _grid("stable", optSurfBC=2, optSurfFlux=hetero, zTemperature=zT, z0T=0.05)
    u, v, TH = _sfc_fields(g, 5, th0=0.5, th_amp=0.1)
    z = np.zeros((g.nx, g.ny))
    M, us, q2, qa, iob, most = O.SurfaceFlux(g, u, v, TH, (z, z, z, z, z + 1, z + 1), -0.5)
    assert qa < 0
    z1, z0, zref = _z1_m(g), g.z0m, (zT if zT > 0 else 0.05)
    th_ref = np.mean(TH[:, :, 0]) + g.T_0
    # the Obukhov length in metres: u* in m/s, q in K m/s (u_scale = TH_scale = 1)
    Lob = -us ** 3 * th_ref / (KAPPA * GRAV * qa)
    assert np.all(Lob > 0)
    np.testing.assert_allclose(most[0], 5 * z1 / Lob, rtol=1e-12)
    np.testing.assert_allclose(most[1], 5 * z0 / Lob, rtol=1e-12)
    np.testing.assert_allclose(most[2], 5 * z1 / Lob, rtol=1e-12)
    np.testing.assert_allclose(most[3], 5 * zref / Lob, rtol=1e-12)
    np.testing.assert_allclose(most[4], 1 + 5 * z1 / Lob, rtol=1e-12)
    np.testing.assert_allclose(most[5], 1 + 5 * z1 / Lob, rtol=1e-12)
    # the nondimensional inverse length the step carries: z_scale / L
    np.testing.assert_allclose(iob, g.z_scale / Lob, rtol=1e-12)


def test_prescribed_temperature_uses_the_previous_stability_in_the_flux():
    """u* and q are formed with the MOST functions carried in from the
    previous step (stable log-linear), then updated."""
    g = _grid("stable", optSurfBC=2, optSurfFlux=0, zTemperature=0.0, z0T=0.05)
    u, v, TH = _sfc_fields(g, 6, th0=0.3, th_amp=0.0)
    z1, z0, zT = _z1_m(g), g.z0m, 0.05
    Lprev = 25.0
    one = np.ones((g.nx, g.ny))
    most_in = (5 * z1 / Lprev * one, 5 * z0 / Lprev * one, 5 * z1 / Lprev * one,
               5 * zT / Lprev * one, (1 + 5 * z1 / Lprev) * one, (1 + 5 * z1 / Lprev) * one)
    M, us, q2, qa, iob, most = O.SurfaceFlux(g, u, v, TH, most_in, 0.0)
    Mbar = np.mean(np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0]))
    ustar = KAPPA * Mbar / (math.log(z1 / z0) + 5 * (z1 - z0) / Lprev)
    q = KAPPA * ustar * (0.0 - 0.3) / (math.log(z1 / zT) + 5 * (z1 - zT) / Lprev)
    np.testing.assert_allclose(us, ustar, rtol=1e-13)
    np.testing.assert_allclose(qa, q, rtol=1e-13)


# ---------------------------------------------------------------------------
# Advection_Dealias
# ---------------------------------------------------------------------------

class Modes:
    """A field sum_m P_m(z) cos(kx_m x + ky_m y + phi_m) on the [0, 2pi)^2 grid."""

    def __init__(self, terms):
        self.t = list(terms)      # (P (nz,), kx, ky, phi)

    @staticmethod
    def one(P, kx, ky, phi):
        return Modes([(np.asarray(P, float), kx, ky, phi)])

    def __add__(self, o):
        return Modes(self.t + o.t)

    def __neg__(self):
        return Modes([(-P, a, b, f) for P, a, b, f in self.t])

    def __sub__(self, o):
        return self + (-o)

    def __mul__(self, o):
        # cos A cos B = (cos(A + B) + cos(A - B)) / 2
        out = []
        for P, a, b, f in self.t:
            for Q, c, d, h in o.t:
                out.append((0.5 * P * Q, a + c, b + d, f + h))
                out.append((0.5 * P * Q, a - c, b - d, f - h))
        return Modes(out)

    def zmap(self, fn):
        return Modes([(fn(P), a, b, f) for P, a, b, f in self.t])

    def truncate(self, nx, ny):
        """the modes a sharp cut at the n-grid's |k| < n/2 keeps (Nyquist too
        is dropped)"""
        return Modes([t for t in self.t if abs(t[1]) < nx // 2 and abs(t[2]) < ny // 2])

    def has(self, kx, ky):
        return any((abs(a), abs(b)) == (kx, ky) and np.any(P != 0) for P, a, b, f in self.t)

    def eval(self, g):
        x = np.arange(g.nx)[:, None, None] * g.dx
        y = np.arange(g.ny)[None, :, None] * g.dy
        out = np.zeros((g.nx, g.ny, g.nz))
        for P, a, b, f in self.t:
            out += P[None, None, :] * np.cos(a * x + b * y + f)
        return out


def _stag_uvp(P):
    """w-node profile averaged to uvp levels, with the wall and top rows of
    the rotational advection (first level: half the first w level, top: the
    top w value)"""
    nz = len(P)
    out = np.empty(nz)
    out[1:nz - 1] = 0.5 * (P[1:nz - 1] + P[2:nz])
    out[0] = 0.5 * P[1]
    out[nz - 1] = P[nz - 1]
    return out


def test_advection_dealias_band_limited_products_are_exact_and_cut_at_half_n():
    g = _grid("neutral", nx=32, ny=32, nz=5)
    assert g.optDealias == 1 and g.mx == 48
    rng = np.random.default_rng(7)
    prof = lambda: rng.uniform(0.5, 1.5, g.nz)
    # every input mode has |k| <= 9 < n/3: exact on the 48-point padded grid
    u = Modes.one(prof(), 9, 2, 0.3)
    v = Modes.one(prof(), 0, 9, 0.7)
    w = Modes.one(prof(), 2, -5, 1.1)
    dudy = Modes.one(prof(), 0, 9, 0.2)
    dudz = Modes.one(prof(), 1, 7, -0.4)
    dvdx = Modes.one(prof(), 8, 1, 0.9)
    dvdz = Modes.one(prof(), 5, -3, 0.1)
    dwdx = Modes.one(prof(), 6, 0, -1.3)
    dwdy = Modes.one(prof(), 0, 8, 0.5)
    ev = lambda m: m.eval(g)
    Cx, Cy, Cz = O.Advection_Dealias(g, ev(u), ev(v), ev(w), ev(dudy), ev(dudz), ev(dvdx),
                                     ev(dvdz), ev(dwdx), ev(dwdy))
    n = g.nx
    # (omega x u)_x = v (du/dy - dv/dx) + w (du/dz - dw/dx) averaged to uvp nodes
    ax = v * (dudy - dvdx)
    bx = (w * (dudz - dwdx)).zmap(_stag_uvp)
    ex = (ax + bx).truncate(n, n)
    ay = u * (dvdx - dudy)
    by = (w * (dvdz - dwdy)).zmap(_stag_uvp)
    ey = (ay + by).truncate(n, n)
    # (omega x u)_z at interior w nodes from the uvp-node u, v averaged to w nodes
    def to_w(P):
        out = np.zeros_like(P)
        out[1:-1] = 0.5 * (P[:-2] + P[1:-1])
        return out
    inner = lambda P: np.concatenate([[0.0], P[1:-1], [0.0]])
    ez = (u.zmap(to_w) * (dwdx - dudz).zmap(inner)
          + v.zmap(to_w) * (dwdy - dvdz).zmap(inner)).truncate(n, n)
    scale = max(np.abs(Cx).max(), np.abs(Cy).max(), np.abs(Cz).max())
    for got, exp in ((Cx, ex), (Cy, ey), (Cz, ez)):
        np.testing.assert_allclose(got, exp.eval(g), atol=1e-12 * scale)
    # the pairs above n/2 are really there before the cut: v dudy makes ky 18,
    # u dvdx kx 17, v dwdy ky 17; after the cut they are gone, not folded
    # back to 32 - 18 = 14, 32 - 17 = 15 as the 32-point product aliases them
    assert (v * dudy).has(0, 18) and (u * dvdx).has(17, 3) and (v * dwdy).has(0, 17)
    spec = lambda F: np.abs(np.fft.rfft2(F, axes=(0, 1))).max(axis=2)
    for F, (kx, ky) in ((Cx, (0, 14)), (Cy, (15, 3)), (Cz, (0, 15))):
        assert spec(F)[kx, ky] < 1e-10 * scale * n * n
    # the non-dealiased product does fold them back
    Nx, Ny, Nz = O.Advection_NoDealias(g, ev(u), ev(v), ev(w), ev(dudy), ev(dudz), ev(dvdx),
                                       ev(dvdz), ev(dwdx), ev(dwdy))
    assert spec(Nx)[0, 14] > 1e-3 * scale * n * n


# ---------------------------------------------------------------------------
# Wong-Lilly beta quintic
# ---------------------------------------------------------------------------

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
    out = np.zeros_like(coeffs[0])
    for c in coeffs:
        out = out * b + c
    return out


def _wl_grid():
    cfg, *_ = cases.initial_fields("convective", nx=16, ny=16, nz=6, optSgs=2)
    return O.Grid(cfg)


def _wl_momentum(g):
    u, v, w = (_smooth(g, 60 + i, 3.0) for i in range(3))
    Sij = [0.5 * _smooth(g, 70 + i, 3.0) for i in range(6)]
    r = O.LASDD_WL(g, u, v, w, *Sij)
    F1 = lambda F: O.Filtering_Level1(g, O.FFT(F))
    F2 = lambda F: O.Filtering_Level2(g, O.FFT(F))
    uu = (r["u_"], r["v_"], r["w_"])
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    Lij = [F1(uu[a] * uu[b]) - F1(uu[a]) * F1(uu[b]) for a, b in pairs]
    Qij = [F2(uu[a] * uu[b]) - F2(uu[a]) * F2(uu[b]) for a, b in pairs]
    Sh = [F1(s) for s in Sij]
    Sd = [F2(s) for s in Sij]
    return r, Lij, Qij, Sh, Sd


def test_wl_momentum_quintic_is_the_two_level_condition():
    """Level 1 (width TFR Delta): WL model tensor M(b) = (1 - TFR^(4/3) b) S^
    (the strain is linear, so the filtered grid-level term and the test-level
    term share S^); level 2 (TFR^2 Delta): N(b) = (1 - TFR^(8/3) b^2) S^^.
    The reference contracts the level-2 Leonard term Q with the level-1
    strain (a1_terms, DynamicSGS_LASDD_WL.py:236-238), so its numerator is
    <Q.N'(b)> with N'(b) = (1 - TFR^(8/3) b^2) S^.  Cross-multiplied,
    <Q.N'> <M.M> - <L.M> <N.N> is the quintic (the constant 2 L^(4/3) of
    each model term is a common factor and left out)."""
    g = _wl_grid()
    r, Lij, Qij, Sh, Sd = _wl_momentum(g)
    a = g.TFR ** (4 / 3)
    for b in (0.2, 0.6, 1.0, 1.45, 3.1):
        M = [(1 - a * b) * s for s in Sh]
        N = [(1 - a * a * b * b) * s for s in Sd]
        Np = [(1 - a * a * b * b) * s for s in Sh]
        cond = (O.PlanarMean(_contract(Qij, Np)) * O.PlanarMean(_contract(M, M))
                - O.PlanarMean(_contract(Lij, M)) * O.PlanarMean(_contract(N, N)))
        got = _poly(r["beta_coeffs"], b)
        np.testing.assert_allclose(got, cond, rtol=1e-9, atol=1e-12 * np.abs(cond).max(),
                                   err_msg=f"beta={b}")


def test_wl_scalar_quintic_is_the_two_level_condition():
    g = _wl_grid()
    r, Lij, Qij, Sh, Sd = _wl_momentum(g)
    TH = _smooth(g, 80, 3.0)
    dT = [_smooth(g, 81 + i, 3.0) for i in range(3)]
    s = O.ScalarLASDD_WL(g, r, TH, *dT)
    F1 = lambda F: O.Filtering_Level1(g, O.FFT(F))
    F2 = lambda F: O.Filtering_Level2(g, O.FFT(F))
    nz = g.nz
    THz = np.zeros_like(TH)                        # dTHdz averaged to uvp nodes
    THz[:, :, 1:nz - 1] = 0.5 * (dT[2][:, :, 1:nz - 1] + dT[2][:, :, 2:nz])
    THz[:, :, 0] = dT[2][:, :, 0]
    THz[:, :, nz - 1] = dT[2][:, :, nz - 1]
    grads = (dT[0], dT[1], THz)
    dh = [F1(x) for x in grads]
    dd = [F2(x) for x in grads]
    uu = (r["u_"], r["v_"], r["w_"])
    LT = [F1(uu[i] * TH) - F1(uu[i]) * F1(TH) for i in range(3)]
    QT = [F2(uu[i] * TH) - F2(uu[i]) * F2(TH) for i in range(3)]
    dot = lambda A, B: A[0] * B[0] + A[1] * B[1] + A[2] * B[2]
    a = g.TFR ** (4 / 3)
    for b in (0.2, 0.6, 1.0, 1.45, 3.1):
        X = [(1 - a * b) * x for x in dh]
        Y = [(1 - a * a * b * b) * x for x in dd]
        Yp = [(1 - a * a * b * b) * x for x in dh]
        cond = (O.PlanarMean(dot(QT, Yp)) * O.PlanarMean(dot(X, X))
                - O.PlanarMean(dot(LT, X)) * O.PlanarMean(dot(Y, Y)))
        got = _poly(s["beta_coeffs"], b)
        np.testing.assert_allclose(got, cond, rtol=1e-9, atol=1e-12 * np.abs(cond).max(),
                                   err_msg=f"beta={b}")
    # the up and down terms are the level-1 WL model vector
    # X(beta) = L^(4/3) (dTH^ - TFR^(4/3) beta dTH^^): T_up = LT.X, T_dn = X.X
    bb = s["beta2_1D"].reshape(1, 1, -1)
    Xb = [g.L ** (4 / 3) * (dh[i] - a * bb * dd[i]) for i in range(3)]
    np.testing.assert_allclose(s["T_up"], dot(LT, Xb), rtol=1e-10,
                               atol=1e-13 * np.abs(s["T_up"]).max())
    np.testing.assert_allclose(s["T_dn"], dot(Xb, Xb), rtol=1e-10,
                               atol=1e-13 * np.abs(s["T_dn"]).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wl_quintic_vanishes_where_a_level_model_term_vanishes(seed):
    """The condition carries the factors (1 - TFR^(4/3) b) (level 1) and
    (1 - TFR^(8/3) b^2) (level 2), so for any plane means the quintic has a
    double root at b = TFR^(-4/3) and a root at b = -TFR^(-4/3)."""
    g = _wl_grid()
    rng = np.random.default_rng(seed)
    a1, a3, a6, a8 = (rng.standard_normal(4) for _ in range(4))
    co = O._wl_quintic(g, a1, a3, a6, a8)
    r0 = g.TFR ** (-4 / 3)
    scale = np.max(np.abs(np.stack(co)), axis=0)
    assert np.all(np.abs(_poly(co, r0)) < 1e-12 * scale)
    assert np.all(np.abs(_poly(co, -r0)) < 1e-12 * scale)
    d = [c * (5 - i) for i, c in enumerate(co[:-1])]          # derivative, degree 4
    assert np.all(np.abs(_poly(d, r0)) < 1e-11 * scale)
