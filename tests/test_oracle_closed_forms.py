"""Closed-form checks of the oracle's PDE and SGS terms (CPU only).

Each test builds fields whose result follows from the mathematics or the
physics (vector identities, the Smagorinsky model's definition, exact
derivatives of single modes and linear profiles), not from the oracle's
formula, so a dropped term, a wrong sign, index or averaging fails one of
them.  Grid: the nondimensional horizontal domain is [0, 2 pi)^2; u, v, TH
live on uvp levels z = (k + 1/2) dz, w on w levels z = k dz.
"""
import numpy as np
import pytest
from scipy import ndimage

import alfa_oracle as O
from alfa_amd import cases


def grid(nx=16, ny=12, nz=10, **over):
    cfg, *_ = cases.initial_fields("convective", nx=nx, ny=ny, nz=nz)
    cfg.update(over)
    g = O.Grid(cfg)
    x = np.arange(g.nx) * g.dx
    y = np.arange(g.ny) * g.dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    X = np.repeat(X[:, :, None], g.nz, axis=2)
    Y = np.repeat(Y[:, :, None], g.nz, axis=2)
    zu = (np.arange(g.nz) + 0.5) * g.dz
    zw = np.arange(g.nz) * g.dz
    Zu = np.broadcast_to(zu, X.shape).copy()
    Zw = np.broadcast_to(zw, X.shape).copy()
    return g, X, Y, Zu, Zw


def test_advection_is_omega_cross_u_for_horizontal_flow():
    # u = (A sin y, B sin x, 0): omega x u = (u.grad)u - grad(|u|^2/2)
    g, X, Y, Zu, Zw = grid()
    A, B = 1.3, -0.7
    u, v, w = A * np.sin(Y), B * np.sin(X), np.zeros_like(X)
    zero = np.zeros_like(X)
    Cx, Cy, Cz = O.Advection_NoDealias(g, u, v, w, A * np.cos(Y), zero, B * np.cos(X), zero,
                                       zero, zero)
    ex = A * B * np.sin(X) * np.cos(Y) - B * B * np.sin(X) * np.cos(X)
    ey = A * B * np.sin(Y) * np.cos(X) - A * A * np.sin(Y) * np.cos(Y)
    assert np.allclose(Cx, ex, atol=1e-13)
    assert np.allclose(Cy, ey, atol=1e-13)
    assert np.all(Cz == 0)


def test_advection_vertical_shear_on_the_staggered_grid():
    # u = a z (uvp nodes), w = W above the ground (w nodes), v = 0:
    # (omega x u)_x = (du/dz) w = a W ; (omega x u)_z = -(du/dz) u = -a (a z_w)
    g, X, Y, Zu, Zw = grid()
    a, W = 0.8, 0.3
    u = a * Zu
    v = np.zeros_like(X)
    w = np.full_like(X, W)
    w[:, :, 0] = 0.0
    dudz = np.full_like(X, a)
    zero = np.zeros_like(X)
    Cx, Cy, Cz = O.Advection_NoDealias(g, u, v, w, zero, dudz, zero, zero, zero, zero)
    nz = g.nz
    assert np.allclose(Cx[:, :, 1:nz - 1], a * W, atol=1e-13)
    # first uvp level sits between the wall (w = 0) and the first w level
    assert np.allclose(Cx[:, :, 0], 0.5 * a * W, atol=1e-13)
    assert np.allclose(Cy, 0, atol=1e-13)
    assert np.allclose(Cz[:, :, 1:nz - 1], -a * (a * Zw[:, :, 1:nz - 1]), atol=1e-12)
    assert np.all(Cz[:, :, 0] == 0) and np.all(Cz[:, :, nz - 1] == 0)


def test_scalar_advection_is_u_dot_grad_theta():
    g, X, Y, Zu, Zw = grid()
    U, V, W, b = 0.9, -0.4, 0.25, 1.7
    u, v, w = np.full_like(X, U), np.full_like(X, V), np.full_like(X, W)
    w[:, :, 0] = 0.0
    dsdx, dsdy = np.cos(X) * np.cos(Y), -np.sin(X) * np.sin(Y)   # theta = sin x cos y + b z
    dsdz = np.full_like(X, b)
    adv = O.ScalarAdvection(g, u, v, w, dsdx, dsdy, dsdz)
    nz = g.nz
    horiz = U * dsdx + V * dsdy
    assert np.allclose(adv[:, :, 1:nz - 1], horiz[:, :, 1:nz - 1] + W * b, atol=1e-13)
    assert np.allclose(adv[:, :, 0], horiz[:, :, 0] + 0.5 * W * b, atol=1e-13)


@pytest.mark.parametrize("opt", [0, 1])
def test_buoyancy_of_a_zero_mean_perturbation(opt):
    # theta = Theta(z) + A sin x cos y: the buoyancy at a w level is g theta'/theta_ref
    # averaged from the two uvp levels around it; 0 at the ground
    g, X, Y, Zu, Zw = grid(optBuoyancy=opt)
    A = 0.6
    Theta = 2.0 * Zu
    pert = A * np.sin(X) * np.cos(Y)
    b = O.Buoyancy(g, Theta + pert)
    T0 = g.T_0_nondim
    ref = T0 + Theta if opt == 1 else np.full_like(Theta, T0)
    r = pert / ref
    exp = np.zeros_like(b)
    exp[:, :, 1:] = g.g_nondim * 0.5 * (r[:, :, 1:] + r[:, :, :-1])
    assert np.allclose(b, exp, rtol=1e-12, atol=1e-10)
    assert np.all(b[:, :, 0] == 0)


def test_div_stress_and_div_flux_single_modes_and_linear_profiles():
    g, X, Y, Zu, Zw = grid()
    c, d = 0.5, -1.25
    txx = np.sin(X)
    tyy = np.cos(2 * Y)
    txy = np.sin(X) * np.cos(Y)
    txz = c * Zw + np.cos(X)
    tyz = c * Zw + np.sin(Y)
    tzz = d * Zu
    dtx, dty, dtz = O.DivStress(g, txx, tyy, tzz, txy, txz, tyz)
    nz = g.nz
    i = slice(1, nz - 1)   # both z stencils are interior here
    ex = np.cos(X) - np.sin(X) * np.sin(Y) + c
    ey = np.cos(X) * np.cos(Y) - 2 * np.sin(2 * Y) + c
    ez = -np.sin(X) + np.cos(Y) + d
    assert np.allclose(dtx[:, :, i], ex[:, :, i], atol=1e-12)
    assert np.allclose(dty[:, :, i], ey[:, :, i], atol=1e-12)
    assert np.allclose(dtz[:, :, i], ez[:, :, i], atol=1e-12)
    q = O.DivFlux(g, np.cos(3 * X), np.sin(Y), c * Zw)
    assert np.allclose(q[:, :, i], (-3 * np.sin(3 * X) + np.cos(Y) + c)[:, :, i], atol=1e-12)


def test_strain_rates_of_a_shear_and_a_cellular_flow():
    # u = (a z + sin y, sin x, 0): S12 = (cos y + cos x)/2, S13 = a/2, others 0,
    # |S| = sqrt(2 S_ij S_ij) = 2 sqrt(S12^2 + S13^2)
    g, X, Y, Zu, Zw = grid()
    a = 0.9
    zero = np.zeros_like(X)
    dudy, dvdx = np.cos(Y), np.cos(X)
    dudz = np.full_like(X, a)
    S12e = 0.5 * (np.cos(Y) + np.cos(X))
    Se = 2 * np.sqrt(S12e ** 2 + (a / 2) ** 2)
    nz = g.nz
    for f in (O.StrainsUVPnodes, O.StrainsWnodes):
        S11, S22, S33, S12, S13, S23, S = f(g, zero, dvdx, zero, dudy, zero, zero, dudz, zero, zero)
        i = slice(1, nz - 1)
        assert np.allclose(S12[:, :, i], S12e[:, :, i], atol=1e-14)
        assert np.allclose(S13[:, :, i], a / 2, atol=1e-14)
        assert np.allclose(S[:, :, i], Se[:, :, i], atol=1e-13)
        for t in (S11, S22, S33, S23):
            assert np.allclose(t, 0)


def test_smagorinsky_stresses_of_a_pure_shear():
    # Smagorinsky: tau_ij = -2 (Cs Delta)^2 |S| S_ij with |S| = sqrt(2 S_ij S_ij);
    # a pure shear S12 = s gives |S| = 2|s| and tau_12 = -4 Cs2 Delta^2 |s| s
    g, X, Y, Zu, Zw = grid()
    s, cs2 = 0.37, 0.021
    zero = np.zeros_like(X)
    S12 = np.full_like(X, s)
    S = O._Smag(zero, zero, zero, S12, zero, zero)
    assert np.allclose(S, 2 * abs(s))
    Cs2 = np.full_like(X, cs2)
    txx, tyy, tzz, txy = O.StressesUVP(g, zero.copy(), zero.copy(), zero.copy(), S12, S, Cs2)
    assert np.allclose(txy[:, :, :-1], -4 * cs2 * g.L ** 2 * abs(s) * s)
    assert np.all(txy[:, :, -1] == 0)
    txz, tyz = O.StressesW_interior(g, S12, zero, S, Cs2)
    assert np.allclose(txz[:, :, 1:-1], -4 * cs2 * g.L ** 2 * abs(s) * s)


def test_cs2_is_the_box_filtered_ratio_clipped_to_the_unit_interval():
    rng = np.random.default_rng(3)
    LM = rng.uniform(-0.5, 1.5, (8, 6, 4))
    MM = rng.uniform(0.2, 1.0, (8, 6, 4))
    MM[2, 3, 1] = 0.0
    MM[:, :, 3] = 1e-12           # |MM| below the 1e-10 floor everywhere on a level
    cs2 = O.Cs2_3D_from_LM_MM(LM, MM)
    a = ndimage.uniform_filter(LM, size=(3, 3, 1), mode="wrap")
    b = ndimage.uniform_filter(MM, size=(3, 3, 1), mode="wrap")
    r = a / b
    keep = (np.abs(b) >= 1e-10) & (r >= 0) & (r <= 1)
    assert np.allclose(cs2[keep], r[keep], rtol=1e-12)
    assert np.all(cs2[~keep] == 0)
    assert np.all(cs2[:, :, 3] == 0)
    assert keep.any() and (~keep[:, :, :3]).any()


def test_rayleigh_layer_is_a_raised_cosine_ramp():
    # 0 below z_damping, 1/RelaxTime at the top, sin^2 ramp in between, and 0
    # at the staggered level above the top of the domain
    g, *_ = grid(nz=40)
    full, stag = O.Initialize_RayleighDampingLayer(g)
    inv = 1.0 / g.RelaxTime_nondim
    zd, top = g.z_damping_nondim, g.l_z / g.z_scale
    for prof, z in ((full, np.arange(g.nz) * g.dz), (stag, (np.arange(g.nz) + 0.5) * g.dz)):
        inside = (z >= zd) & (z <= top * (1 + 1e-12))
        ramp = np.where(inside, inv * np.sin(0.5 * np.pi * (z - zd) / (top - zd)) ** 2, 0.0)
        assert np.allclose(prof, ramp, rtol=1e-12, atol=1e-14)
        assert np.all(np.diff(prof[inside]) >= -1e-15)
        assert np.all(prof[z < zd] == 0)
    assert np.isclose(full[-1], inv)


def test_rhs_scalar_damps_only_the_fluctuation_in_the_sponge():
    g, X, Y, Zu, Zw = grid(nz=40)
    th = 3.0 + 0.5 * np.sin(X)          # plane mean 3, fluctuation 0.5 sin x
    zero = np.zeros_like(X)
    R = O.RHS_Scalar(g, th, zero, zero)
    rs = g.RayleighDampCoeff_stag
    exp = -rs[None, None, :] * 0.5 * np.sin(X)
    assert np.allclose(R, exp, atol=1e-12)
    assert np.any(rs > 0) and np.any(rs == 0)
    adv, divq = np.cos(Y), np.sin(2 * X)
    R2 = O.RHS_Scalar(g, np.full_like(X, 3.0), adv, divq)
    assert np.allclose(R2, -adv - divq, atol=1e-13)


def test_rhs_momentum_sponge_and_buoyancy_terms():
    g, X, Y, Zu, Zw = grid(nz=40, f_coriolis=0.0)
    u = 2.0 + 0.3 * np.cos(Y)
    v = -1.0 + 0.2 * np.sin(X)
    w = 0.1 * np.sin(X + Y)
    zero = np.zeros_like(X)
    buoy = np.cos(X)
    Ru, Rv, Rw = O.RHS_Momentum(g, u, v, w, zero, zero, zero, buoy, zero, zero, zero)
    rs, rf = g.RayleighDampCoeff_stag, g.RayleighDampCoeff
    assert np.allclose(Ru, -rs * 0.3 * np.cos(Y), atol=1e-12)
    assert np.allclose(Rv, -rs * 0.2 * np.sin(X), atol=1e-12)
    assert np.allclose(Rw, buoy - rf * w, atol=1e-12)


def test_wong_lilly_stresses_and_fluxes():
    # Wong-Lilly: tau_ij = -2 C_WL Delta^(4/3) S_ij and q_i = -C Delta^(4/3) dtheta/dx_i,
    # both free of |S|; C on w levels is the mean of the two uvp levels around it
    g, X, Y, Zu, Zw = grid()
    s, c = 0.37, 0.021
    zero = np.zeros_like(X)
    S12 = np.full_like(X, s)
    C = np.full_like(X, c)
    txx, tyy, tzz, txy = O.StressesUVP_WL(g, zero.copy(), zero.copy(), zero.copy(), S12, C)
    d43 = g.L ** (4 / 3)
    assert np.allclose(txy[:, :, :-1], -2 * c * d43 * s)
    assert np.all(txy[:, :, -1] == 0)
    Cz = c * (1 + Zu)                      # varies with height
    txz, tyz = O.StressesW_interior_WL(g, S12, zero, Cz)
    cw = c * (1 + Zw)                      # uvp mean at w levels = C at z_w (linear)
    assert np.allclose(txz[:, :, 1:-1], -2 * d43 * cw[:, :, 1:-1] * s)
    gx = np.cos(X)
    qx, qy = O.ScalarFluxesUVP_WL(g, gx, zero.copy(), C)
    assert np.allclose(qx[:, :, :-1], -c * d43 * gx[:, :, :-1])
    qz = O.ScalarFluxesW_interior_WL(g, np.full_like(X, 2.0), Cz)
    assert np.allclose(qz[:, :, 1:-1], -d43 * cw[:, :, 1:-1] * 2.0)


def test_rhs_forcing_profiles_geostrophic_balance():
    """Main.py:167-188 forcing: with u = Ug(z) - Ugal and v = Vg(z) the
    Coriolis terms vanish for any (time-varying) geostrophic profile, so with
    no other terms RHS_u, RHS_v, RHS_TH are exactly the large-scale advection
    tendencies; without advection they are zero."""
    g, *_ = grid(optDamping=0, Ugal=3.0)
    rng = np.random.default_rng(5)
    Ug, Vg = 4.0 + rng.uniform(-1, 1, g.nz), rng.uniform(-1, 1, g.nz)
    Ua, Va, Ta = (rng.uniform(-1, 1, g.nz) for _ in range(3))
    shape = (g.nx, g.ny, g.nz)
    u = np.broadcast_to(Ug - g.Ugal, shape).copy()
    v = np.broadcast_to(Vg, shape).copy()
    z = np.zeros(shape)
    Ru, Rv, Rw = O.RHS_Momentum(g, u, v, z, z, z, z, z, z, z, z, Ug=Ug, Vg=Vg, Uadv=Ua, Vadv=Va)
    np.testing.assert_allclose(Ru, np.broadcast_to(Ua, shape), atol=1e-15)
    np.testing.assert_allclose(Rv, np.broadcast_to(Va, shape), atol=1e-15)
    assert not np.any(Rw)
    Ru0, Rv0, _ = O.RHS_Momentum(g, u, v, z, z, z, z, z, z, z, z, Ug=Ug, Vg=Vg)
    assert np.max(np.abs(Ru0)) < 1e-15 and np.max(np.abs(Rv0)) < 1e-15
    RT = O.RHS_Scalar(g, z, z, z, THadv=Ta)
    np.testing.assert_allclose(RT, np.broadcast_to(Ta, shape), atol=1e-15)


def test_moisture_terms_closed_forms():
    """AB2_Q with its top rule, the virtual-temperature buoyancy and the
    MOST moisture flux in the neutral limit (psi = 0: the log law)."""
    g, X, Y, Zu, Zw = grid(q_inversion=2e-6, optBuoyancy=0, zMoisture=0.0)
    rng = np.random.default_rng(4)
    shape = X.shape
    Q = 0.01 + 1e-3 * rng.standard_normal(shape)
    R, Rp = 0.3 * np.ones(shape), -0.1 * np.ones(shape)
    Qn = O.AB2_Q(g, Q, R, Rp)
    inc = g.dt_nondim * (1.5 * 0.3 + 0.5 * 0.1)
    np.testing.assert_allclose(Qn[:, :, :-1], Q[:, :, :-1] + inc, rtol=0, atol=1e-15)
    np.testing.assert_allclose(Qn[:, :, -1], Qn[:, :, -2] + g.q_inversion_nondim * g.dz,
                               rtol=0, atol=1e-15)
    # buoyancy: TH uniform, H = a cos x on every level -> THv' / T_0 = 0.61 a cos x
    a = 2e-3
    b = O.Buoyancy(g, np.zeros(shape), a * np.cos(X))
    # (THv = 300 (1 + 0.61 H) carries ~3e-14 of rounding, times g_nondim ~3e3)
    np.testing.assert_allclose(b[:, :, 1:], g.g_nondim * 0.61 * a * np.cos(X[:, :, 1:]),
                               rtol=1e-12, atol=1e-10)
    assert not np.any(b[:, :, 0])
    # surface moisture flux, neutral (psi_h = psi_h0 = 0), prescribed Q_s
    ustar = 0.3 + 0.1 * np.cos(X[:, :, 0])
    z2 = np.zeros((g.nx, g.ny))
    MOST = (z2, z2, z2, z2, 1 + z2, 1 + z2)
    for het in (0, 1):
        g2 = O.Grid(dict(g.cfg, optMoistureSurfBC=2, optSurfFlux=het))
        qm = O.SurfaceMoistureFlux(g2, Q, ustar, MOST, 0.015)
        q0 = Q[:, :, 0] if het else np.mean(Q[:, :, 0])
        lg = np.log(0.5 * g.dz * g.z_scale / g.z0T)
        np.testing.assert_allclose(qm, ustar * 0.4 * (0.015 - q0) / lg, rtol=1e-13)
    g3 = O.Grid(dict(g.cfg, optMoistureSurfBC=0, MoistureFlux=1.5e-4))
    assert np.all(O.SurfaceMoistureFlux(g3, Q, ustar, MOST) == 1.5e-4)


def test_moisture_step_keeps_uniform_q():
    """A uniform Q with no surface flux and no lapse stays uniform: every
    moisture term (gradients, advection, SGS flux, sponge) vanishes."""
    cfg, u, v, w, TH = cases.initial_fields("convective", nx=16, ny=16, nz=12)
    cfg.update(optMoisture=1, optMoistureSurfBC=0, MoistureFlux=0.0, q_inversion=0.0)
    g = O.Grid(cfg)
    s = O.initial_state(g, u, v, w, TH, np.full(u.shape, 0.008))
    th = O.ThomasPressureInit(g)
    ser = cases.surface_series(cfg, 4)
    for _ in range(2):
        s = O.step(g, s, th, None if ser is None else ser[s.iteration - 1])
    np.testing.assert_allclose(s.Q, 0.008, rtol=0, atol=1e-17)
