"""Closed-form checks of the oracle's statistics (ComputeStats,
src/utilities/Statistics.py:47-330) on the CPU.

Fields are single horizontal modes with per-level amplitudes, so every
plane mean, variance and covariance is known in closed form (the mean of
cos^2 over whole periods is 1/2, of cos*sin is 0): the expected values come
from the amplitudes, not from the oracle's formulas.  The staggered vertical
fluxes use the averages of neighbouring amplitudes that the w-node
interpolation implies, with level 0 taken as is.
"""
import numpy as np

import alfa_oracle as O
from alfa_amd import cases


def setup(Ugal=5.0, nx=16, ny=12, nz=9, seed=3):
    cfg, *_ = cases.initial_fields("stable", nx=nx, ny=ny, nz=nz, Ugal=Ugal)
    g = O.Grid(cfg)
    rng = np.random.default_rng(seed)
    x = np.arange(g.nx) * 2 * np.pi / g.nx
    y = np.arange(g.ny) * 2 * np.pi / g.ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    cx, sy, c2x = np.cos(X)[..., None], np.sin(Y)[..., None], np.cos(2 * X)[..., None]
    amp = {k: rng.uniform(0.5, 2.0, g.nz) for k in "ABDabcdef"}
    return g, amp, cx, sy, c2x


def fields(g, amp, cx, sy, c2x):
    A, B, D = amp["A"], amp["B"], amp["D"]
    a, b, c, d = amp["a"], amp["b"], amp["c"], amp["d"]
    u = A + a * cx            # stored u (Ugal removed)
    v = B + b * cx + amp["e"] * c2x
    w = c * cx + amp["f"] * sy
    TH = D + d * cx
    return u, v, w, TH


def test_means_variances_and_level_fluxes():
    g, amp, cx, sy, c2x = setup()
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    z = np.zeros_like(u)
    acc = O.ComputeStats(g, O.InitializeStats(g), u, v, w, TH, z, z, z,
                         np.ones((g.nx, g.ny)), np.ones((g.nx, g.ny)), 0.0, z, z, z, z,
                         None, None, None, None, None)
    a, b, c, d, e, f = (amp[k] for k in "abcdef")
    np.testing.assert_allclose(acc["U"], amp["A"] + g.Ugal, rtol=1e-13)
    np.testing.assert_allclose(acc["V"], amp["B"], rtol=1e-13)
    np.testing.assert_allclose(acc["W"], 0.0, atol=1e-13)
    np.testing.assert_allclose(acc["TH"], amp["D"] + g.T_0_nondim, rtol=1e-13)
    np.testing.assert_allclose(acc["u2"], a * a / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["v2"], (b * b + e * e) / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["w2"], (c * c + f * f) / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["TH2"], d * d / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["uv"], a * b / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["uTH"], a * d / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["vTH"], b * d / 2, rtol=1e-12)


def test_staggered_vertical_fluxes():
    g, amp, cx, sy, c2x = setup()
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    z = np.zeros_like(u)
    acc = O.ComputeStats(g, O.InitializeStats(g), u, v, w, TH, z, z, z,
                         np.ones((g.nx, g.ny)), np.ones((g.nx, g.ny)), 0.0, z, z, z, z,
                         None, None, None, None, None)
    a, b, c, d = (amp[k] for k in "abcd")

    def wnode(p):   # amplitude on w nodes: level 0 as is, else the mean of k-1, k
        q = p.copy()
        q[1:] = 0.5 * (p[:-1] + p[1:])
        return q

    np.testing.assert_allclose(acc["uw"], wnode(a) * c / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["vw"], wnode(b) * c / 2, rtol=1e-12)
    np.testing.assert_allclose(acc["wTH"], wnode(d) * c / 2, rtol=1e-12)


def test_fluctuation_statistics_ignore_level_constants():
    g, amp, cx, sy, c2x = setup()
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    z = np.zeros_like(u)
    args = (z, z, z, np.ones((g.nx, g.ny)), np.ones((g.nx, g.ny)), 0.0, z, z, z, z,
            None, None, None, None, None)
    base = O.ComputeStats(g, O.InitializeStats(g), u, v, w, TH, *args)
    shift = np.linspace(-3, 3, g.nz)
    moved = O.ComputeStats(g, O.InitializeStats(g), u + shift, v - 2 * shift, w, TH + shift,
                           *args)
    for k in ("u2", "v2", "w2", "TH2", "uv", "uTH", "vTH", "uw", "vw", "wTH"):
        np.testing.assert_allclose(moved[k], base[k], rtol=1e-10, atol=1e-13, err_msg=k)
    np.testing.assert_allclose(moved["U"], base["U"] + shift, rtol=1e-13)


def test_gradients_surface_sgs_and_coefficients():
    g, amp, cx, sy, c2x = setup(Ugal=0.0)
    rng = np.random.default_rng(9)
    shape = (g.nx, g.ny, g.nz)
    lin = np.broadcast_to(np.arange(g.nz, dtype=float), shape)
    dudz = 2.0 * lin + cx          # plane mean 2k
    dvdz = -lin + 0 * cx
    dTHdz = 0.5 + 0 * lin
    M = 3.0 + np.cos(np.arange(g.nx) * 2 * np.pi / g.nx)[:, None] * np.ones((1, g.ny))
    ustar = np.full((g.nx, g.ny), 0.3)
    ustar[::2] = 0.4                # mean of squares (0.09 + 0.16) / 2
    txy, txz, tyz, qz = (k * lin + cx for k in (1.0, 2.0, 3.0, 4.0))
    profs = [rng.uniform(0, 1, g.nz) for _ in range(5)]
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    acc = O.ComputeStats(g, O.InitializeStats(g), u, v, w, TH, dudz, dvdz, dTHdz, M, ustar,
                         0.12, txy, txz, tyz, qz, *profs)
    k = np.arange(g.nz)
    np.testing.assert_allclose(acc["dUdz"], 2.0 * k / g.z_scale, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(acc["dVdz"], -k / g.z_scale, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(acc["dTHdz"], 0.5 / g.z_scale, rtol=1e-13)
    assert abs(acc["M_sfc"] - 3.0) < 1e-13
    assert abs(acc["ustar"] - np.sqrt((0.09 + 0.16) / 2)) < 1e-14
    assert acc["qz_sfc"] == 0.12
    for key, s in (("txy", 1.0), ("txz", 2.0), ("tyz", 3.0), ("qz", 4.0)):
        np.testing.assert_allclose(acc[key], s * k, rtol=1e-13, atol=1e-13)
    for key, p in zip(("Cs2_1", "Cs2_2", "Cs2PrRatio", "Beta1", "Beta2"), profs):
        np.testing.assert_array_equal(acc[key], p)


def test_accumulate_then_average():
    g, amp, cx, sy, c2x = setup()
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    z = np.zeros_like(u)
    args = (z, z, z, np.ones((g.nx, g.ny)), np.ones((g.nx, g.ny)), 0.1, z, z, z, z,
            None, None, None, None, None)
    acc = O.InitializeStats(g)
    acc = O.ComputeStats(g, acc, u, v, w, TH, *args)
    acc = O.ComputeStats(g, acc, 3 * u, v, w, TH, *args)
    avg = O.AverageStats(acc, 2)
    a = amp["a"]
    # u2 of 3u is 9 a^2 / 2: the average of the two samples is 5 a^2 / 2
    np.testing.assert_allclose(avg["u2"], 5 * a * a / 2, rtol=1e-12)
    assert abs(avg["qz_sfc"] - 0.1) < 1e-15
    assert avg["Ugal"] == g.Ugal and np.all(avg["ZeRo1D"] == 0)
    assert set(O.STATS_MOIST_PROFILES) <= set(avg) and avg["qm_sfc"] == 0.0


def test_moisture_statistics():
    g, amp, cx, sy, c2x = setup()
    u, v, w, TH = fields(g, amp, cx, sy, c2x)
    c, d = amp["c"], amp["d"]
    Q = 0.01 + 1e-3 * d * cx          # same mode as w's cos part
    z = np.zeros_like(u)
    k = np.arange(g.nz, dtype=float)
    acc = O.ComputeStats(g, O.InitializeStats(g), u, v, w, TH, z, z, z,
                         np.ones((g.nx, g.ny)), np.ones((g.nx, g.ny)), 0.0, z, z, z, z,
                         None, None, None, None, None, Q=Q, dQdz=k + 0 * u, qHz=2 * k + 0 * u,
                         qm_sfc_avg=3e-5)
    a = 1e-3 * d
    wn = a.copy()
    wn[1:] = 0.5 * (a[:-1] + a[1:])
    np.testing.assert_allclose(acc["Q"], 0.01, rtol=1e-13)
    np.testing.assert_allclose(acc["Q2"], a * a / 2, rtol=1e-10)
    np.testing.assert_allclose(acc["wQ"], wn * c / 2, rtol=1e-10)
    np.testing.assert_allclose(acc["dQdz"], k / g.z_scale, rtol=1e-13)
    np.testing.assert_allclose(acc["qHz"], 2 * k, rtol=1e-13)
    assert acc["qm_sfc"] == 3e-5
