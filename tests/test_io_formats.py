"""The reference's run-input formats (alfa_amd.io), on the CPU.

vel.npy / TH.npy are Fortran-order flattenings (first index fastest:
element (i, j, k) at i + nx*(j + ny*k)), checked element by element from
that definition; SurfaceBC.npz round-trips and is validated as
Initialization.py:107-160 validates it.
"""
import numpy as np
import pytest

from alfa_amd import cases, io


def test_initial_fields_layout_and_round_trip(tmp_path):
    cfg, u, v, w, TH = cases.initial_fields("stable", nx=6, ny=5, nz=4, seed=2, Ugal=5.0)
    Ugal, T0 = cfg["Ugal"], cfg["T_0"]
    io.write_initial_fields(tmp_path, u + Ugal, v, w, TH + T0)
    vel = np.load(tmp_path / "vel.npy")
    th = np.load(tmp_path / "TH.npy")
    nx, ny, nz = u.shape
    assert vel.shape == (nx * ny * nz, 3) and th.shape == (nx * ny * nz,)
    for (i, j, k) in [(0, 0, 0), (5, 0, 0), (0, 4, 0), (2, 3, 1), (5, 4, 3)]:
        n = i + nx * (j + ny * k)
        assert vel[n, 0] == u[i, j, k] + Ugal
        assert vel[n, 1] == v[i, j, k] and vel[n, 2] == w[i, j, k]
        assert th[n] == TH[i, j, k] + T0
    ru, rv, rw, rth = io.read_initial_fields(tmp_path, cfg)
    np.testing.assert_allclose(ru, u, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(rv, v)
    np.testing.assert_array_equal(rw, w)
    np.testing.assert_allclose(rth, TH, rtol=0, atol=1e-12)


def test_initial_fields_shape_checks(tmp_path):
    cfg, u, v, w, TH = cases.initial_fields("neutral", nx=4, ny=4, nz=4)
    io.write_initial_fields(tmp_path, u, v, w, TH)
    bad = dict(cfg, nz=5)
    with pytest.raises(ValueError):
        io.read_initial_fields(tmp_path, bad)


def test_surface_bc_round_trip_and_validation(tmp_path):
    cfg, *_ = cases.initial_fields("stable", nx=8, ny=8, nz=8)
    cfg = dict(cfg, SimTime=1.0, dt=0.05)
    n = io.nsteps_of(cfg)
    assert n == 20                              # ceil(1.0 / 0.05)
    t = np.arange(n + 1) * cfg["dt"]
    T_s = cfg["T_0"] - 0.25 * t / 3600.0         # the GABLS1 linear cooling
    p = tmp_path / "SurfaceBC.npz"
    io.write_surface_bc(p, T_s, cfg["dt"], 2)
    with np.load(p) as f:
        assert set(f.keys()) == {"data_series", "t_series", "dt_surf", "optSurfBC"}
        np.testing.assert_array_equal(f["t_series"], t)
    got = io.read_surface_bc(p, cfg)
    np.testing.assert_allclose(got, cases.surface_series(cfg, n), rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        io.read_surface_bc(p, dict(cfg, dt=0.1, SimTime=2.0))      # dt mismatch
    with pytest.raises(ValueError):
        io.read_surface_bc(p, dict(cfg, optSurfBC=1))              # mode mismatch
    with pytest.raises(ValueError):
        io.read_surface_bc(p, dict(cfg, SimTime=2.0))              # length mismatch
    # heat flux series: returned as is (u_scale = TH_scale = 1)
    q = np.full(n + 1, 0.24)
    io.write_surface_bc(p, q, cfg["dt"], 1)
    np.testing.assert_array_equal(io.read_surface_bc(p, dict(cfg, optSurfBC=1)), q)


def test_geowind_series_and_compact_profiles(tmp_path):
    cfg, *_ = cases.initial_fields("convective", nx=8, ny=8, nz=6)
    cfg = dict(cfg, SimTime=2.0, dt=0.5, optGeoWind=1)
    n = io.nsteps_of(cfg) + 1
    nz = cfg["nz"]
    t = np.arange(n) * cfg["dt"]
    z = np.arange(nz)
    Ug = 5.0 + 0.1 * t[:, None] + 0.01 * z[None, :]
    Vg = -1.0 + 0.2 * t[:, None] * np.ones((1, nz))
    p = tmp_path / "GeoWind.npz"
    np.savez(p, Ug_series=Ug, Vg_series=Vg, dt_geo=cfg["dt"], optGeoWind=np.int32(1))
    a, b = io.read_geowind(p, cfg)
    np.testing.assert_array_equal(a, Ug)
    np.testing.assert_array_equal(b, Vg)
    # compact form: profiles at t = 0, 1.5, 3 s, linear in time, so the
    # interpolation onto t = i dt reproduces the linear law exactly
    tp = np.array([0.0, 1.5, 3.0])
    Up = 5.0 + 0.1 * tp[:, None] + 0.01 * z[None, :]
    Vp = -1.0 + 0.2 * tp[:, None] * np.ones((1, nz))
    np.savez(p, t_profile=tp, Ug_profile=Up, Vg_profile=Vp, dt_geo=cfg["dt"],
             optGeoWind=np.int32(1))
    a, b = io.read_geowind(p, cfg)
    np.testing.assert_allclose(a, Ug, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b, Vg, rtol=0, atol=1e-13)
    with pytest.raises(ValueError):   # profiles ending before SimTime
        np.savez(p, t_profile=tp[:2], Ug_profile=Up[:2], Vg_profile=Vp[:2], dt_geo=cfg["dt"],
                 optGeoWind=np.int32(1))
        io.read_geowind(p, cfg)


def test_adv_forcing_scaling_and_defaults(tmp_path):
    cfg, *_ = cases.initial_fields("convective", nx=8, ny=8, nz=6)
    cfg = dict(cfg, SimTime=1.0, dt=0.5, optAdvection=1)
    n, nz = io.nsteps_of(cfg) + 1, cfg["nz"]
    U = np.full((n, nz), 1e-4)
    V = np.full((n, nz), -2e-4)
    p = tmp_path / "AdvForcing.npz"
    np.savez(p, Uadv_series=U, Vadv_series=V, dt_adv=cfg["dt"], optAdvection=np.int32(1))
    a, b, c = io.read_adv_forcing(p, cfg)
    zs = cfg["l_x"] / (2 * np.pi)            # tendencies [m/s^2] times z_scale
    np.testing.assert_allclose(a, 1e-4 * zs, rtol=1e-15)
    np.testing.assert_allclose(b, -2e-4 * zs, rtol=1e-15)
    assert c.shape == (n, nz) and not np.any(c)
    with pytest.raises(ValueError):
        io.read_adv_forcing(p, dict(cfg, optAdvection=0))


def test_moisture_inputs(tmp_path):
    cfg, u, *_ = cases.initial_fields("convective", nx=6, ny=4, nz=5)
    cfg = dict(cfg, SimTime=1.0, dt=0.5, optMoistureSurfBC=2)
    Q = 0.008 + 1e-4 * np.arange(u.size).reshape(u.shape)
    io.write_moisture(tmp_path, Q)
    q = np.load(tmp_path / "Q.npy")
    nx, ny, nz = u.shape
    i, j, k = 5, 3, 4
    assert q[i + nx * (j + ny * k)] == Q[i, j, k]
    np.testing.assert_array_equal(io.read_moisture(tmp_path, cfg), Q)
    n = io.nsteps_of(cfg) + 1
    p = tmp_path / "MoistureSurfaceBC.npz"
    ser = np.linspace(0.009, 0.0085, n)
    np.savez(p, data_series=ser, dt_moist=cfg["dt"], optMoistureSurfBC=np.int32(2))
    np.testing.assert_array_equal(io.read_moisture_surface_bc(p, cfg), ser)
    with pytest.raises(ValueError):
        io.read_moisture_surface_bc(p, dict(cfg, optMoistureSurfBC=1))
