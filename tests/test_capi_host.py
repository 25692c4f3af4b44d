"""C-ABI library checks that need no GPU: the in-tree libgamil_dycore.so loads
and exports every function declared in include/gamil_dycore.h, and the
host-side initial-condition generators behave as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gamil_dycore.h")
G = 9.80616
A = 6.37122e6
OMEGA = 2 * np.pi / 86400.0


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gd):
    lib = ctypes.CDLL(gd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(gd.EXPORTS)


def test_balanced_ic_properties(gd):
    N, M = 180, 91
    u, v, gdf, ghs = gd.ic_balanced(N, M, seed=11, nmax=20, rms_height=150.0, mean_depth=8000.0)
    assert np.all(u[0] == 0) and np.all(u[-1] == 0)
    assert np.all(ghs == 0)
    lat = -0.5 * np.pi + np.arange(M) * np.pi / (M - 1)
    w = np.cos(lat); w[0] = w[-1] = 0.0
    h = gdf / G - 8000.0
    mean = (w[:, None] * h).sum() / (w.sum() * N)
    rms = np.sqrt((w[:, None] * (h - mean) ** 2).sum() / (w.sum() * N))
    assert abs(mean) < 1e-9 and abs(rms - 150.0) < 1e-9
    # geostrophic balance f u = -(g/a) dh/dlat at mid latitudes (2nd-order differences)
    dlat = np.pi / (M - 1)
    j = np.arange(20, 40)
    dh = (h[j + 1] - h[j - 1]) / (2 * dlat)
    hb = 0.5 * (dh + np.roll(dh, -1, axis=1))  # to the half-lon u points
    f = 2 * OMEGA * np.sin(lat[j])[:, None]
    resid = f * u[j] + G / A * hb
    assert np.abs(resid).max() < 2e-2 * np.abs(f * u[j]).max()


def test_balanced_ic_is_seeded_and_deterministic(gd):
    a = gd.ic_balanced(72, 37, seed=1, nmax=8)
    b = gd.ic_balanced(72, 37, seed=1, nmax=8)
    c = gd.ic_balanced(72, 37, seed=2, nmax=8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.abs(a[2] - c[2]).max() > 1.0


def test_mountains(gd):
    N, M = 180, 91
    u, v, gdf, ghs = gd.ic_balanced(N, M, seed=3, nmax=10, n_mountains=3, mountain_seed=5)
    hs = ghs / G
    assert 500.0 <= hs.max() <= 2000.0 * 1.2
    assert hs.min() >= 0.0
    # the free surface gd + ghs does not see the mountains
    u2, v2, gd2, ghs2 = gd.ic_balanced(N, M, seed=3, nmax=10)
    np.testing.assert_allclose(gdf + ghs, gd2 + ghs2, rtol=1e-14)


def test_rossby_haurwitz_ic_closed_form(gd):
    """rossby_haurwitz_wave_test_mod.F90:33-85 against the numpy closed form."""
    from test_oracle import rh_state
    N, M = 96, 49
    got = gd.ic_rossby_haurwitz(N, M)
    ref = rh_state(N, M)
    for x, y in zip(got, ref):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-9)


def test_mountain_zonal_flow_ic(gd):
    """mountain_zonal_flow_test_mod.F90:28-97: cone of 2000 m at (270E, 30N),
    u = 20 cos(lat), balanced depth."""
    N, M = 180, 91
    u, v, gdf, ghs = gd.ic_mountain_zonal_flow(N, M)
    assert np.isclose(ghs.max() / G, 2000.0 * (1 - 0.0), rtol=0.05)
    lat = -0.5 * np.pi + np.arange(M) * np.pi / (M - 1)
    np.testing.assert_allclose(u[1:-1, 0], 20.0 * np.cos(lat[1:-1]), rtol=1e-13, atol=1e-13)
    assert np.all(v == 0) and np.all(u[0] == 0)


def test_cutoff_table_from_lat0(gd):
    cut = gd.cutoff_from_lat0(720, 361, 75.0)
    assert len(cut) == 29  # rows at 89.5 .. 75.5 deg from each pole
    assert cut[0] < cut[-1]  # cutoff rises away from the pole
    assert all(1 <= c <= 359 for c in cut)
    lat = np.deg2rad(-90 + 0.5)
    assert cut[0] == int(np.floor(360 * np.cos(lat) / np.cos(np.deg2rad(75.0))))


def test_serve_rejects_bad_requests(gd):
    """Argument checks of gd_serve run before any device call."""
    N, M = 72, 37
    ic = gd.ic_balanced(N, M, seed=1, nmax=4)
    ok = dict(num_lon=N, num_lat=M, time_step_size=600.0, nsteps=3, arrival_s=0.0, state=ic)
    with pytest.raises(gd.DycoreError, match="bad request 0"):
        gd.serve([dict(ok, time_step_size=0.0)])
    with pytest.raises(gd.DycoreError, match="bad request 1"):
        gd.serve([ok, dict(ok, arrival_s=-1.0)])
    with pytest.raises(gd.DycoreError, match="max_slots"):
        gd.serve([ok], max_slots=0)
    res, wall = gd.serve([])
    assert res == [] and wall >= 0


def test_create_rejects_bad_configs_before_any_device_call(gd):
    """gd_create validates the configuration before it touches a device:
    isp needs subcycles >= 1 like csp2 (isp_splitting divides dt and the
    accumulator scale by it, dycore_mod.F90:689-752)."""
    with pytest.raises(gd.DycoreError, match="subcycles") as e:
        gd.Dycore(72, 37, 600.0, split_scheme="isp", subcycles=0)
    assert e.value.code == gd.GD_ERR_INVALID_ARG
    with pytest.raises(gd.DycoreError, match="subcycles"):
        gd.Dycore(72, 37, 600.0, split_scheme="csp2", subcycles=-1)
