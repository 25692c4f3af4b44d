"""Pins of the FP64 oracle (CPU only).

Each check is chosen so a plausible slip in the restatement (a dropped term,
a wrong sign, index or weight) fails it without retyping the oracle's own
formulas:
  * filter: the reference's own FFTPACK 5.1 (golden vectors committed in
    tests/golden, and the live library in oracle/_ref when present), plus
    brute-force projection properties on tiny rows;
  * mesh/coef: closed forms and the pole-cos reset;
  * operators: the four antisymmetry identities of check_antisymmetry
    (dycore_mod.F90:794-851), the exact global mass budget of the flux form
    including the pole sums, rest and zonal-symmetry special cases;
  * time integration: exact quadratic-energy conservation of the
    beta-corrected predict-correct step (no filter, no diffusion);
  * diffusion: annihilates constants, conserves mass exactly, and reproduces
    the spherical-harmonic eigenvalue -n(n+1)/a^2 to O(dlat^2);
  * diag: total mass of a constant depth approaches 4 pi a^2 gd.
"""
import os

import numpy as np
import pytest

from conftest import rel_err

A = 6.37122e6
OMEGA = 2 * np.pi / 86400.0
G = 9.80616
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fftpack_filter.npz")


def rh_state(N, M):
    """Rossby-Haurwitz R=4 state from the closed form of
    rossby_haurwitz_wave_test_mod.F90:22-30 (numpy, independent of the C code)."""
    R, omg, gd0 = 4.0, 7.848e-6, 8.0e3 * G
    dlon, dlat = 2 * np.pi / N, np.pi / (M - 1)
    flat = -0.5 * np.pi + np.arange(M) * dlat
    hlat = flat[:-1] + 0.5 * dlat
    flon = np.arange(N) * dlon
    hlon = flon + 0.5 * dlon
    fc = np.cos(flat); fc[0] = fc[-1] = 0.0
    fs = np.sin(flat); fs[0], fs[-1] = -1.0, 1.0
    c, s = fc[:, None], fs[:, None]
    u = A * omg * (c + R * c ** (R - 1) * s ** 2 * np.cos(R * hlon) - c ** (R + 1) * np.cos(R * hlon))
    hc, hs = np.cos(hlat)[:, None], np.sin(hlat)[:, None]
    v = -A * omg * R * hc ** (R - 1) * hs * np.sin(R * flon)
    Af = 0.5 * omg * (2 * OMEGA + omg) * c ** 2 + 0.25 * omg ** 2 * (
        (R + 1) * c ** (2 * R + 2) + (2 * R ** 2 - R - 2) * c ** (2 * R) - 2 * R ** 2 * c ** (2 * R - 2))
    Bf = 2 * (OMEGA + omg) * omg * c ** R * (R ** 2 + 2 * R + 2 - (R + 1) ** 2 * c ** 2) / (R + 1) / (R + 2)
    Cf = 0.25 * omg ** 2 * c ** (2 * R) * ((R + 1) * c ** 2 - R - 2)
    gd = gd0 + A ** 2 * (Af + Bf * np.cos(R * flon) + Cf * np.cos(2 * R * flon))
    return u, v, gd, np.zeros_like(gd)


def smooth_state(N, M, seed=0):
    """Random smooth state (few zonal/meridional waves), u = 0 on the poles."""
    rng = np.random.default_rng(seed)
    dlon, dlat = 2 * np.pi / N, np.pi / (M - 1)
    lat = -0.5 * np.pi + np.arange(M) * dlat
    hlat = lat[:-1] + 0.5 * dlat
    lon = np.arange(N) * dlon
    def field(lats, lons, amp):
        f = np.zeros((len(lats), len(lons)))
        for _ in range(6):
            k, l = rng.integers(0, 5), rng.integers(1, 5)
            ph = rng.uniform(0, 2 * np.pi)
            f += amp * rng.standard_normal() * np.cos(k * lons[None, :] + ph) * np.cos(lats[:, None]) ** 2 * np.cos(l * lats[:, None])
        return f
    u = field(lat, lon + 0.5 * dlon, 15.0); u[0] = u[-1] = 0.0
    v = field(hlat, lon, 8.0)
    gd = G * (8000.0 + field(lat, lon, 120.0))
    ghs = G * np.abs(field(lat, lon, 300.0))
    return u, v, gd, ghs


# ---------------------------------------------------------------------------
# filter
# ---------------------------------------------------------------------------

def test_filter_matches_reference_fftpack_golden(oracle_mod):
    d = np.load(GOLDEN)
    for n in (12, 16, 24, 90, 180, 360, 720):
        x, cuts, ys = d[f"x_{n}"], d[f"cuts_{n}"], d[f"y_{n}"]
        # one oracle per cutoff: the row filtered is full row 2 (cutoff entry 1)
        for c, y in zip(cuts, ys):
            if c == 0:
                continue
            o = oracle_mod.Oracle(n, 9, 100.0, cutoff=[int(c)])
            assert rel_err(o.filter_full(2, x), y) < 1e-13, (n, c)


def test_filter_matches_live_reference_fftpack(oracle_mod):
    if not oracle_mod.ref_fftpack_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    for n in (36, 144, 1440):
        fft = oracle_mod.RefFFTPACK(n)
        for c in (1, 5, n // 6, n // 2 - 1):
            o = oracle_mod.Oracle(n, 9, 100.0, cutoff=[c])
            x = rng.standard_normal(n)
            mask = np.zeros(n); mask[: 2 * c + 2] = 1
            assert rel_err(o.filter_full(2, x), fft.filter(x, mask)) < 1e-13


def test_fftpack_layout_golden():
    """FFTPACK5 rfft1f output: [a0, a1, b1, a2, b2, ...] scaled 1/N and 2/N
    (rfftf1.f:87-108), a_{N/2} scaled 1/N."""
    d = np.load(GOLDEN)
    for n in (12, 16, 24, 90):
        x, c = d[f"x_{n}"], d[f"coef_{n}"]
        t = 2 * np.pi * np.arange(n) / n
        assert abs(c[0] - x.mean()) < 1e-14
        for k in range(1, (n - 1) // 2 + 1):
            assert abs(c[2 * k - 1] - 2.0 / n * (x * np.cos(k * t)).sum()) < 1e-13
            assert abs(c[2 * k] - 2.0 / n * (x * np.sin(k * t)).sum()) < 1e-13
        assert abs(c[n - 1] - (x * np.cos(n // 2 * t)).sum() / n) < 1e-13


@pytest.mark.parametrize("n", [8, 12, 15, 16])
def test_filter_projection_brute_force(oracle_mod, n):
    """Brute force on tiny rows: the filter matrix is a symmetric idempotent
    projector that keeps constants and cos((c+1) x) but removes sin((c+1) x)."""
    for c in range(1, (n - 2) // 2 + 1):
        o = oracle_mod.Oracle(n, 9, 100.0, cutoff=[c])
        P = np.array([o.filter_full(2, e) for e in np.eye(n)]).T
        assert np.abs(P - P.T).max() < 1e-14
        assert np.abs(P @ P - P).max() < 1e-13
        t = 2 * np.pi * np.arange(n) / n
        assert np.abs(P @ np.ones(n) - 1).max() < 1e-14
        if 2 * (c + 1) < n:
            assert np.abs(P @ np.cos((c + 1) * t) - np.cos((c + 1) * t)).max() < 1e-13
            assert np.abs(P @ np.sin((c + 1) * t)).max() < 1e-13
            if c + 2 < n / 2:
                assert np.abs(P @ np.cos((c + 2) * t)).max() < 1e-13
        assert np.isclose(np.trace(P), min(2 * c + 2, n))


def test_filter_table_quirks(oracle_mod):
    """filter_mod.F90:44-99: K nonzero entries flag K full rows per hemisphere,
    K half rows in the south but K-1 in the north, where each flagged north
    half row uses the mask of the row on its equatorward side (entry j-1)."""
    N, M = 48, 25
    o = oracle_mod.Oracle(N, M, 100.0, cutoff=[5, 3, 2])
    ff, hf, fc, hc = o.filter_table()
    assert list(np.nonzero(ff)[0]) == [1, 2, 3, M - 4, M - 3, M - 2]
    assert list(fc[[1, 2, 3]]) == [5, 3, 2] and list(fc[[M - 2, M - 3, M - 4]]) == [5, 3, 2]
    assert list(np.nonzero(hf)[0]) == [0, 1, 2, M - 3, M - 2]
    assert list(hc[[0, 1, 2]]) == [5, 3, 2]
    # north: half row M-1 (1-based) flagged by entry 2, masked by entry 1
    assert list(hc[[M - 2, M - 3, M - 4]]) == [5, 3, 2]
    assert hf[M - 4] == 0


# ---------------------------------------------------------------------------
# mesh / coefficients
# ---------------------------------------------------------------------------

def test_mesh_and_coefficients_closed_form(oracle_mod):
    N, M = 40, 21
    o = oracle_mod.Oracle(N, M, 100.0)
    dlat = np.pi / (M - 1)
    lat = -0.5 * np.pi + np.arange(M) * dlat
    assert np.allclose(o.coef("full_lat"), lat, atol=1e-15)
    assert np.allclose(o.coef("half_lat"), lat[:-1] + 0.5 * dlat, atol=1e-15)
    fc = o.coef("full_cos")
    assert fc[0] == 0.0 and fc[-1] == 0.0
    assert np.allclose(fc[1:-1], np.cos(lat[1:-1]), atol=1e-15)
    assert np.allclose(o.coef("full_f"), 2 * OMEGA * np.sin(lat), atol=1e-18)
    c = o.coef("full_c")
    assert c[0] == 0 and c[-1] == 0
    assert np.allclose(c[1:-1], np.tan(lat[1:-1]) / A, rtol=1e-14)
    assert np.allclose(o.coef("half_dlon"), A * 2 * np.pi / N * np.cos(lat[:-1] + 0.5 * dlat), rtol=1e-14)
    # reset_cos_lat_at_poles (dycore_mod.F90:159-173), applied by dycore_run
    o.reset_cos_lat_at_poles()
    fc2 = o.coef("full_cos")
    hc = o.coef("half_cos")
    assert fc2[0] == hc[0] * 0.25 and fc2[-1] == hc[-1] * 0.25
    assert np.isclose(o.coef("full_dlat")[0], A * dlat * hc[0] / 4, rtol=1e-15)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------

def _ready(oracle_mod, N, M, state, **kw):
    o = oracle_mod.Oracle(N, M, 600.0, **kw)
    o.set_initial_state(*state)
    o.start()
    return o


def test_antisymmetry_identities(oracle_mod):
    """check_antisymmetry (dycore_mod.F90:794-851): with the center scheme
    advection is energy-neutral, and PGF/mass divergence and the Coriolis
    pair cancel in the IAP inner product."""
    N, M = 72, 37
    o = _ready(oracle_mod, N, M, smooth_state(N, M, 3), use_zonal_tend_filter=False)
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    sums = o.check_antisymmetry(0, o.old)
    # scale: magnitude of the individual products
    T, S = oracle_mod, o
    U = o.state(o.old, T.IAP_U); V = o.state(o.old, T.IAP_V)
    gw = o.state(o.old, T.GD)
    fc = o.coef("full_cos")  # after the pole reset done by start()
    hc = o.coef("half_cos")
    def mag(f, w, cos):
        return np.abs(o.tend(0, f) * w * cos[:, None]).sum()
    scale = [mag(T.T_U_ADV_LON, U, fc) + mag(T.T_V_ADV_LAT, V, hc),
             mag(T.T_U_PGF, U, fc) + mag(T.T_MASS_DIV_LON, gw, fc),
             mag(T.T_FV, U, fc) + mag(T.T_FU, V, hc),
             mag(T.T_V_PGF, V, hc) + mag(T.T_MASS_DIV_LAT, gw, fc)]
    for k in range(4):
        assert abs(sums[k]) < 1e-12 * scale[k], (k, sums[k], scale[k])


def test_mass_budget_closes_with_pole_sums(oracle_mod):
    """sum_j cos_j sum_i dgd = 0: the flux form telescopes and the pole-row
    sums with the reset pole cos (dycore_mod.F90:159-173, 572-596) close it."""
    N, M = 64, 33
    o = _ready(oracle_mod, N, M, smooth_state(N, M, 5), use_zonal_tend_filter=False)
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    dgd = o.tend(0, oracle_mod.T_DGD)
    fc = o.coef("full_cos")
    total = (fc[:, None] * dgd).sum()
    assert abs(total) < 1e-14 * (fc[:, None] * np.abs(dgd)).sum()
    # without the pole treatment the budget would not close
    assert abs((fc[1:-1, None] * dgd[1:-1]).sum()) > 1e3 * abs(total)


def test_rest_state_has_no_tendency(oracle_mod):
    N, M = 36, 19
    u = np.zeros((M, N)); v = np.zeros((M - 1, N))
    gd = np.full((M, N), G * 8000.0); ghs = np.zeros((M, N))
    o = _ready(oracle_mod, N, M, (u, v, gd, ghs))
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    for f in (oracle_mod.T_DU, oracle_mod.T_DV, oracle_mod.T_DGD):
        assert np.all(o.tend(0, f) == 0.0)


def test_zonally_symmetric_state(oracle_mod):
    """u(lat) only, v = 0, gd(lat): zonal advection, zonal PGF and zonal mass
    divergence vanish exactly."""
    N, M = 48, 25
    lat = -0.5 * np.pi + np.arange(M) * np.pi / (M - 1)
    u = np.tile((20 * np.cos(lat) ** 2)[:, None], (1, N)); u[0] = u[-1] = 0
    v = np.zeros((M - 1, N))
    gd = np.tile((G * (8000 - 200 * np.sin(lat) ** 2))[:, None], (1, N))
    o = _ready(oracle_mod, N, M, (u, v, gd, np.zeros((M, N))))
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    for f in (oracle_mod.T_U_ADV_LON, oracle_mod.T_U_PGF, oracle_mod.T_MASS_DIV_LON,
              oracle_mod.T_V_ADV_LON, oracle_mod.T_V_ADV_LAT, oracle_mod.T_MASS_DIV_LAT):
        assert np.abs(o.tend(0, f)).max() < 1e-20, f


def test_iap_round_trip(oracle_mod):
    """u -> U = (G_i+G_i+1)/2 u -> u' = 2U/(G_i+G_i+1) is the identity to 1 ulp."""
    N, M = 40, 21
    st = smooth_state(N, M, 8)
    o = _ready(oracle_mod, N, M, st)
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    o.update_state(0.0, 0, o.old, o.new)
    u0, u1 = o.state(o.old, oracle_mod.U), o.state(o.new, oracle_mod.U)
    assert np.all(np.abs(u1 - u0) <= 2.3e-16 * np.abs(u0) + 1e-300)


@pytest.mark.parametrize("split", ["none", "csp2"])
def test_predict_correct_conserves_energy_exactly(oracle_mod, split):
    """qcon_modified beta makes the IAP quadratic energy exactly conserved
    (no filter, no diffusion): drift at round-off over 20 steps; mass also."""
    N, M = 72, 37
    o = oracle_mod.Oracle(N, M, 300.0, split_scheme=split, subcycles=3,
                          use_zonal_tend_filter=False)
    o.set_initial_state(*rh_state(N, M))
    m0, e0 = o.start()
    mass, energy, beta, ns = o.run(20)
    assert ns == 0
    assert np.abs(energy - e0).max() < 1e-13 * e0
    assert np.abs(mass - m0).max() < 1e-13 * m0
    # without the correction energy drifts much more
    o2 = oracle_mod.Oracle(N, M, 300.0, split_scheme=split, subcycles=3, qcon_modified=False,
                           use_zonal_tend_filter=False)
    o2.set_initial_state(*rh_state(N, M))
    o2.start()
    _, en2, b2, ns2 = o2.run(20)
    n_ok = ns2 - 1 if ns2 else 20  # the uncorrected unsplit run may blow up
    assert np.all(b2[:n_ok] == 1.0)
    assert ns2 or np.abs(en2 - e0).max() > 1e3 * np.abs(energy - e0).max()


# ---------------------------------------------------------------------------
# diffusion and diag
# ---------------------------------------------------------------------------

def test_diffusion_constant_and_mass(oracle_mod):
    N, M = 48, 25
    u = np.zeros((M, N)); v = np.zeros((M - 1, N))
    o = _ready(oracle_mod, N, M, (u, v, np.full((M, N), 7.0e4), np.zeros((M, N))),
               use_diffusion=True, diffusion_coef=1e6)
    o.ordinary_diffusion(600.0, o.old)
    assert np.all(o.state(o.old, oracle_mod.GD) == 7.0e4)
    st = smooth_state(N, M, 11)
    o = _ready(oracle_mod, N, M, st, use_diffusion=True, diffusion_coef=1e6,
               use_zonal_tend_filter=False)
    m0, _ = o.diag(o.old)
    o.ordinary_diffusion(600.0, o.old)
    m1, _ = o.diag(o.old)
    assert abs(m1 - m0) < 1e-14 * m0
    assert np.abs(o.state(o.old, oracle_mod.GD) - st[2]).max() > 1e-3


def test_diffusion_spherical_harmonic_eigenvalue(oracle_mod):
    """The scalar operator applied to Y_3^2 gives -n(n+1)/a^2 Y to O(dlat^2)."""
    errs = []
    for N, M in ((72, 37), (144, 73)):
        lat = -0.5 * np.pi + np.arange(M) * np.pi / (M - 1)
        lon = np.arange(N) * 2 * np.pi / N
        Y = (np.cos(lat) ** 2 * np.sin(lat))[:, None] * np.cos(2 * lon)[None, :]  # P_3^2 cos(2 lon)
        gd0 = 1.0e5 + 100.0 * Y
        # nu*dt = a^2 makes the single Euler update O(1e3), far above round-off
        o = _ready(oracle_mod, N, M, (np.zeros((M, N)), np.zeros((M - 1, N)), gd0, np.zeros((M, N))),
                   use_diffusion=True, diffusion_coef=A ** 2, use_zonal_tend_filter=False)
        o.ordinary_diffusion(1.0, o.old)  # gd <- gd + dt*nu*lap(gd)
        lap = o.state(o.old, oracle_mod.GD) - gd0
        expect = -12.0 * 100.0 * Y
        inner = slice(3, M - 3)
        errs.append(np.abs(lap[inner] - expect[inner]).max() / np.abs(expect).max())
    assert errs[0] < 2e-2 and errs[1] < errs[0] / 3.0  # second-order convergence


def test_diag_mass_of_constant_depth(oracle_mod):
    N, M = 180, 91
    gd = 1.0e4
    o = _ready(oracle_mod, N, M, (np.zeros((M, N)), np.zeros((M - 1, N)), np.full((M, N), gd),
                                  np.zeros((M, N))))
    mass, energy = o.diag(o.old)
    assert abs(mass - 4 * np.pi * A ** 2 * gd) < 2e-4 * 4 * np.pi * A ** 2 * gd
    # energy of a resting state: sum gd^2 cos
    fc = o.coef("full_cos")
    assert np.isclose(energy, gd ** 2 * N * fc.sum(), rtol=1e-13)


def test_isp_restatement_runs(oracle_mod):
    """isp_splitting (dycore_mod.F90:689-752) is restated in the oracle; as
    written it rescales beta by 4/dt and does not stay bounded."""
    N, M = 36, 19
    o = oracle_mod.Oracle(N, M, 600.0, split_scheme="isp", subcycles=3, use_zonal_tend_filter=False)
    o.set_initial_state(*rh_state(N, M))
    o.start()
    mass, energy, beta, ns = o.run(3)
    assert np.all(np.isfinite(mass[:2]))


def test_diffusion_fourth_order_eigenvalue(oracle_mod):
    """diffusion_order = 4 applies the scalar operator twice and subtracts
    (sign (-1)**(4/2+1) = -1): on Y_3^2 the update is -dt nu (n(n+1)/a^2)^2 Y
    to O(dlat^2) (diffusion_mod.F90:107, 172-179, 195-197)."""
    errs = []
    for N, M in ((72, 37), (144, 73)):
        lat = -0.5 * np.pi + np.arange(M) * np.pi / (M - 1)
        lon = np.arange(N) * 2 * np.pi / N
        Y = (np.cos(lat) ** 2 * np.sin(lat))[:, None] * np.cos(2 * lon)[None, :]
        gd0 = 1.0e5 + 100.0 * Y
        o = _ready(oracle_mod, N, M, (np.zeros((M, N)), np.zeros((M - 1, N)), gd0, np.zeros((M, N))),
                   use_diffusion=True, diffusion_order=4, diffusion_coef=A ** 4,
                   use_zonal_tend_filter=False)
        o.ordinary_diffusion(1.0, o.old)
        delta = o.state(o.old, oracle_mod.GD) - gd0
        expect = -(12.0 ** 2) * 100.0 * Y
        # the second application sees the one-sided polar rows of the first:
        # compare away from the poles, over a fixed latitude band
        inner = np.abs(lat) <= np.radians(60.0)
        errs.append(np.abs(delta[inner] - expect[inner]).max() / np.abs(expect).max())
    assert errs[0] < 2e-2 and errs[1] < errs[0] / 3.0


def test_filter_rescale_is_ill_conditioned_for_pure_waves(oracle_mod):
    """For a pure zonal-wavenumber state (Rossby-Haurwitz wave) the filter's
    s1 = sum(tend * weight) on a row vanishes analytically and is only
    round-off: sum|t w| / |sum t w| exceeds 1e12 on the polar rows, so the
    reference's rescale tend * s1/s2 (dycore_mod.F90:212-219) multiplies the
    filtered tendency by a ratio of round-off residues.  A balanced random
    state is well conditioned."""
    import __graft_entry__ as ge
    gd = ge.load_package()
    N, M = 144, 73

    def kappa(ic):
        o = oracle_mod.Oracle(N, M, 600.0, use_zonal_tend_filter=False)
        o.set_initial_state(*ic)
        o.start()
        o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
        du = o.tend(0, oracle_mod.T_DU)
        U = o.state(o.old, oracle_mod.IAP_U)
        rows = list(range(1, 8)) + list(range(M - 8, M - 1))
        return max(np.abs(du[j] * U[j]).sum() / abs((du[j] * U[j]).sum()) for j in rows)

    assert kappa(gd.ic_rossby_haurwitz(N, M)) > 1e12
    assert kappa(gd.ic_balanced(N, M, seed=3, nmax=10, n_mountains=2, mountain_seed=103)) < 1e5
