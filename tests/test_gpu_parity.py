"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

FP64 mode: the device evaluates the reference expressions with the same
operation order and no FMA contraction, so pointwise results agree to the
last bits; only sums (pole fluxes, filter products, the rescale and beta
inner products) run in another order.  Tolerances below are therefore
close to FP64 round-off: 1e-13 relative for one operator evaluation and
1e-12 .. 1e-11 for short integrations.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _pair(gd, oracle_mod, N, M, dt, ic="balanced", seed=3, nmax=10, mountains=0, **kw):
    if ic == "balanced":
        u, v, h, ghs = gd.ic_balanced(N, M, seed=seed, nmax=nmax, n_mountains=mountains,
                                      mountain_seed=seed + 100)
    elif ic == "rh":
        u, v, h, ghs = gd.ic_rossby_haurwitz(N, M)
    else:
        u, v, h, ghs = gd.ic_mountain_zonal_flow(N, M)
    cut = kw.pop("cutoff", [])
    o = oracle_mod.Oracle(N, M, dt, cutoff=cut, **kw)
    o.set_initial_state(u, v, h, ghs)
    o.start()
    d = gd.Dycore(N, M, dt, zonal_tend_filter_cutoff_wavenumber=cut, **kw)
    d.set_state(u, v, h, ghs)
    return o, d


@pytest.mark.parametrize("pass_,pid", [("all", 0), ("fast", 1), ("slow", 2)])
@pytest.mark.parametrize("adv", ["center_diff", "upwind"])
def test_space_operators(gd, oracle_mod, pass_, pid, adv):
    N, M = 72, 37
    o, d = _pair(gd, oracle_mod, N, M, 600.0, uv_adv_scheme=adv, mountains=2, cutoff=[5, 4, 3, 2])
    o.space_operators(o.old, 0, pid)
    du, dv, dgd = d.space_operators(pass_)
    T = oracle_mod
    for name, a, b in (("du", du, o.tend(0, T.T_DU)), ("dv", dv, o.tend(0, T.T_DV)),
                       ("dgd", dgd, o.tend(0, T.T_DGD))):
        err = np.abs(a - b).max(axis=1) / max(np.abs(b).max(), 1e-300)
        print(name, "worst rows", np.argsort(err)[-3:], err[np.argsort(err)[-3:]])
        assert err.max() < 1e-12, name
    if pass_ == "slow":
        assert np.all(dgd == 0)
    d.close()


def test_space_operators_unfiltered_rows_bitwise(gd, oracle_mod):
    """Rows off the filter and off the poles involve no reduction: identical bits."""
    N, M = 96, 49
    o, d = _pair(gd, oracle_mod, N, M, 600.0, cutoff=[4, 3])
    o.space_operators(o.old, 0, 0)
    du, dv, dgd = d.space_operators("all")
    ff, hf, _, _ = d.filter_table()
    rows = [j for j in range(2, M - 2) if not ff[j] and not ff[j - 1] and not ff[j + 1]]
    T = oracle_mod
    np.testing.assert_array_equal(du[rows], o.tend(0, T.T_DU)[rows])
    np.testing.assert_array_equal(dgd[rows], o.tend(0, T.T_DGD)[rows])
    hrows = [j for j in rows if not hf[j]]
    np.testing.assert_array_equal(dv[hrows], o.tend(0, T.T_DV)[hrows])
    d.close()


# Grid sizes exercise every FFT code path of the device filter: radix 4/3
# (72 -> 36), radix 8/8 (128 -> 64), radix 3/5 with an odd half length
# (90 -> 45), odd N (75, full complex transform), radix 7 (70 -> 35), a
# prime half length handled by the direct O(R^2) pass (94 -> 47), radix
# 8 x 2 x 3 x 3 (288 -> 144), the plans fixed at compile time (180, 360,
# 720 longitudes: half lengths 90, 180, 360), the 1440-longitude grid, and
# the 16-points-per-thread path up to the largest supported row (2560, 4096).
@pytest.mark.parametrize("N,M,cut", [(72, 37, [6, 5, 4, 3, 2, 1]), (128, 65, [30, 20, 10, 60, 61]),
                                     (90, 46, [44, 1, 43]), (75, 38, [36, 20, 1, 7]),
                                     (70, 36, [34, 2, 17]), (94, 48, [46, 3, 12]),
                                     (288, 145, [143, 100, 71, 9, 1]),
                                     (180, 91, [89, 60, 20, 3, 1]), (360, 181, [179, 120, 45, 7, 1]),
                                     (720, 361, [359, 300, 200, 120, 60, 30, 15, 8, 4, 2, 1]),
                                     (1440, 721, [719, 500, 360, 101, 33, 2]),
                                     (2560, 41, [1279, 700, 31, 2]), (4096, 33, [2047, 1000, 64, 1])])
def test_filter_rows(gd, oracle_mod, N, M, cut):
    o, d = _pair(gd, oracle_mod, N, M, 600.0, cutoff=cut)
    rng = np.random.default_rng(1)
    # the device's flag and cutoff table is filter_init's (filter_mod.F90:35-103)
    for x, y in zip(d.filter_table(), o.filter_table()):
        np.testing.assert_array_equal(x, y)
    # the rows to compare are chosen by the oracle's flags
    ff, hf, _, _ = o.filter_table()
    x = rng.standard_normal((M, N))
    y = d.filter_rows(x, half=False)
    for j in range(M):
        ref = o.filter_full(j + 1, x[j]) if ff[j] else x[j]
        assert rel_err(y[j], ref) < 1e-13, j
    xh = rng.standard_normal((M - 1, N))
    yh = d.filter_rows(xh, half=True)
    for j in range(M - 1):
        ref = o.filter_half(j + 1, xh[j]) if hf[j] else xh[j]
        assert rel_err(yh[j], ref) < 1e-13, j
    d.close()


@pytest.mark.parametrize("pass_,pid,dt", [("all", 0, 300.0), ("fast", 1, 100.0), ("slow", 2, 300.0)])
def test_predict_correct(gd, oracle_mod, pass_, pid, dt):
    N, M = 72, 37
    o, d = _pair(gd, oracle_mod, N, M, 600.0, mountains=1, cutoff=[6, 4, 2])
    beta_ref = o.predict_correct(dt, o.old, o.new, pid, True)
    beta = d.predict_correct(dt, pass_)
    assert abs(beta[0] - beta_ref) < 1e-13 * abs(beta_ref)
    u, v, h = d.get_state()
    assert rel_err(h, o.state(o.new, oracle_mod.GD)) < 1e-14
    assert rel_err(u, o.state(o.new, oracle_mod.U)) < 1e-12
    assert rel_err(v, o.state(o.new, oracle_mod.V)) < 1e-12
    d.close()


@pytest.mark.parametrize("order", [2, 4, 6])
def test_ordinary_diffusion(gd, oracle_mod, order):
    # higher orders apply the second-order operator order/2 times
    # (diffusion_mod.F90:107-179); the coefficient keeps dt*nu*lap^k small
    N, M = 72, 37
    nu = {2: 5e5, 4: 5e5 * 4e10, 6: 5e5 * 1.6e21}[order]
    o, d = _pair(gd, oracle_mod, N, M, 600.0, use_diffusion=True, diffusion_coef=nu,
                 diffusion_order=order, cutoff=[5, 3, 2])
    o.ordinary_diffusion(600.0, o.old)
    d.ordinary_diffusion(600.0)
    u, v, h = d.get_state()
    assert rel_err(h, o.state(o.old, oracle_mod.GD)) < 1e-14
    assert rel_err(u, o.state(o.old, oracle_mod.U)) < 1e-12
    assert rel_err(v, o.state(o.old, oracle_mod.V)) < 1e-12
    d.close()


@pytest.mark.parametrize("split", ["csp2", "none"])
def test_steps_vs_oracle(gd, oracle_mod, split):
    N, M, dt = 72, 37, 900.0
    kw = dict(split_scheme=split, subcycles=3, use_diffusion=True, diffusion_coef=2e5,
              cutoff=[6, 5, 4, 3, 2])
    o, d = _pair(gd, oracle_mod, N, M, dt, mountains=2, **kw)
    om, oe, ob, ns = o.run(8)
    mass, energy, beta, _ = d.step(8)
    u, v, h = d.get_state()
    # 8 steps x 5 integrator calls: the ~1e-13 reduction-order differences of
    # the filtered rows compound through the dynamics
    assert rel_err(h, o.state(o.old, oracle_mod.GD)) < 1e-11
    assert rel_err(u, o.state(o.old, oracle_mod.U)) < 1e-10
    assert rel_err(v, o.state(o.old, oracle_mod.V)) < 1e-10
    assert np.all(np.abs(mass[:, 0] - om) <= 1e-13 * np.abs(om))
    assert np.all(np.abs(energy[:, 0] - oe) <= 1e-12 * np.abs(oe))
    assert np.all(np.abs(beta[:, 0] - ob) <= 1e-10)
    d.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_isp_steps_vs_oracle(gd, oracle_mod, precision):
    """isp_splitting (dycore_mod.F90:689-752): slow tendency saved, three fast
    stages per subcycle with the accumulated fast tendency, two slow
    corrections, beta = <R,F>/<R,R> * 4/dt applied as (dt/2) beta."""
    N, M, dt = 72, 37, 900.0
    kw = dict(split_scheme="isp", subcycles=3, use_diffusion=True, diffusion_coef=2e5,
              cutoff=[6, 5, 4])
    o, d0 = _pair(gd, oracle_mod, N, M, dt, mountains=1, **kw)
    d0.close()
    u0, v0, h0, ghs = gd.ic_balanced(N, M, seed=3, nmax=10, n_mountains=1, mountain_seed=103)
    d = gd.Dycore(N, M, dt, zonal_tend_filter_cutoff_wavenumber=kw.pop("cutoff"),
                  precision=precision, **kw)
    d.set_state(u0, v0, h0, ghs)
    om, oe, ob, ns = o.run(3)
    mass, energy, beta, _ = d.step(3)
    u, v, h = d.get_state()
    tol = 1e-11 if precision == "fp64" else 2e-5
    assert rel_err(h, o.state(o.old, oracle_mod.GD)) < tol
    assert rel_err(u, o.state(o.old, oracle_mod.U)) < 100 * tol
    assert rel_err(v, o.state(o.old, oracle_mod.V)) < 100 * tol
    assert np.all(np.abs(beta[:, 0] - ob) <= 100 * tol * np.abs(ob))
    d.close()


# The Rossby-Haurwitz wave and the Shamir-Paldor wave are pure zonal
# wavenumber patterns: on every row the filter's inner product s1 = sum(tend *
# weight) vanishes in exact arithmetic and is left as round-off of 1e-16 of
# its terms, which still passes the reference's |s1| > 1e-16 test
# (dycore_mod.F90:212-219), so the rescale s1/s2 is a ratio of round-off
# residues that depends on the summation order (tests/test_oracle.py::
# test_filter_rescale_is_ill_conditioned_for_pure_waves).  Those two cases are
# compared without the filter, on a grid and time step whose pole rows are
# stable without it (fast step 75 s, 2.5 deg).
@pytest.mark.parametrize("case,use_filter", [("steady_geostrophic", True), ("jet_zonal_flow", True),
                                             ("shallow_water_waves", False), ("rossby_haurwitz", False),
                                             ("mountain_zonal_flow", True)])
def test_reference_cases_vs_oracle(gd, oracle_mod, case, use_filter):
    """Each of the reference's test cases stepped on the device and by the
    oracle from the same initial condition."""
    N, M, dt = (144, 73, 600.0) if use_filter else (72, 37, 300.0)
    ic = {"steady_geostrophic": lambda: gd.ic_steady_geostrophic(N, M),
          "jet_zonal_flow": lambda: gd.ic_jet_zonal_flow(N, M),
          "shallow_water_waves": lambda: gd.ic_shallow_water_waves(N, M)[:4],
          "rossby_haurwitz": lambda: gd.ic_rossby_haurwitz(N, M),
          "mountain_zonal_flow": lambda: gd.ic_mountain_zonal_flow(N, M)}[case]()
    kw = dict(split_scheme="csp2", subcycles=4, use_diffusion=True, diffusion_coef=1e5,
              use_zonal_tend_filter=use_filter)
    cut = gd.cutoff_from_lat0(N, M, 70.0) if use_filter else []
    o = oracle_mod.Oracle(N, M, dt, cutoff=cut, **kw)
    o.set_initial_state(*ic)
    o.start()
    with gd.Dycore(N, M, dt, zonal_tend_filter_cutoff_wavenumber=cut, **kw) as d:
        d.set_state(*ic)
        om, oe, ob, _ = o.run(4)
        mass, energy, beta, _ = d.step(4)
        u, v, h = d.get_state()
    assert rel_err(h, o.state(o.old, oracle_mod.GD)) < 1e-11
    assert rel_err(u, o.state(o.old, oracle_mod.U)) < 1e-10
    assert rel_err(v, o.state(o.old, oracle_mod.V)) < 1e-9
    assert np.all(np.abs(mass[:, 0] - om) <= 1e-13 * np.abs(om))
    assert np.all(np.abs(energy[:, 0] - oe) <= 1e-12 * np.abs(oe))
