"""Mixed precision against FP64 and ensemble batching on the device.

Mixed mode evaluates the momentum stencils and the momentum filter rows in
FP32 while the state update, mass-flux divergence (and its filter rows), pole
sums and inner products stay FP64.  Tolerances: one FP32 operator
evaluation 5e-4 relative to the field maximum; the mass budget of a mixed run
tracks the FP64 run to 1e-13 relative per day.
"""
import numpy as np
import pytest

from conftest import l2_rel, rel_err

pytestmark = pytest.mark.gpu


def _cfg(gd, N, M, **kw):
    base = dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=1e5,
                zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 75.0))
    base.update(kw)
    return base


def test_mixed_space_operators_vs_oracle(gd, oracle_mod):
    N, M = 180, 91
    u, v, h, ghs = gd.ic_balanced(N, M, seed=5, nmax=12, n_mountains=3, mountain_seed=9)
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    o = oracle_mod.Oracle(N, M, 480.0, cutoff=cut)
    o.set_initial_state(u, v, h, ghs)
    o.start()
    with gd.Dycore(N, M, 480.0, zonal_tend_filter_cutoff_wavenumber=cut, precision="mixed") as d:
        d.set_state(u, v, h, ghs)
        for pname, pid in (("all", 0), ("fast", 1), ("slow", 2)):
            o.space_operators(o.old, 0, pid)
            du, dv, dgd = d.space_operators(pname)
            # FP32 tendencies: du/dv are small differences of nearly balanced
            # Coriolis and pressure-gradient terms, and next to the poles the
            # zonal advection u2*V(i+1) - u1*V(i-1) cancels across neighbours
            # 480 m apart; both enlarge FP32 rounding ~1000x relative to the
            # largest tendency
            assert rel_err(du, o.tend(0, oracle_mod.T_DU)) < 5e-4, pname
            assert rel_err(dv, o.tend(0, oracle_mod.T_DV)) < 5e-4, pname
            # mass continuity is FP64 in mixed mode
            if pname != "slow":
                assert rel_err(dgd, o.tend(0, oracle_mod.T_DGD)) < 1e-12, pname


@pytest.mark.parametrize("N,M,dt,steps", [(360, 181, 240.0, 360)])
def test_mixed_one_day_vs_fp64(gd, N, M, dt, steps):
    """1 day at 360x181: height l2 within 1e-6 of FP64, mass budget identical to 1e-13."""
    ic = gd.ic_balanced(N, M, seed=21, nmax=20)
    out = {}
    for prec in ("fp64", "mixed"):
        with gd.Dycore(N, M, dt, precision=prec, **_cfg(gd, N, M)) as d:
            d.set_state(*ic)
            m0, _ = d.diag()
            mass, energy, beta, _ = d.step(steps)
            out[prec] = (d.get_state(), mass[:, 0], energy[:, 0], m0[0])
    (u64, v64, h64), mass64, en64, m064 = out["fp64"]
    (u32, v32, h32), mass32, en32, m032 = out["mixed"]
    assert m064 == m032
    assert l2_rel(h32, h64) < 1e-6
    assert l2_rel(u32, u64) < 1e-4
    assert abs(mass32[-1] - mass64[-1]) <= 1e-13 * abs(mass64[-1])
    assert np.all(np.abs(en32 - en64) <= 1e-8 * np.abs(en64))


def test_ensemble_members_match_single_runs_bitwise(gd):
    """Each member of a batched ensemble equals its own single run, bit for bit."""
    N, M, dt = 180, 91, 480.0
    cfg = _cfg(gd, N, M, subcycles=4)
    ics = [gd.ic_balanced(N, M, seed=3, nmax=10, pert_seed=100 + k, pert_nmax=10, pert_rms=5.0)
           for k in range(3)]
    singles = []
    for ic in ics:
        with gd.Dycore(N, M, dt, **cfg) as d:
            d.set_state(*ic)
            mass, energy, beta, _ = d.step(20)
            singles.append((d.get_state(), mass[:, 0], energy[:, 0], beta[:, 0]))
    with gd.Dycore(N, M, dt, members=3, **cfg) as e:
        for k, ic in enumerate(ics):
            e.set_state(*ic, member=k)
        mass, energy, beta, _ = e.step(20)
        for k in range(3):
            (u, v, h) = e.get_state(member=k)
            (us, vs, hs), ms, es, bs = singles[k]
            np.testing.assert_array_equal(h, hs)
            np.testing.assert_array_equal(u, us)
            np.testing.assert_array_equal(v, vs)
            np.testing.assert_array_equal(mass[:, k], ms)
            np.testing.assert_array_equal(energy[:, k], es)
            np.testing.assert_array_equal(beta[:, k], bs)
    # the members differ from each other
    assert np.abs(singles[0][0][2] - singles[1][0][2]).max() > 0


def test_nan_guard_reports_step(gd):
    """An absurd time step drives the run to NaN; gd_step reports the first NaN step
    (diag_mod.F90:79-87 stops the reference with an error there)."""
    N, M = 72, 37
    ic = gd.ic_balanced(N, M, seed=2, nmax=6)
    with gd.Dycore(N, M, 2.0e5, split_scheme="none", use_zonal_tend_filter=False) as d:
        d.set_state(*ic)
        mass, energy, beta, nan_step = d.step(200, raise_on_nan=False)
        assert nan_step >= 1
        assert np.all(np.isfinite(mass[: nan_step - 1, 0]))
        assert not np.isfinite(mass[nan_step - 1, 0]) or not np.isfinite(energy[nan_step - 1, 0])
        with pytest.raises(gd.DycoreError) as ei:
            d.step(5)
        assert ei.value.code == gd.GD_ERR_NAN


def test_rejects_bad_state_and_config(gd):
    N, M = 72, 37
    u, v, h, ghs = gd.ic_balanced(N, M, seed=2, nmax=6)
    with gd.Dycore(N, M, 600.0) as d:
        bad = u.copy()
        bad[0, 3] = 1.0  # u must vanish on the pole rows
        with pytest.raises(gd.DycoreError):
            d.set_state(bad, v, h, ghs)
        with pytest.raises(gd.DycoreError):
            d.set_state(u, v, -h, ghs)
    with pytest.raises(gd.DycoreError):
        gd.Dycore(N, M, 600.0, uv_adv_scheme="weno")  # not provided (incomplete in the reference)
    with pytest.raises(gd.DycoreError):
        gd.Dycore(N, M, 600.0, use_diffusion=True, diffusion_order=1)
    with pytest.raises(gd.DycoreError):
        gd.Dycore(N, M, 600.0, zonal_tend_filter_cutoff_wavenumber=[0, 4])
    with pytest.raises(gd.DycoreError):
        gd.Dycore(N, M, 600.0, zonal_tend_filter_cutoff_wavenumber=[N // 2])
    with pytest.raises(gd.DycoreError) as ei:  # filter rows longer than 4096 points
        gd.Dycore(4104, 9, 60.0, zonal_tend_filter_cutoff_wavenumber=[4])
    assert ei.value.code == gd.GD_ERR_NOT_SUPPORTED


def test_time_kernel_leaves_state_untouched(gd):
    """The measurement entry point replays single kernels on scratch slots
    only: the state X and the next steps are unchanged."""
    N, M = 72, 37
    u, v, h, ghs = gd.ic_balanced(N, M, seed=5, nmax=6)
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=2e5,
              zonal_tend_filter_cutoff_wavenumber=[6, 5, 4, 3])
    with gd.Dycore(N, M, 900.0, **kw) as d, gd.Dycore(N, M, 900.0, **kw) as ref:
        d.set_state(u, v, h, ghs)
        ref.set_state(u, v, h, ghs)
        d.step(2)
        ref.step(2)
        for k in ("tend", "filter", "update"):
            assert d.time_kernel(k, 5) > 0.0
        for a, b in zip(d.get_state(), ref.get_state()):
            np.testing.assert_array_equal(a, b)
        ma, ea, _, _ = d.step(2)
        mb, eb, _, _ = ref.step(2)
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(ea, eb)


def test_active_members_and_member_copy(gd):
    """Only the active prefix is stepped; a copied member continues bit for bit
    (the serving pools compact their runs this way)."""
    N, M, dt = 72, 37, 900.0
    cfg = _cfg(gd, N, M, subcycles=3)
    ics = [gd.ic_balanced(N, M, seed=3, nmax=8, pert_seed=40 + k, pert_nmax=8, pert_rms=5.0)
           for k in range(3)]
    with gd.Dycore(N, M, dt, members=3, **cfg) as e, gd.Dycore(N, M, dt, **cfg) as one:
        for k, ic in enumerate(ics):
            e.set_state(*ic, member=k)
        frozen = e.get_state(member=2)
        e.set_active_members(2)
        mass, _, _, _ = e.step(5)
        for a, b in zip(e.get_state(member=2), frozen):
            np.testing.assert_array_equal(a, b)
        one.set_state(*ics[0])
        m1, _, _, _ = one.step(5)
        np.testing.assert_array_equal(mass[:, 0], m1[:, 0])
        e.copy_member(2, 0)
        e.set_active_members(3)
        mass, _, _, _ = e.step(4)
        m1, _, _, _ = one.step(4)
        np.testing.assert_array_equal(mass[:, 2], m1[:, 0])
        np.testing.assert_array_equal(mass[:, 0], m1[:, 0])
        for a, b in zip(e.get_state(member=2), one.get_state()):
            np.testing.assert_array_equal(a, b)
        with pytest.raises(gd.DycoreError):
            e.set_active_members(4)


def test_seams_leave_inactive_members_alone(gd):
    """predict_correct and ordinary_diffusion step the active members only:
    an inactive member keeps its state bit for bit, and space_operators
    refuses an inactive member instead of returning a stale tendency."""
    N, M = 72, 37
    kw = _cfg(gd, N, M, subcycles=3, diffusion_coef=2e5)
    print("seams with 3 members, 2 active", flush=True)
    with gd.Dycore(N, M, 600.0, members=3, **kw) as d, gd.Dycore(N, M, 600.0, **kw) as one:
        ics = [gd.ic_balanced(N, M, seed=31 + m, nmax=6) for m in range(3)]
        for m, ic in enumerate(ics):
            d.set_state(*ic, member=m)
        one.set_state(*ics[1])
        before = d.get_state(member=2)
        d.set_active_members(2)
        beta = d.predict_correct(300.0, "all")
        d.ordinary_diffusion(300.0)
        for x, y in zip(d.get_state(member=2), before):
            np.testing.assert_array_equal(x, y)
        # an active member advanced exactly as a single run does
        b1 = one.predict_correct(300.0, "all")
        one.ordinary_diffusion(300.0)
        assert beta[1] == b1[0]
        for x, y in zip(d.get_state(member=1), one.get_state()):
            np.testing.assert_array_equal(x, y)
        with pytest.raises(gd.DycoreError, match="not active"):
            d.space_operators("all", member=2)
        d.space_operators("all", member=1)
