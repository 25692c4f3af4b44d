"""The fused launches (k_phase: a tendency phase and its polar filter rows in
one launch; k_dphase: the last diffusion application and its filter rows)
against k_tend / k_diffuse followed by k_filter (GD_FUSED=0 when the handle
is created).  Both evaluate every value with the same expressions in
the same order, so they must agree bit for bit: states, per-step mass, energy
and beta, for every split scheme, both precisions, both advection schemes,
even and odd num_lon, and ensembles with a subset of active members.  The
slow pass's dgd rows are zero: k_phase gives them T = +0 without a transform,
k_filter transforms them, so the csp2 cases also check that shortcut.
"""
import os
from contextlib import contextmanager

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@contextmanager
def fused_env(on):
    old = os.environ.get("GD_FUSED")
    os.environ["GD_FUSED"] = "1" if on else "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["GD_FUSED"]
        else:
            os.environ["GD_FUSED"] = old


def families(gd, N, M, dt, fused, **kw):
    with fused_env(fused):
        with gd.Dycore(N, M, dt, **kw) as d:
            d.set_state(*gd.ic_balanced(N, M, seed=3, nmax=6))
            return set(d.profile_kernels(1))


def run(gd, fused, N, M, dt, steps, members=1, active=None, **kw):
    with fused_env(fused):
        with gd.Dycore(N, M, dt, members=members, **kw) as d:
            for m in range(members):
                d.set_state(*gd.ic_balanced(N, M, seed=11, nmax=8, n_mountains=2, mountain_seed=5,
                                            pert_seed=100 + m, pert_nmax=6, pert_rms=5.0), member=m)
            if active is not None:
                d.set_active_members(active)
            mass, energy, beta, _ = d.step(steps)
            n = members if active is None else active
            states = [d.get_state(member=m) for m in range(n)]
    return states, mass, energy, beta


def assert_same(a, b):
    sa, ma, ea, ba = a
    sb, mb, eb, bb = b
    for x, y in zip(sa, sb):
        for f, g in zip(x, y):
            np.testing.assert_array_equal(f, g)
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(ea, eb)
    np.testing.assert_array_equal(ba, bb)


CASES = [
    # (N, M, dt, steps, options)
    (72, 37, 900.0, 4, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=2e5)),
    (72, 37, 900.0, 4, dict(split_scheme="isp", subcycles=3, use_diffusion=True, diffusion_coef=2e5)),
    (72, 37, 600.0, 4, dict(split_scheme="none", use_diffusion=False)),
    (72, 37, 900.0, 4, dict(split_scheme="csp2", subcycles=3, uv_adv_scheme="upwind")),
    (45, 23, 900.0, 4, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=2e5)),
    (360, 181, 240.0, 3, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=1e5)),
    (1024, 33, 30.0, 2, dict(split_scheme="csp2", subcycles=2, use_diffusion=True, diffusion_coef=1e4)),
    # the benchmarked shape (compile-time 360-point FFT plan with the radix-8
    # swizzle, three segments per row) and the 180-longitude plan
    (720, 361, 120.0, 3, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=2.5e4)),
    (180, 91, 480.0, 3, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=1e5)),
    # eight longitudes per thread in the filter tiles (1440: the compile-time
    # 720-point plan, six segments per row)
    (1440, 61, 60.0, 2, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=6.25e3)),
    (72, 37, 900.0, 4, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_order=4,
                            diffusion_coef=2e5 * 4e10)),
]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_phase_bitwise(gd, case, prec):
    N, M, dt, steps, opts = CASES[case]
    kw = dict(opts, precision=prec, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    print(f"fused vs unfused N={N} M={M} {opts} {prec}", flush=True)
    assert "phase" in families(gd, N, M, dt, True, **kw)
    assert "phase" not in families(gd, N, M, dt, False, **kw)
    assert_same(run(gd, True, N, M, dt, steps, **kw), run(gd, False, N, M, dt, steps, **kw))


def test_fused_phase_ensemble_active_subset(gd):
    N, M = 72, 37
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=2e5, precision="mixed",
              zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    print("fused ensemble 3 members, 2 active", flush=True)
    assert_same(run(gd, True, N, M, 900.0, 4, members=3, active=2, **kw),
                run(gd, False, N, M, 900.0, 4, members=3, active=2, **kw))


def test_fused_phase_off_above_2048_longitudes(gd):
    """num_lon > 2048 keeps k_tend + k_filter even when fusing is asked for;
    past 1024 the phases fuse by default up to 6144 row tiles."""
    N, M = 2050, 17
    kw = dict(split_scheme="csp2", subcycles=2, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    fam = families(gd, N, M, 30.0, True, **kw)
    assert "phase" not in fam and "filter" in fam
    old = os.environ.pop("GD_FUSED", None)
    try:
        N, M = 1440, 37
        with gd.Dycore(N, M, 60.0, members=30, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0)) as d:
            assert not d.fused_phases()  # 30 x 37 x 6 = 6,660 row tiles
            d.set_active_members(27)
            assert d.fused_phases()      # 5,994
    finally:
        if old is not None:
            os.environ["GD_FUSED"] = old


def test_fused_phase_rule_by_launch_size(gd):
    """Without GD_FUSED, phases are fused while active members x num_lat x
    segments <= 24576 row tiles (360x181: 362 tiles per member)."""
    old = os.environ.pop("GD_FUSED", None)
    try:
        N, M = 360, 181
        with gd.Dycore(N, M, 240.0, members=68, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 75.0)) as d:
            assert not d.fused_phases()
            d.set_active_members(67)
            assert d.fused_phases()
            d.set_active_members(68)
            assert not d.fused_phases()
    finally:
        if old is not None:
            os.environ["GD_FUSED"] = old
