"""The wide phase launch (k_wphase: workgroups of two tiles; a filter row's
segments evaluated two at a time and filtered with one or two longitudes
per thread; row workgroups of two consecutive row tiles, of one row or two,
with a spare tile past the last) against the fused launch k_phase
(GD_FUSED=2 and GD_FUSED=1 when the handle is created).  Each tile does
k_phase's arithmetic and the filter's sums keep filter_row's order, so the
two must agree bit for bit: states, per-step mass, energy and beta, for rows
of two and three segments (255 to 762 longitudes), even and odd num_lon,
every split scheme, both precisions and an ensemble with a subset of active
members.
"""
import os
from contextlib import contextmanager

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@contextmanager
def form_env(form):
    old = os.environ.get("GD_FUSED")
    os.environ["GD_FUSED"] = str(form)
    try:
        yield
    finally:
        if old is None:
            del os.environ["GD_FUSED"]
        else:
            os.environ["GD_FUSED"] = old


def run(gd, form, N, M, dt, steps, members=1, active=None, **kw):
    with form_env(form):
        with gd.Dycore(N, M, dt, members=members, **kw) as d:
            assert d.fused_phases() == form
            for m in range(members):
                d.set_state(*gd.ic_balanced(N, M, seed=11, nmax=8, n_mountains=2, mountain_seed=5,
                                            pert_seed=100 + m, pert_nmax=6, pert_rms=5.0), member=m)
            if active is not None:
                d.set_active_members(active)
            fam = set(d.profile_kernels(1))
            mass, energy, beta, _ = d.step(steps)
            n = members if active is None else active
            states = [d.get_state(member=m) for m in range(n)]
    return fam, (states, mass, energy, beta)


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
    (720, 361, 120.0, 3, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=2.5e4)),
    # three segments: workgroups pairing tiles of two rows, pole-row tiles
    # paired with others (a spare tile: test_wide_phase_spare_tile)
    (600, 91, 150.0, 3, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=4e4)),
    (762, 37, 120.0, 2, dict(split_scheme="none", use_diffusion=True, diffusion_coef=2.5e4)),
    (360, 181, 240.0, 3, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=1e5)),
    (360, 181, 240.0, 2, dict(split_scheme="isp", subcycles=3, use_diffusion=True, diffusion_coef=1e5)),
    (360, 91, 300.0, 3, dict(split_scheme="none", uv_adv_scheme="upwind")),
    (480, 241, 180.0, 2, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=6e4,
                              diffusion_order=4)),
    # odd rows (complex transform), and a last segment of 121 and of 1 longitude
    (375, 49, 300.0, 3, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=1e5)),
    (255, 37, 600.0, 3, dict(split_scheme="csp2", subcycles=3)),
]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_wide_phase_bitwise(gd, case, prec):
    N, M, dt, steps, opts = CASES[case]
    kw = dict(opts, precision=prec, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    print(f"wide vs fused N={N} M={M} {opts} {prec}", flush=True)
    fw, rw = run(gd, 2, N, M, dt, steps, **kw)
    ff, rf = run(gd, 1, N, M, dt, steps, **kw)
    assert "wphase" in fw and "phase" not in fw
    assert "phase" in ff and "wphase" not in ff
    assert_same(rw, rf)


def test_wide_phase_ensemble_active_subset(gd):
    N, M = 360, 91
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=1e5, precision="mixed",
              zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    print("wide ensemble 3 members, 2 active", flush=True)
    assert_same(run(gd, 2, N, M, 300.0, 3, members=3, active=2, **kw)[1],
                run(gd, 1, N, M, 300.0, 3, members=3, active=2, **kw)[1])


def test_wide_phase_only_for_two_or_three_segments(gd):
    """GD_FUSED=2 falls back to k_phase where a row is one segment or more
    than three."""
    for N, M in ((180, 91), (763, 37), (1024, 33)):
        with form_env(2):
            with gd.Dycore(N, M, 60.0, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0)) as d:
                assert d.fused_phases() == 1


def far_rows(d, M):
    """Rows with a component outside the filter, and the pole rows: the rows
    of the fused launches' row tiles."""
    ff, hf, _, _ = d.filter_table()
    return sum(1 for j in range(M) if j in (0, M - 1) or not ff[j] or (j <= M - 2 and not hf[j]))


def test_wide_phase_spare_tile(gd):
    """Rows of three segments and an odd number of far rows: an odd tile
    count, so the last row workgroup has a spare tile."""
    N = 600
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=4e4, precision="mixed")
    for M in range(61, 101, 2):
        cut = gd.cutoff_from_lat0(N, M, 60.0)
        with gd.Dycore(N, M, 150.0, zonal_tend_filter_cutoff_wavenumber=cut, **kw) as d:
            nf = far_rows(d, M)
        if nf % 2 == 1:
            break
    else:
        pytest.fail("no grid with an odd far-row count")
    print(f"wide spare tile N={N} M={M} far rows {nf}", flush=True)
    kw["zonal_tend_filter_cutoff_wavenumber"] = cut
    assert_same(run(gd, 2, N, M, 150.0, 3, **kw)[1], run(gd, 1, N, M, 150.0, 3, **kw)[1])
