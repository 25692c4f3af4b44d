"""The corrector update k_update in rounds of eight or of four points and
partials per thread (GD_UPD_NARROW=0 / 1 when the handle is created; by
default the handle picks by the size of a workgroup's band).  A thread adds
its beta partials in the same order either way and every point's update is
independent, so both give the same bits: states, per-step mass, energy and
beta, for csp2, isp and unsplit runs, both precisions and an ensemble with a
subset of active members."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@contextmanager
def narrow_env(v):
    old = os.environ.get("GD_UPD_NARROW")
    os.environ["GD_UPD_NARROW"] = str(v)
    try:
        yield
    finally:
        if old is None:
            del os.environ["GD_UPD_NARROW"]
        else:
            os.environ["GD_UPD_NARROW"] = old


def run(gd, narrow, N, M, dt, steps, members=1, active=None, **kw):
    with narrow_env(narrow):
        with gd.Dycore(N, M, dt, members=members, **kw) as d:
            for m in range(members):
                d.set_state(*gd.ic_balanced(N, M, seed=13, nmax=8, n_mountains=2, mountain_seed=4,
                                            pert_seed=200 + m, pert_nmax=6, pert_rms=5.0), member=m)
            if active is not None:
                d.set_active_members(active)
            mass, energy, beta, _ = d.step(steps)
            n = members if active is None else active
            return [d.get_state(member=m) for m in range(n)], mass, energy, beta


CASES = [
    (720, 361, 120.0, 2, 1, None, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=2.5e4)),
    (360, 181, 240.0, 3, 1, None, dict(split_scheme="csp2", subcycles=6, use_diffusion=True, diffusion_coef=1e5)),
    (360, 181, 240.0, 2, 1, None, dict(split_scheme="isp", subcycles=3, use_diffusion=True, diffusion_coef=1e5)),
    (180, 91, 480.0, 3, 1, None, dict(split_scheme="none")),
    (180, 91, 480.0, 3, 12, 9, dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=1e5)),
]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_update_rounds_bitwise(gd, case, prec):
    N, M, dt, steps, members, active, opts = CASES[case]
    kw = dict(opts, precision=prec, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))
    print(f"k_update rounds N={N} M={M} members={members} active={active} {opts} {prec}", flush=True)
    a = run(gd, 0, N, M, dt, steps, members, active, **kw)
    b = run(gd, 1, N, M, dt, steps, members, active, **kw)
    for x, y in zip(a[0], b[0]):
        for f, g in zip(x, y):
            np.testing.assert_array_equal(f, g)
    for k in (1, 2, 3):
        np.testing.assert_array_equal(a[k], b[k])
