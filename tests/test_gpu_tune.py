"""The tile autotuner (gd_autotune) picks among launch shapes that cannot
change a result, so a tuned run gives the bits of the default run; tuning
leaves the members' state as it was; a cached entry is applied without
timing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _kw(gd, N, M, prec):
    return dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=2e5, precision=prec,
                zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0))


def _ics(gd, N, M, members):
    return [gd.ic_balanced(N, M, seed=17, nmax=8, n_mountains=2, mountain_seed=3, pert_seed=300 + m,
                           pert_nmax=6, pert_rms=5.0) for m in range(members)]


def _run(gd, N, M, dt, steps, members, prec, tune, active=None):
    ics = _ics(gd, N, M, members)
    with gd.Dycore(N, M, dt, members=members, **_kw(gd, N, M, prec)) as d:
        for m, ic in enumerate(ics):
            d.set_state(*ic, member=m)
        if active:
            d.set_active_members(active)
        choice = None
        if tune:
            before = [d.get_state(member=m) for m in range(members)]
            choice = d.autotune(steps=5)
            for m in range(members):  # tuning put the state back
                for x, y in zip(d.get_state(member=m), before[m]):
                    np.testing.assert_array_equal(x, y)
        mass, energy, beta, _ = d.step(steps)
        n = active or members
        return [d.get_state(member=m) for m in range(n)], mass[:, :n], energy[:, :n], beta[:, :n], choice


@pytest.mark.parametrize("N,M,dt,members,active,prec", [(72, 37, 900.0, 3, None, "mixed"),
                                                      (72, 37, 900.0, 4, 3, "fp64"),
                                                      (360, 181, 240.0, 8, None, "mixed"),
                                                      (360, 181, 240.0, 1, None, "fp64")])
def test_tuned_run_gives_default_bits(gd, N, M, dt, members, active, prec):
    print(f"autotune {N}x{M} x{members} active={active} {prec}", flush=True)
    a = _run(gd, N, M, dt, 4, members, prec, False, active)
    b = _run(gd, N, M, dt, 4, members, prec, True, active)
    choice = b[4]
    assert choice["upd_wpm"] >= 1 and choice["ms_per_step"] > 0
    for x, y in zip(a[0], b[0]):
        for f, g in zip(x, y):
            np.testing.assert_array_equal(f, g)
    for k in (1, 2, 3):
        np.testing.assert_array_equal(a[k], b[k])


def test_autotune_cache(gd, tmp_path):
    N, M = 720, 361
    cache = tmp_path / "tune.txt"
    ic = gd.ic_balanced(N, M, seed=20241, nmax=20, n_mountains=3, mountain_seed=777)
    kw = dict(_kw(gd, N, M, "mixed"), subcycles=6, diffusion_coef=2.5e4)
    print("autotune cache 720x361", flush=True)
    with gd.Dycore(N, M, 120.0, **kw) as d:
        d.set_state(*ic)
        first = d.autotune(steps=10, cache=cache)
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "N=720 M=361" in lines[0]
    with gd.Dycore(N, M, 120.0, **kw) as d:
        d.set_state(*ic)
        again = d.autotune(steps=10, cache=cache)
        d.step(2)
    assert again == first  # taken from the cache: same shapes, same recorded time
    assert len(cache.read_text().splitlines()) == 1
