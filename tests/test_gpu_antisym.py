"""check_antisymmetry (dycore_mod.F90:794-851) on the device
(gd_check_antisymmetry): the ten inner products of the operator terms with
the state, against the oracle's og_check_antisymmetry on small grids, and the
four identities at full size, where no fixture is needed: with the center
scheme advection is energy-neutral and the PGF / mass-divergence and Coriolis
pairs cancel, to round-off relative to the magnitudes of their terms.  The
upwind scheme breaks the advection identity only, which shows the check can
fail."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the terms each identity pairs: advection; u_pgf + mass_div_lon; fv - fu;
# v_pgf + mass_div_lat
PAIRS = [(0, 1, 4, 5), (3, 8), (2, 6), (7, 9)]


def scales(mags):
    return [sum(mags[k] for k in p) for p in PAIRS]


def state(gd, N, M, seed):
    return gd.ic_balanced(N, M, seed=seed, nmax=min(20, M // 3), n_mountains=3, mountain_seed=seed + 1)


@pytest.mark.parametrize("scheme", ["center_diff", "upwind"])
@pytest.mark.parametrize("N,M", [(72, 37), (144, 73), (45, 23)])
def test_terms_vs_oracle(gd, oracle_mod, N, M, scheme):
    u, v, h, ghs = state(gd, N, M, 11)
    assert np.abs(ghs).max() > 0
    cut = gd.cutoff_from_lat0(N, M, 60.0)
    o = oracle_mod.Oracle(N, M, 600.0, cutoff=cut, uv_adv_scheme=scheme)
    o.set_initial_state(u, v, h, ghs)
    o.start()
    o.space_operators(o.old, 0, oracle_mod.ALL_PASS)
    ref = o.check_antisymmetry(0, o.old)
    with gd.Dycore(N, M, 600.0, zonal_tend_filter_cutoff_wavenumber=cut, uv_adv_scheme=scheme) as d:
        d.set_state(u, v, h, ghs)
        terms, mags, sums = d.check_antisymmetry()
    sc = scales(mags)
    print(f"antisymmetry {N}x{M} {scheme}: device {sums} oracle {ref} scales {sc}", flush=True)
    for k in range(4):
        assert abs(sums[k] - ref[k]) <= 1e-12 * sc[k], (k, sums[k], ref[k], sc[k])
    assert all(m > 0 for m in mags)


@pytest.mark.parametrize("N,M,prec", [(720, 361, "fp64"), (720, 361, "mixed"), (360, 181, "fp64")])
def test_identities_vanish_at_full_size(gd, N, M, prec):
    """The benchmark grids with mountains (ghs != 0), after a few steps so
    the state is the model's own: every identity vanishes to round-off."""
    u, v, h, ghs = state(gd, N, M, 20241)
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    dt = 120.0 if N == 720 else 240.0
    with gd.Dycore(N, M, dt, zonal_tend_filter_cutoff_wavenumber=cut, precision=prec, split_scheme="csp2",
                   subcycles=6, use_diffusion=True, diffusion_coef=2.5e4 if N == 720 else 1e5) as d:
        d.set_state(u, v, h, ghs)
        d.step(6)
        terms, mags, sums = d.check_antisymmetry()
    sc = scales(mags)
    print(f"antisymmetry {N}x{M} {prec}: sums {sums} relative {[abs(s) / c for s, c in zip(sums, sc)]}", flush=True)
    for k in range(4):
        assert abs(sums[k]) <= 1e-12 * sc[k], (k, sums[k], sc[k])


def test_upwind_breaks_only_the_advection_identity(gd):
    N, M = 360, 181
    u, v, h, ghs = state(gd, N, M, 7)
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    with gd.Dycore(N, M, 240.0, zonal_tend_filter_cutoff_wavenumber=cut, uv_adv_scheme="upwind") as d:
        d.set_state(u, v, h, ghs)
        terms, mags, sums = d.check_antisymmetry()
    sc = scales(mags)
    print(f"antisymmetry upwind {N}x{M}: sums {sums} scales {sc}", flush=True)
    assert abs(sums[0]) > 1e-6 * sc[0]
    for k in (1, 2, 3):
        assert abs(sums[k]) <= 1e-12 * sc[k], (k, sums[k], sc[k])
