"""The captured step graphs: gd_step replays a graph of eight steps while at
least eight remain and a one-step graph for the rest.  Both record the
launches of enqueue_step, so a run stepped in one call, one step per call,
or with every kernel launched eagerly (GD_NO_GRAPH=1) gives the same bits:
states, per-step mass, energy and beta."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(gd, chunks, eager=False, members=1, **kw):
    N, M = 180, 91
    old = os.environ.get("GD_NO_GRAPH")
    if eager:
        os.environ["GD_NO_GRAPH"] = "1"
    try:
        with gd.Dycore(N, M, 480.0, members=members, zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 60.0),
                       **kw) as d:
            for m in range(members):
                d.set_state(*gd.ic_balanced(N, M, seed=21, nmax=8, n_mountains=2, mountain_seed=6,
                                            pert_seed=300 + m, pert_rms=5.0), member=m)
            out = [d.step(n)[:3] for n in chunks]
            states = [d.get_state(member=m) for m in range(members)]
    finally:
        if eager:
            if old is None:
                del os.environ["GD_NO_GRAPH"]
            else:
                os.environ["GD_NO_GRAPH"] = old
    mass, energy, beta = (np.concatenate([o[k] for o in out]) for k in range(3))
    return states, mass, energy, beta


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_multi_step_graph_gives_the_bits_of_single_steps(gd, prec):
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=1e5, precision=prec)
    print(f"step graphs {prec}", flush=True)
    a = run(gd, [21], members=3, **kw)              # two eight-step replays and five single ones
    b = run(gd, [1] * 21, members=3, **kw)          # single-step replays
    c = run(gd, [21], eager=True, members=3, **kw)  # no graphs
    for x in (b, c):
        for sa, sx in zip(a[0], x[0]):
            for f, g in zip(sa, sx):
                np.testing.assert_array_equal(f, g)
        for k in (1, 2, 3):
            np.testing.assert_array_equal(a[k], x[k])
