"""Serving (gd_serve): continuous batching of runs of uneven resolution and
length.  Each result must be identical alone and in the mix, and identical to
the same run on its own handle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = dict(precision="fp64", split_scheme="csp2", subcycles=3, use_diffusion=True,
            filter_lat0=75.0)


def fnv1a(arrays):
    h = 1469598103934665603
    for a in arrays:
        for b in np.ascontiguousarray(a, dtype=np.float64).tobytes():
            h ^= b
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def request(gd, N, M, dt, coef, nsteps, arrival, seed):
    ic = gd.ic_balanced(N, M, seed=seed, nmax=8, n_mountains=1, mountain_seed=seed + 1)
    return dict(num_lon=N, num_lat=M, time_step_size=dt, diffusion_coef=coef, nsteps=nsteps,
                arrival_s=arrival, state=ic)


def mix(gd):
    return [request(gd, 72, 37, 1200.0, 4e5, 30, 0.00, 1),
            request(gd, 144, 73, 600.0, 1e5, 40, 0.00, 2),
            request(gd, 72, 37, 1200.0, 4e5, 45, 0.01, 3),
            request(gd, 96, 49, 900.0, 2e5, 25, 0.02, 4),
            request(gd, 144, 73, 600.0, 1e5, 20, 0.03, 5),
            request(gd, 72, 37, 1200.0, 4e5, 12, 0.04, 6)]


def test_serve_alone_equals_mix_equals_handle(gd):
    reqs = mix(gd)
    print("serving the mix", flush=True)
    res, wall = gd.serve(reqs, max_slots=2, chunk_steps=8, **OPTS)
    assert wall > 0
    for k, (q, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (k, r)
        assert r["arrival_s"] <= r["start_s"] <= r["finish_s"]
        print("alone", k, flush=True)
        alone, _ = gd.serve([dict(q, arrival_s=0.0)], max_slots=2, chunk_steps=8, **OPTS)
        a = alone[0]
        assert a["status"] == 0
        assert a["state_hash"] == r["state_hash"], k
        assert a["total_mass"] == r["total_mass"] and a["total_energy"] == r["total_energy"], k
    # the same run on its own handle
    q = reqs[1]
    N, M = q["num_lon"], q["num_lat"]
    with gd.Dycore(N, M, q["time_step_size"], split_scheme="csp2", subcycles=3, use_diffusion=True,
                   diffusion_coef=q["diffusion_coef"],
                   zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(N, M, 75.0)) as d:
        d.set_state(*q["state"])
        mass, energy, _, _ = d.step(q["nsteps"])
        assert fnv1a(d.get_state()) == res[1]["state_hash"]
        assert mass[-1, 0] == res[1]["total_mass"] and energy[-1, 0] == res[1]["total_energy"]


def test_serve_queues_when_slots_are_full(gd):
    """One slot per class: runs of a class follow each other in arrival order."""
    reqs = [request(gd, 72, 37, 1200.0, 4e5, 10, 0.0, s) for s in (7, 8, 9)]
    res, _ = gd.serve(reqs, max_slots=1, chunk_steps=4, **OPTS)
    assert all(r["status"] == 0 and r["slot"] == 0 for r in res)
    assert res[0]["finish_s"] <= res[1]["start_s"] and res[1]["finish_s"] <= res[2]["start_s"]
    assert len({r["state_hash"] for r in res}) == 3


def test_serve_mixed_precision(gd):
    reqs = mix(gd)[:3]
    res, _ = gd.serve(reqs, max_slots=4, chunk_steps=16, **dict(OPTS, precision="mixed"))
    ref, _ = gd.serve(reqs, max_slots=4, chunk_steps=16, **OPTS)
    for r, f in zip(res, ref):
        assert r["status"] == 0
        assert abs(r["total_mass"] - f["total_mass"]) <= 1e-12 * abs(f["total_mass"])
        assert abs(r["total_energy"] - f["total_energy"]) <= 1e-5 * abs(f["total_energy"])
