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


def _alone_hash(gd, q, **opts):
    res, _ = gd.serve([dict(q, arrival_s=0.0)], **opts)
    return res[0]


def test_server_admits_runs_while_others_run(gd):
    """Continuous batching: a run submitted while a long run of its class is
    on the GPU is admitted into the same pool at the next chunk boundary and
    finishes first; both give the bits they give alone."""
    import time
    opts = dict(OPTS, max_slots=4, chunk_steps=10)
    long_q = request(gd, 360, 181, 240.0, 1e5, 400, 0.0, 21)
    short_q = request(gd, 360, 181, 240.0, 1e5, 20, 0.0, 22)
    print("server: long run, then a short one while it runs", flush=True)
    with gd.Server(**opts) as srv:
        a = srv.submit(360, 181, 240.0, 400, long_q["state"], diffusion_coef=1e5)
        time.sleep(0.05)
        b = srv.submit(360, 181, 240.0, 20, short_q["state"], diffusion_coef=1e5)
        assert srv.poll(a) is None  # still running
        rb = srv.poll(b, wait=True)
        ra = srv.poll(a, wait=True)
    assert ra["status"] == 0 and rb["status"] == 0
    assert rb["start_s"] < ra["finish_s"] and rb["finish_s"] < ra["finish_s"]
    assert rb["slot"] == 1  # the long run held slot 0 of the same pool
    assert ra["state_hash"] == _alone_hash(gd, long_q, **opts)["state_hash"]
    assert rb["state_hash"] == _alone_hash(gd, short_q, **opts)["state_hash"]


def test_server_evicts_a_nan_run(gd):
    """A run that blows up (winds 50x those of its balanced state) ends with
    GD_ERR_NAN, the reference's NaN guard (diag_mod.F90:79-87); the runs
    beside it in the pool keep the bits they have alone."""
    good = [request(gd, 180, 91, 480.0, 1e5, 120, 0.0, s) for s in (31, 32, 33)]
    u, v, g, hs = gd.ic_balanced(180, 91, seed=34, nmax=8)
    bad = dict(good[0], state=(u * 50.0, v * 50.0, g, hs))
    opts = dict(OPTS, max_slots=4, chunk_steps=10)
    print("server: nan run alone", flush=True)
    alone_bad = _alone_hash(gd, bad, **opts)
    assert alone_bad["status"] == gd.GD_ERR_NAN
    print("server: nan run in the mix", flush=True)
    res, _ = gd.serve([good[0], bad, good[1], good[2]], **opts)
    assert res[1]["status"] == gd.GD_ERR_NAN
    for q, r in zip(good, (res[0], res[2], res[3])):
        assert r["status"] == 0
        assert r["state_hash"] == _alone_hash(gd, q, **opts)["state_hash"]


def test_serve_benchmark_classes_alone_equal_mix(gd):
    """The serving stream's classes (180x91 / 360x181 / 720x361 at 480 / 240 /
    120 s, mixed precision, filter and diffusion on), short runs: every
    result is the same alone and in the mix."""
    classes = [(180, 91, 480.0, 1e5), (360, 181, 240.0, 1e5), (720, 361, 120.0, 2.5e4)]
    reqs = []
    for k, (n, steps) in enumerate([(0, 40), (1, 30), (2, 20), (0, 25), (2, 15), (1, 35)]):
        N, M, dt, coef = classes[n]
        reqs.append(request(gd, N, M, dt, coef, steps, 0.004 * k, 50 + k))
    opts = dict(OPTS, precision="mixed", subcycles=6, max_slots=3, chunk_steps=12)
    print("serving the benchmark classes", flush=True)
    res, _ = gd.serve(reqs, **opts)
    for k, (q, r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (k, r)
        assert r["state_hash"] == _alone_hash(gd, q, **opts)["state_hash"], k


def test_serve_budget_too_small_for_one_run(gd):
    reqs = [request(gd, 72, 37, 1200.0, 4e5, 5, 0.0, 61)]
    res, _ = gd.serve(reqs, hbm_budget_bytes=1024, **OPTS)
    assert res[0]["status"] == gd.GD_ERR_INVALID_ARG
