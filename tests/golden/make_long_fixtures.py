"""Writes the long-run fixtures of the benchmark configurations: the FP64
oracle (oracle/gd_oracle.c) stepped from the seeded initial states of the
1-degree 1-day run, the 0.5-degree 10-day run, the 0.25-degree run over 6
hours (the 2-day case is defined too; it takes the oracle many hours) and
members 0 and 63 of the 64-member 1-degree 5-day ensemble.  Each fixture
holds the case's parameters, the per-step total mass and energy, and the
final u, v, gd (every `stride`-th row and column), so the GPU tests step the
device at full size and compare without running the oracle.

    python tests/golden/make_long_fixtures.py [case ...]      (CPU, minutes to
                                                               hours per case)

The initial states come from the library's host-side generator
(gd_ic_balanced), the filter tables from gd_cutoff_from_lat0; neither needs a
GPU.  The run loop is that of dycore_run (dycore_mod.F90:131-140).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OPTS = dict(split_scheme="csp2", subcycles=6, qcon_modified=True, uv_adv_scheme="center_diff",
            use_diffusion=True, diffusion_order=2)
LAT0 = 75.0
BAL = dict(slope=-3.0, rms_height=150.0, mean_depth=8000.0)

# name: grid, time step, diffusion coefficient, steps, IC, subsampling stride
CASES = {
    "360x181_1day": dict(N=360, M=181, dt=240.0, coef=1.0e5, steps=360, stride=1,
                         ic=dict(seed=20240, nmax=20, **BAL)),
    "720x361_10day": dict(N=720, M=361, dt=120.0, coef=2.5e4, steps=7200, stride=4,
                          ic=dict(seed=20241, nmax=20, n_mountains=3, mountain_seed=777, **BAL)),
    # the 0.25-degree run over 6 hours (360 steps); the full 2-day run is
    # 2880 oracle steps at 4-12 s each on 8 CPU cores (many hours)
    "1440x721_6h": dict(N=1440, M=721, dt=60.0, coef=6.25e3, steps=360, stride=8,
                        ic=dict(seed=20242, nmax=80, **BAL)),
    "1440x721_2day": dict(N=1440, M=721, dt=60.0, coef=6.25e3, steps=2880, stride=8,
                          ic=dict(seed=20242, nmax=80, **BAL)),
    "ens360_5day_m0": dict(N=360, M=181, dt=240.0, coef=1.0e5, steps=1800, stride=2, members=64,
                           member=0, ic=dict(seed=20243, nmax=20, pert_seed=5000, pert_nmax=20,
                                             pert_rms=5.0, **BAL)),
    "ens360_5day_m63": dict(N=360, M=181, dt=240.0, coef=1.0e5, steps=1800, stride=2, members=64,
                            member=63, ic=dict(seed=20243, nmax=20, pert_seed=5063, pert_nmax=20,
                                               pert_rms=5.0, **BAL)),
}


def fixture_path(name):
    return os.path.join(HERE, f"long_{name}.npz")


def make(name):
    import __graft_entry__ as ge
    from oracle import oracle

    gd = ge.load_package()
    c = CASES[name]
    N, M, s = c["N"], c["M"], c["stride"]
    u, v, h, ghs = gd.ic_balanced(N, M, **c["ic"])
    cut = gd.cutoff_from_lat0(N, M, LAT0)
    oracle.set_threads(min(8, os.cpu_count() or 1))
    o = oracle.Oracle(N, M, c["dt"], cutoff=cut, diffusion_coef=c["coef"], **OPTS)
    o.set_initial_state(u, v, h, ghs)
    m0, e0 = o.start()
    t0 = time.time()
    mass, energy, beta, ns = o.run(c["steps"])
    assert ns == 0, f"{name}: the oracle reached NaN at step {ns}"
    fu, fv, fg = (o.state(o.old, f) for f in (oracle.U, oracle.V, oracle.GD))
    params = dict(c, lat0=LAT0, opts=OPTS, cutoff=cut)
    np.savez_compressed(fixture_path(name), params=json.dumps(params), mass0=m0, energy0=e0,
                        mass=mass, energy=energy, beta=beta, u=fu[::s, ::s], v=fv[::s, ::s],
                        gd=fg[::s, ::s])
    print(f"{name}: {c['steps']} steps in {time.time() - t0:.0f} s, "
          f"{os.path.getsize(fixture_path(name)) / 1e6:.2f} MB", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or [c for c in CASES if c != "1440x721_2day"]:
        make(n)
