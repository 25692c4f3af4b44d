"""The benchmark configurations at full size against the FP64 oracle.

The fixtures (tests/golden/long_*.npz, written by make_long_fixtures.py on
the CPU) hold the oracle's per-step mass and energy and its final state after
the configuration's run: 360x181 for 1 day, 720x361 for 10 days over three
mountains, 1440x721 for 6 hours (spectrum to n = 80; the oracle's 2-day run
takes many CPU hours) and members 0 and 63 of the 64-member 360x181 ensemble
after 5 days.  The device runs the same
seeded states through the C ABI in FP64 and in mixed precision (the ensemble
as one 64-member handle).

Tolerances: FP64 agrees with the oracle to round-off growth (height l2 1e-12,
winds 1e-10, mass and energy 1e-12 relative); mixed, and mixed with the FP32
matrix-core momentum filter, meet the stated target, height l2 <= 1e-5
relative with mass drift <= 1e-12 of the initial mass, and energy within
1e-8 of the oracle's; with the BF16 matrix-core filter the height l2 bound is
1e-4 (mass drift still <= 1e-12: the mass rows keep the FP64 filter).  Full-size properties with neither
filter nor diffusion: mass conserved to round-off, and with the
energy-conserving beta the quadratic energy too.
"""
import json
import os

import numpy as np
import pytest

from conftest import l2_rel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["360x181_1day", "720x361_10day", "1440x721_6h", "ens360_5day_m0", "ens360_5day_m63"]


def _fixture(name):
    path = os.path.join(GOLDEN, f"long_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not written (tests/golden/make_long_fixtures.py)")
    f = np.load(path)
    return json.loads(str(f["params"])), f


_RUNS = {}


def _device_run(gd, p, prec):
    """The device's final state of member p['member'] and its mass/energy
    series; cached per (case, precision) for the two members of the ensemble."""
    key = (p["N"], p["M"], p["steps"], p.get("members", 1), json.dumps(p["ic"], sort_keys=True)
           if "members" not in p else p["ic"]["seed"], prec)
    if key in _RUNS:
        return _RUNS[key][p.get("member", 0)]
    N, M, members = p["N"], p["M"], p.get("members", 1)
    kw = dict(p["opts"], diffusion_coef=p["coef"], zonal_tend_filter_cutoff_wavenumber=p["cutoff"],
              precision=prec, members=members)
    print(f"device run {N}x{M} x{members} {p['steps']} steps {prec}", flush=True)
    out = {}
    with gd.Dycore(N, M, p["dt"], **kw) as d:
        for m in range(members):
            ic = dict(p["ic"])
            if members > 1:
                ic["pert_seed"] = 5000 + m
            d.set_state(*gd.ic_balanced(N, M, **ic), member=m)
        m0, e0 = d.diag()
        mass, energy, _, _ = d.step(p["steps"])
        keep = [p.get("member", 0)] if members == 1 else [0, 63]
        for m in keep:
            out[m] = (d.get_state(member=m), m0[m], mass[:, m], energy[:, m])
    _RUNS[key] = out
    return out[p.get("member", 0)]


PRECS = ["fp64", "mixed", "mixed_mfma", "mixed_bf16"]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", CASES)
def test_config_vs_oracle(gd, name, prec):
    if prec.startswith("mixed_") and name == "360x181_1day":
        pytest.skip("the matrix-core filter modes run on the other cases")
    if prec == "mixed_mfma" and name == "720x361_10day":
        # 7200 steps at about 17 ms (one run: 15 of 16 matrix-core rows are
        # padding) take two minutes; the mode runs at full size on the
        # 64-member ensemble cases, and the BF16 mode on this one (measured
        # here once: height l2 1.6e-7, mass drift 0)
        pytest.skip("FP32 matrix-core filter: full-size run on the ensemble cases")
    p, f = _fixture(name)
    s = p["stride"]
    (u, v, h), m0, mass, energy = _device_run(gd, p, prec)
    ou, ov, oh = f["u"], f["v"], f["gd"]
    eh, eu, ev = l2_rel(h[::s, ::s], oh), l2_rel(u[::s, ::s], ou), l2_rel(v[::s, ::s], ov)
    linf_h = np.abs(h[::s, ::s] - oh).max() / np.abs(oh).max()
    dm = abs(mass[-1] - m0) / m0
    de = np.abs(energy - f["energy"]).max() / abs(float(f["energy0"]))
    om = np.abs(mass - f["mass"]).max() / float(f["mass0"])
    print(f"{name} {prec}: gd l2 {eh:.2e} linf {linf_h:.2e}, u l2 {eu:.2e}, v l2 {ev:.2e}, "
          f"mass drift {dm:.1e} (oracle series {om:.1e}), energy vs oracle {de:.1e}", flush=True)
    assert np.all(np.isfinite(h))
    assert abs(m0 - float(f["mass0"])) <= 1e-13 * float(f["mass0"])
    assert dm <= 1e-12
    if prec == "fp64":
        assert eh <= 1e-12 and eu <= 1e-10 and ev <= 1e-10
        assert om <= 1e-12 and de <= 1e-12
    elif prec == "mixed_bf16":
        # BF16 split operands in the momentum filter: stated, looser
        assert eh <= 1e-4 and eu <= 1e-3 and ev <= 1e-3
    else:
        assert eh <= 1e-5 and linf_h <= 1e-4
        assert eu <= 1e-4 and ev <= 1e-4
        assert de <= 1e-8


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_size_conservation_without_filter_and_diffusion(gd, prec):
    """720x361, 60 steps with neither filter nor diffusion: the flux form
    conserves mass to round-off and the beta-corrected integrator conserves
    the quadratic energy (dycore_mod.F90:754-792) to round-off growth."""
    N, M = 720, 361
    print(f"conservation 720x361 {prec}", flush=True)
    with gd.Dycore(N, M, 20.0, split_scheme="csp2", subcycles=2, use_zonal_tend_filter=False,
                   precision=prec) as d:
        d.set_state(*gd.ic_balanced(N, M, seed=20241, nmax=20))
        m0, e0 = d.diag()
        mass, energy, beta, _ = d.step(60)
    assert np.abs(mass[:, 0] - m0[0]).max() <= 1e-13 * m0[0]
    tol = 1e-12 if prec == "fp64" else 1e-8
    assert np.abs(energy[:, 0] - e0[0]).max() <= tol * e0[0]
    assert np.all(beta[:, 0] != 1.0)
