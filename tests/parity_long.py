"""Long-run parity report: the device after N model days against the FP64
oracle from the same seeded state (height and wind normalised l2 / l-inf
errors, mass and total-energy drift).  Not collected by pytest: the oracle
needs minutes to hours of CPU for these runs, so the two halves run apart.

  python tests/parity_long.py oracle --workload 720x361 --days 10 --out FILE
      (CPU) runs the FP64 oracle and saves its final state and per-step
      mass/energy to FILE (.npz)
  python tests/parity_long.py device --workload 720x361 --days 10 --oracle FILE
      (GPU) runs the device in FP64 and mixed precision and prints one JSON
      line per precision with the errors against FILE

Workloads: 720x361 = the 0.5-degree 10-day benchmark workload (3 mountains,
dt 120 s); 360x181 = the 1-degree flat-topography workload (dt 240 s);
ens360 = member --member of the 64-member 360x181 ensemble (common balanced
state plus its own rms-5 m perturbation; the device runs the whole ensemble
and the member is compared).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "720x361": dict(N=720, M=361, dt=120.0, coef=2.5e4,
                    ic=dict(seed=20241, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0,
                            n_mountains=3, mountain_seed=777)),
    "360x181": dict(N=360, M=181, dt=240.0, coef=1.0e5,
                    ic=dict(seed=20240, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0)),
    "ens360": dict(N=360, M=181, dt=240.0, coef=1.0e5, members=64,
                   ic=dict(seed=20243, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0,
                           pert_nmax=20, pert_rms=5.0)),
}
LAT0 = 75.0
OPTS = dict(split_scheme="csp2", subcycles=6, qcon_modified=True, uv_adv_scheme="center_diff",
            use_diffusion=True, diffusion_order=2)


def member_ic(gd, w, member):
    kw = dict(w["ic"])
    if "members" in w:
        kw["pert_seed"] = 5000 + member
    return gd.ic_balanced(w["N"], w["M"], **kw)


def setup(name, member=0):
    import __graft_entry__ as ge
    gd = ge.load_package()
    w = WORKLOADS[name]
    ic = member_ic(gd, w, member)
    cut = gd.cutoff_from_lat0(w["N"], w["M"], LAT0)
    return gd, w, ic, cut


def run_oracle(args):
    from oracle.oracle import Oracle, set_threads, U, V, GD
    gd, w, ic, cut = setup(args.workload, args.member)
    set_threads(args.threads)
    steps = int(round(args.days * 86400 / w["dt"]))
    o = Oracle(w["N"], w["M"], w["dt"], cutoff=cut, diffusion_coef=w["coef"], **OPTS)
    o.set_initial_state(*ic)
    o.start()
    mass, energy = [], []
    t0 = time.time()
    done = 0
    while done < steps:
        n = min(args.chunk, steps - done)
        m, e, _, ns = o.run(n)
        if ns:
            raise SystemExit(f"oracle NaN at step {done + ns}")
        mass.append(m)
        energy.append(e)
        done += n
        print(f"oracle {done}/{steps} steps, {time.time() - t0:.0f} s", flush=True)
    np.savez(args.out, u=o.state(o.old, U), v=o.state(o.old, V), gd=o.state(o.old, GD),
             mass=np.concatenate(mass), energy=np.concatenate(energy),
             workload=args.workload, days=args.days, member=args.member)


def norms(a, b):
    d = a - b
    return {"l2": float(np.sqrt((d ** 2).sum() / (b ** 2).sum())),
            "linf": float(np.abs(d).max() / np.abs(b).max())}


def run_device(args):
    gd, w, ic, cut = setup(args.workload, args.member)
    ref = np.load(args.oracle, allow_pickle=False)
    steps = int(round(args.days * 86400 / w["dt"]))
    assert str(ref["workload"]) == args.workload and float(ref["days"]) == args.days
    assert (int(ref["member"]) if "member" in ref.files else 0) == args.member
    nm = w.get("members", 1)
    k = args.member if nm > 1 else 0
    for prec in ("fp64", "mixed"):
        with gd.Dycore(w["N"], w["M"], w["dt"], precision=prec, diffusion_coef=w["coef"],
                       zonal_tend_filter_cutoff_wavenumber=cut, members=nm, **OPTS) as d:
            for q in range(nm):
                d.set_state(*(ic if q == k else member_ic(gd, w, q)), member=q)
            m0, e0 = d.diag()
            mass, energy, _, _ = d.step(steps)
            u, v, h = d.get_state(member=k)
            m0, e0, mass, energy = m0[k:k + 1], e0[k:k + 1], mass[:, k:k + 1], energy[:, k:k + 1]
        rm, re = ref["mass"], ref["energy"]
        out = {
            "workload": args.workload, "days": args.days, "precision": prec, "member": args.member,
            "height": norms(h, ref["gd"]), "u": norms(u, ref["u"]), "v": norms(v, ref["v"]),
            "mass_drift_device": float((mass[-1, 0] - m0[0]) / m0[0]),
            "mass_drift_oracle": float((rm[-1] - m0[0]) / m0[0]),
            "mass_vs_oracle": float(abs(mass[-1, 0] - rm[-1]) / abs(rm[-1])),
            "energy_drift_device": float((energy[-1, 0] - e0[0]) / e0[0]),
            "energy_drift_oracle": float((re[-1] - e0[0]) / e0[0]),
            "energy_vs_oracle": float(abs(energy[-1, 0] - re[-1]) / abs(re[-1])),
            "max_mass_vs_oracle_per_step": float(np.max(np.abs(mass[:, 0] - rm) / np.abs(rm))),
        }
        print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["oracle", "device"])
    ap.add_argument("--workload", default="720x361", choices=sorted(WORKLOADS))
    ap.add_argument("--days", type=float, default=10.0)
    ap.add_argument("--out")
    ap.add_argument("--oracle")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=120)
    ap.add_argument("--member", type=int, default=0)
    args = ap.parse_args()
    run_oracle(args) if args.mode == "oracle" else run_device(args)


if __name__ == "__main__":
    main()
