"""Benchmark: model-days/sec of the barotropic step at 720x361 (0.5 deg),
dt = 120 s, 10 model days (7200 steps), csp2 with 6 subcycles, polar filter
and diffusion on; mixed precision (the headline value) against plain FP64 on
the same GPU, plus p50/p99 run latency over repeated full runs.

Prints ONE JSON line (rank 0).  A "step" is one time step of the path over
the whole grid (time_integrate + diag_run, dycore_mod.F90:131-140).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("model-days/sec at 720x361 (mixed vs FP64 plain, same GPU); "
          "p50/p99 run latency")

# The 720x361 (0.5 degree) 10-day workload; values not fixed by the workload
# description are stated here and in namelists/cfg2_720x361.nml.
WORKLOAD = dict(num_lon=720, num_lat=361, time_step_size=120.0, split_scheme="csp2",
                subcycles=6, qcon_modified=True, uv_adv_scheme="center_diff",
                use_diffusion=True, diffusion_order=2, diffusion_coef=2.5e4,
                use_zonal_tend_filter=True)
FILTER_LAT0 = 75.0
IC = dict(seed=20241, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0,
          n_mountains=3, mountain_seed=777)
MODEL_DAYS = 10


def tend_launch_bytes(N, M, prec):
    """Algorithmic HBM bytes of one launch of the tendency kernel in the
    second predictor of a fast-pass call (S2 = S0 + dt/2 L(S1), keep Tb):
    read S1 (gd, U, V), ghs and S0 (gd, U, V); write S2 (gd, U, V) and Tb
    (du, dv, dgd).  gd, U, u live on N x M full rows, V on N x (M-1) half
    rows; state and stored tendencies are FP64 in both precision modes."""
    full, half = N * M, N * (M - 1)
    state = 8 * (full + full + half)            # gd, U, V
    tend = 8 * (full + half) + 8 * full         # du, dv, dgd
    return state + 8 * full + state + state + tend


def traffic_from_profile(kernel):
    """HBM bytes per launch of the same kernel launch measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a 256 MiB copy), as
    committed in profiles/<kernel>_720x361_traffic.txt; None when the
    summary is absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_720x361_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return int(float(vals["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def make_dycore(gd, prec, members=1):
    cut = gd.cutoff_from_lat0(WORKLOAD["num_lon"], WORKLOAD["num_lat"], FILTER_LAT0)
    return gd.Dycore(zonal_tend_filter_cutoff_wavenumber=cut, precision=prec, members=members,
                     **WORKLOAD), cut


def run_gpu(gd, prec, ic, steps, warmup, repeats):
    d, cut = make_dycore(gd, prec)
    u, v, h, ghs = ic
    d.set_state(u, v, h, ghs)
    if warmup:
        d.step(warmup)
    d.set_state(u, v, h, ghs)
    t0 = time.perf_counter()
    mass, energy, beta, _ = d.step(steps)   # returns after a stream synchronize
    t1 = time.perf_counter()
    elapsed = t1 - t0
    full = MODEL_DAYS * 86400 / WORKLOAD["time_step_size"]
    lat = []
    for _ in range(repeats):
        d.set_state(u, v, h, ghs)
        s0 = time.perf_counter()
        d.step(int(full))
        lat.append(time.perf_counter() - s0)
    prof = d.profile_kernels(60)
    fused = d.fused_phases()
    kern_us = d.time_kernel("phase" if fused else "tend", 200)
    d.close()
    return dict(elapsed=elapsed, mass=mass[:, 0], energy=energy[:, 0], lat=lat, prof=prof, cut=cut,
                kern_us=kern_us, fused=fused)


def run_batched(gd, prec, ic, members, steps):
    """Aggregate model-days/s of `members` runs of the workload batched in one
    handle (members stepped together by the captured step graph)."""
    d, _ = make_dycore(gd, prec, members=members)
    for m in range(members):
        d.set_state(*ic, member=m)
    d.step(10)
    for m in range(members):
        d.set_state(*ic, member=m)
    t0 = time.perf_counter()
    d.step(steps)
    el = time.perf_counter() - t0
    d.close()
    return members * steps * WORKLOAD["time_step_size"] / 86400.0 / el, el


def pmc_probe(gd, ic, nsteps):
    """Workload of the hardware-counter runs: a plain device-to-device copy of
    256 MiB through the HIP runtime (its FETCH_SIZE / WRITE_SIZE per dispatch
    calibrate the byte counters), then nsteps steps of the mixed workload.
    Run with GD_NO_GRAPH=1 so each kernel is its own dispatch.  The process
    does not import torch: counter collection then sees every dispatch."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    print(f"copy calibration: {nbytes} bytes read and written", flush=True)
    hip.hipFree(src)
    hip.hipFree(dst)
    d, _ = make_dycore(gd, "mixed")
    d.set_state(*ic)
    d.step(nsteps)
    d.close()
    print(f"pmc probe: {nsteps} steps done", flush=True)


# ---- the other workloads (run with --workload; not the default line) -------

def _cfg(gd, N, M, dt, coef, prec, members=1):
    cut = gd.cutoff_from_lat0(N, M, FILTER_LAT0)
    kw = dict(WORKLOAD, num_lon=N, num_lat=M, time_step_size=dt, diffusion_coef=coef)
    return gd.Dycore(zonal_tend_filter_cutoff_wavenumber=cut, precision=prec, members=members, **kw)


def _time_steps(d, steps, warmup):
    d.step(warmup)
    t0 = time.perf_counter()
    d.step(steps)
    return time.perf_counter() - t0


def _kernel_us(d, reps):
    """Replayed duration of the fast-pass second-predictor phase launch:
    k_phase when the handle fuses phases, else k_tend."""
    fused = d.fused_phases()
    return d.time_kernel("phase" if fused else "tend", reps), fused


def _tend_roofline(N, M, members, us, fused):
    """HBM roofline of the fast-pass second-predictor phase launch over all
    members (its algorithmic bytes over its replayed duration)."""
    nbytes = tend_launch_bytes(N, M, "mixed") * members
    achieved = nbytes / (us * 1e-6) / 1e9
    kname = "k_phase" if fused else "k_tend"
    return {"bound": "hbm", "kernel": f"{kname}<float, fast pass> (second predictor)",
            "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4), "traffic": None, "bytes_per_launch": int(nbytes),
            "avg_launch_us": round(us, 3)}


def other_workload(gd, args):
    base = {"n_gpus": 1, "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None,
            "dtype": "f32/f64", "roofline": None, "cpu_baseline": None}
    if args.workload == "ensemble":
        # 64 members at 360x181, dt 240 s, 5 model days (1800 steps); members
        # share the topography-free balanced state plus their own rms-5 m
        # perturbation
        N, M, dt, mem = 360, 181, 240.0, 64
        steps = min(args.steps, 1800)
        ics = [gd.ic_balanced(N, M, seed=20243, nmax=20, pert_seed=5000 + k, pert_nmax=20, pert_rms=5.0)
               for k in range(mem)]
        out, tend_us = {}, {}
        for prec in ("mixed", "fp64"):
            d = _cfg(gd, N, M, dt, 1.0e5, prec, members=mem)
            for k, ic in enumerate(ics):
                d.set_state(*ic, member=k)
            el = _time_steps(d, steps, args.warmup)
            tend_us[prec] = _kernel_us(d, 50)
            d.close()
            out[prec] = mem * steps * dt / 86400.0 / el
        base["roofline"] = _tend_roofline(N, M, mem, *tend_us["mixed"])
        return dict(base, metric="member-days/sec, 64-member ensemble at 360x181 (mixed vs FP64)",
                    value=round(out["mixed"], 3), unit="member-days/s", steps=steps,
                    ms_per_step=round(mem * steps * dt / 86400.0 / out["mixed"] / steps * 1e3, 4),
                    scaling="weak", data="synthetic (seeded balanced state + per-member perturbations)",
                    config={"workload": "ensemble 64 x 360x181, dt=240 s, csp2 x6, diffusion 1e5, "
                                        f"filter cos(lat)/cos({FILTER_LAT0:g})",
                            "fp64_member_days_per_s": round(out["fp64"], 3),
                            "mixed_over_fp64": round(out["mixed"] / out["fp64"], 3)})
    if args.workload == "1440":
        N, M, dt = 1440, 721, 60.0
        steps = min(args.steps, 2880)
        ic = gd.ic_balanced(N, M, seed=20242, nmax=80)
        out, tend_us = {}, {}
        for prec in ("mixed", "fp64"):
            d = _cfg(gd, N, M, dt, 6.25e3, prec)
            d.set_state(*ic)
            el = _time_steps(d, steps, args.warmup)
            tend_us[prec] = _kernel_us(d, 100)
            d.close()
            out[prec] = steps * dt / 86400.0 / el
        base["roofline"] = _tend_roofline(N, M, 1, *tend_us["mixed"])
        return dict(base, metric="model-days/sec at 1440x721 (mixed vs FP64)",
                    value=round(out["mixed"], 4), unit="model-days/s", steps=steps,
                    ms_per_step=round(dt / 86400.0 / out["mixed"] * 1e3, 4), scaling="weak",
                    data="synthetic (seeded balanced state, n <= 80)",
                    config={"workload": "1440x721 (0.25 deg), dt=60 s, csp2 x6, diffusion 6.25e3, "
                                        f"filter cos(lat)/cos({FILTER_LAT0:g})",
                            "fp64_model_days_per_s": round(out["fp64"], 4),
                            "mixed_over_fp64": round(out["mixed"] / out["fp64"], 3)})
    return serving_workload(gd, args, base)


# (num_lon, num_lat, dt, diffusion): diffusion 1e5 at 1 deg and above, scaled
# with the grid spacing squared below (4e5 at 2 deg is unstable next to the
# poles: the oracle reaches NaN at step 124)
SERVE_CLASSES = [(180, 91, 480.0, 1.0e5), (360, 181, 240.0, 1.0e5), (720, 361, 120.0, 2.5e4)]


def serving_workload(gd, args, base):
    """Poisson arrivals of runs with resolution uniform over 180x91, 360x181,
    720x361 (dt 480/240/120 s, diffusion of SERVE_CLASSES)
    and length uniform over 1..10 model days, seeded balanced states.  The
    arrival rate is 0.8 / (mean single-run service time), the service times
    measured alone first on the same GPU."""
    rng = np.random.default_rng(4242)
    nreq = max(4, args.serve_requests)
    # single-run ms/step of each class (mixed), for the arrival rate
    ms = {}
    for (N, M, dt, coef) in SERVE_CLASSES:
        d = _cfg(gd, N, M, dt, coef, "mixed")
        d.set_state(*gd.ic_balanced(N, M, seed=1, nmax=20))
        ms[(N, M)] = _time_steps(d, 200, 20) / 200 * 1e3
        d.close()
    reqs = []
    t = 0.0
    cls = rng.integers(0, 3, nreq)
    days = rng.integers(1, 11, nreq)
    service = [days[k] * 86400 / SERVE_CLASSES[cls[k]][2] * ms[SERVE_CLASSES[cls[k]][:2]] / 1e3
               for k in range(nreq)]
    rate = 0.8 / float(np.mean(service))
    for k in range(nreq):
        N, M, dt, coef = SERVE_CLASSES[cls[k]]
        reqs.append(dict(num_lon=N, num_lat=M, time_step_size=dt, diffusion_coef=coef,
                         nsteps=int(days[k] * 86400 / dt), arrival_s=t,
                         state=gd.ic_balanced(N, M, seed=9000 + k, nmax=20)))
        t += float(rng.exponential(1.0 / rate))
    out = {}
    for prec in ("mixed", "fp64"):
        res, wall = gd.serve(reqs, precision=prec, filter_lat0=FILTER_LAT0, max_slots=16,
                             chunk_steps=60, split_scheme="csp2", subcycles=6)
        assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
        ttr = np.array([r["finish_s"] - r["arrival_s"] for r in res])
        total_days = float(np.sum(days))
        out[prec] = dict(value=total_days / wall, wall=wall, p50=float(np.percentile(ttr, 50)),
                         p99=float(np.percentile(ttr, 99)))
    m = out["mixed"]
    return dict(base, metric="aggregate model-days/sec of a Poisson stream of runs "
                             "(180x91/360x181/720x361, 1-10 days); p50/p99 time-to-result",
                value=round(m["value"], 3), unit="model-days/s", steps=None, ms_per_step=None,
                scaling="weak", data="synthetic (seeded balanced states)",
                config={"workload": f"serving stream of {nreq} runs, rate {rate:.3f}/s "
                                    "(0.8 of the mean single-run service time)",
                        "requests": nreq, "model_days": int(np.sum(days)),
                        "single_run_ms_per_step": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in ms.items()},
                        "fp64_model_days_per_s": round(out["fp64"]["value"], 3),
                        "mixed_over_fp64": round(m["value"] / out["fp64"]["value"], 3)},
                latency_s={"p50": round(m["p50"], 4), "p99": round(m["p99"], 4),
                           "fp64_p50": round(out["fp64"]["p50"], 4),
                           "fp64_p99": round(out["fp64"]["p99"], 4),
                           "stream_wall_s": round(m["wall"], 3)})


def cpu_baseline(ic, nsteps):
    """The FP64 C oracle built with the reference's optimisation level
    (-Ofast), single thread (the reference is serial), on a bounded sample of
    the same workload: the first nsteps steps."""
    from oracle import oracle as orc
    orc.use_timing_build()
    Oracle, set_threads = orc.Oracle, orc.set_threads
    set_threads(1)
    cut = None
    import importlib
    gd = importlib.import_module("gamil_dycore_amd")
    cut = gd.cutoff_from_lat0(WORKLOAD["num_lon"], WORKLOAD["num_lat"], FILTER_LAT0)
    kw = {k: v for k, v in WORKLOAD.items() if k not in ("num_lon", "num_lat", "time_step_size")}
    o = Oracle(WORKLOAD["num_lon"], WORKLOAD["num_lat"], WORKLOAD["time_step_size"],
               split_scheme=kw["split_scheme"], subcycles=kw["subcycles"],
               qcon_modified=kw["qcon_modified"], uv_adv_scheme=kw["uv_adv_scheme"],
               use_diffusion=kw["use_diffusion"], diffusion_order=kw["diffusion_order"],
               diffusion_coef=kw["diffusion_coef"], use_zonal_tend_filter=True, cutoff=cut)
    u, v, h, ghs = ic
    o.set_initial_state(u, v, h, ghs)
    o.start()
    t0 = time.perf_counter()
    o.run(nsteps)
    el = time.perf_counter() - t0
    return nsteps * WORKLOAD["time_step_size"] / 86400.0 / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7200)
    ap.add_argument("--warmup", type=int, default=360)
    ap.add_argument("--repeats", type=int, default=5, help="full 10-day runs for p50/p99 latency")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-members", type=int, default=8,
                    help="members of the batched mixed-precision form timed beside the single run")
    ap.add_argument("--batch-steps", type=int, default=720)
    ap.add_argument("--workload", default="cfg2", choices=["cfg2", "ensemble", "1440", "serving"],
                    help="cfg2: the 720x361 single run (default); the others time the "
                         "ensemble (64 x 360x181), 1440x721 and serving-stream workloads")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--pmc-probe", type=int, default=0, metavar="K",
                    help="counter runs: one 256 MiB device copy (calibration of the byte "
                         "counters) then K steps of the mixed workload, no timing output")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed, not built)"}))
        sys.exit(2)
    import __graft_entry__ as ge
    gd = ge.load_package()
    N, M, dt = WORKLOAD["num_lon"], WORKLOAD["num_lat"], WORKLOAD["time_step_size"]
    ic = gd.ic_balanced(N, M, **IC)
    if args.pmc_probe:
        pmc_probe(gd, ic, args.pmc_probe)
        return
    if args.workload != "cfg2":
        print(json.dumps(other_workload(gd, args)))
        return
    res = {}
    for prec in ("mixed", "fp64"):
        res[prec] = run_gpu(gd, prec, ic, args.steps, args.warmup, args.repeats)
    batched = None
    if args.batch_members > 1:
        bv, bel = run_batched(gd, "mixed", ic, args.batch_members, args.batch_steps)
        batched = (bv, bel)
    days = args.steps * dt / 86400.0
    val = {p: days / r["elapsed"] for p, r in res.items()}
    mixed = res["mixed"]
    # roofline of the dominant kernel: one fast-pass second-predictor phase
    # launch (k_phase: tendency, filter rows and epilogue in one launch; or
    # k_tend when phases are not fused), its algorithmic bytes over its
    # duration replayed back to back (HIP events on the library's stream).
    # Fusing moves no extra bytes: the filter rows stay in LDS.
    prof = mixed["prof"]
    dom = max(prof, key=lambda k: prof[k][1])
    nbytes = tend_launch_bytes(N, M, "mixed")
    avg_s = mixed["kern_us"] * 1e-6
    achieved = nbytes / avg_s / 1e9
    kname = "k_phase" if mixed["fused"] else "k_tend"
    roofline = {"bound": "hbm",
                "kernel": f"{kname}<float, fast pass> (second predictor" +
                          (": tendency + filter rows + epilogue)" if mixed["fused"] else ")"),
                "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": traffic_from_profile(kname),
                "bytes_per_launch": int(nbytes),
                "avg_launch_us": round(mixed["kern_us"], 3),
                "dominant_family_by_time": dom,
                "family_ms_per_step": {k: round(v[1] / 60.0, 4) for k, v in prof.items()}}
    cpu = None
    if not args.no_cpu:
        cv, cel = cpu_baseline(ic, args.cpu_steps)
        cpu = {"value": round(cv, 6), "unit": "model-days/s", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same 720x361 FP64 workload, "
                         f"oracle/gd_oracle.c built -O3 -ffast-math (the reference's -Ofast), "
                         f"single thread ({cel:.1f} s)"}
    lat = np.array(mixed["lat"]) if mixed["lat"] else np.array([np.nan])
    lat64 = np.array(res["fp64"]["lat"]) if res["fp64"]["lat"] else np.array([np.nan])
    out = {
        "metric": METRIC,
        "value": round(val["mixed"], 4),
        "unit": "model-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["elapsed"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded balanced spherical-harmonic IC + 3 Gaussian mountains)",
        "config": {"workload": "barotropic step 720x361 (0.5 deg), dt=120 s, 10 model days, "
                               "csp2 x6, qcon, center_diff, diffusion 2.5e4, polar filter "
                               f"cos(lat)/cos({FILTER_LAT0:g})",
                   "grid": "720x361", "model_days_timed": round(days, 3),
                   "precision": "mixed", "fp64_model_days_per_s": round(val["fp64"], 4),
                   "mixed_over_fp64": round(val["mixed"] / val["fp64"], 3),
                   "filtered_rows_per_hemisphere": len(mixed["cut"]),
                   "batched_mixed": None if batched is None else {
                       "members": args.batch_members, "steps": args.batch_steps,
                       "aggregate_model_days_per_s": round(batched[0], 4),
                       "over_fp64_single_run": round(batched[0] / val["fp64"], 3)},
                   "parallelism": "single GPU, single run"},
        "latency_s": {"runs": int(len(mixed["lat"])), "p50": float(np.percentile(lat, 50)),
                      "p99": float(np.percentile(lat, 99)),
                      "fp64_p50": float(np.percentile(lat64, 50)),
                      "fp64_p99": float(np.percentile(lat64, 99))},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
