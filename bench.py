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

# The 64-member 1-degree ensemble (5 model days in the full workload): a
# common balanced state plus a seeded rms-5 m balanced perturbation per member.
ENS = dict(N=360, M=181, dt=240.0, coef=1.0e5, members=64,
           ic=dict(seed=20243, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0, pert_nmax=20,
                   pert_rms=5.0))


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


def traffic_from_profile(kernel, grid="720x361"):
    """HBM bytes per launch of the same kernel launch measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a 256 MiB copy), as
    committed in profiles/<kernel>_<grid>_traffic.txt; None when the
    summary is absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{grid}_traffic.txt")
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


def pmc_probe(gd, ic, nsteps, grid="720"):
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
    if grid == "360":  # the 1-degree single run (rows of two segments)
        d = _cfg(gd, 360, 181, 240.0, 1.0e5, "mixed")
        d.set_state(*gd.ic_balanced(360, 181, seed=20240, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0))
    elif grid == "1440":  # the 0.25-degree single run
        d = _cfg(gd, 1440, 721, 60.0, 6.25e3, "mixed")
        d.set_state(*gd.ic_balanced(1440, 721, seed=20242, nmax=80))
    elif grid == "ens64":  # the 64-member 360x181 ensemble, default launch shapes
        d = _cfg(gd, ENS["N"], ENS["M"], ENS["dt"], ENS["coef"], "mixed", members=ENS["members"])
        for k, icm in enumerate(_ensemble_ics(gd, ENS["N"], ENS["M"], ENS["members"])):
            d.set_state(*icm, member=k)
    else:
        d, _ = make_dycore(gd, "mixed")
        d.set_state(*ic)
    d.step(nsteps)
    d.close()
    print(f"pmc probe: {nsteps} steps of workload {grid} done", flush=True)


# ---- the other workloads: config.other_workloads of the default line, or one
# ---- line each with --workload

def _cfg(gd, N, M, dt, coef, prec, members=1):
    cut = gd.cutoff_from_lat0(N, M, FILTER_LAT0)
    kw = dict(WORKLOAD, num_lon=N, num_lat=M, time_step_size=dt, diffusion_coef=coef)
    return gd.Dycore(zonal_tend_filter_cutoff_wavenumber=cut, precision=prec, members=members, **kw)


def _time_steps(d, steps, warmup):
    d.step(warmup)
    t0 = time.perf_counter()
    d.step(steps)
    return time.perf_counter() - t0


# the phase launch of each form that fused_phases() reports
KERNEL_OF_FORM = {0: "k_tend", 1: "k_phase", 2: "k_wphase"}


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
    kname = KERNEL_OF_FORM[fused]
    return {"bound": "hbm", "kernel": f"{kname}<float, fast pass> (second predictor)",
            "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4),
            "traffic": traffic_from_profile(kname, f"{N}x{M}" if members == 1 else f"ens{members}x{N}x{M}"),
            "bytes_per_launch": int(nbytes), "avg_launch_us": round(us, 3)}


def _ensemble_ics(gd, N, M, members):
    """The members' initial states, generated on host threads (the generator
    is host code that releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda k: gd.ic_balanced(N, M, **ENS["ic"], pert_seed=5000 + k), range(members)))


def ensemble_workload(gd, steps, warmup, tune=True):
    """The 64-member 360x181 ensemble (dt 240 s, diffusion 1e5): member-days/s
    in mixed precision and FP64, launch shapes from the autotuner (which only
    picks shapes that give the same bits), and the roofline of the dominant
    launch over all members."""
    N, M, dt, mem = ENS["N"], ENS["M"], ENS["dt"], ENS["members"]
    ics = _ensemble_ics(gd, N, M, mem)
    out, kern, tuned = {}, {}, {}
    for prec in ("mixed", "fp64"):
        d = _cfg(gd, N, M, dt, ENS["coef"], prec, members=mem)
        for k, ic in enumerate(ics):
            d.set_state(*ic, member=k)
        if tune and hasattr(d, "autotune"):
            tuned[prec] = d.autotune(steps=min(40, steps))
            for k, ic in enumerate(ics):
                d.set_state(*ic, member=k)
        el = _time_steps(d, steps, warmup)
        kern[prec] = _kernel_us(d, 50)
        d.close()
        out[prec] = mem * steps * dt / 86400.0 / el
    return out, _tend_roofline(N, M, mem, *kern["mixed"]), tuned


def w360_workload(gd, steps, warmup):
    """The 1-degree single run (360x181, dt 240 s, diffusion 1e5; the
    reference's own test resolution): model-days/s in mixed precision and
    FP64, the phase launch form, and the roofline of that phase launch."""
    N, M, dt = 360, 181, 240.0
    ic = gd.ic_balanced(N, M, seed=20240, nmax=20, slope=-3.0, rms_height=150.0, mean_depth=8000.0)
    out, kern = {}, {}
    for prec in ("mixed", "fp64"):
        d = _cfg(gd, N, M, dt, 1.0e5, prec)
        d.set_state(*ic)
        el = _time_steps(d, steps, warmup)
        kern[prec] = _kernel_us(d, 200)
        d.close()
        out[prec] = steps * dt / 86400.0 / el
    return out, _tend_roofline(N, M, 1, *kern["mixed"])


def w1440_workload(gd, steps, warmup):
    """The 1440x721 single run (dt 60 s, spectrum to n = 80, diffusion
    6.25e3): model-days/s in mixed precision and FP64 and the roofline of the
    tendency launch."""
    N, M, dt = 1440, 721, 60.0
    ic = gd.ic_balanced(N, M, seed=20242, nmax=80)
    out, kern = {}, {}
    for prec in ("mixed", "fp64"):
        d = _cfg(gd, N, M, dt, 6.25e3, prec)
        d.set_state(*ic)
        el = _time_steps(d, steps, warmup)
        kern[prec] = _kernel_us(d, 100)
        d.close()
        out[prec] = steps * dt / 86400.0 / el
    return out, _tend_roofline(N, M, 1, *kern["mixed"])


def mm_flops(d, members):
    """Multiply-add flops of one matrix-core filter launch: for every
    momentum row (du on flagged full rows between the poles, dv on flagged
    half rows) of every member, two products of 2 N K flops (K = 2c + 2
    kept basis functions).  Returns (algorithmic, executed) where executed
    counts the padded tiles (N and K up to multiples of 32, members up to
    multiples of 16) as the kernel runs them."""
    N, M = d.N, d.M
    ff, hf, fc, hc = d.filter_table()
    ks = [2 * int(fc[j]) + 2 for j in range(1, M - 1) if ff[j]] + [2 * int(hc[j]) + 2 for j in range(M - 1) if hf[j]]
    pad = lambda x: (x + 31) // 32 * 32
    alg = sum(2 * 2 * N * k for k in ks) * members
    exe = sum(2 * 2 * pad(N) * pad(k) for k in ks) * ((members + 15) // 16 * 16)
    return alg, exe


def filter_crossover(gd, ens_steps, w1440_steps):
    """The polar filter's implementations side by side: the FP64 LDS FFT
    (mixed), FP32 matrix-core products (mixed_mfma) and BF16 split-operand
    products (mixed_bf16), on the 64-member 360x181 ensemble and the
    1440x721 single run; with the matrix-core launch's duration and its
    achieved TFLOP/s against the dense MFMA peaks (FP32 157.3, BF16 2500)."""
    out = {}
    N, M, dt, mem = ENS["N"], ENS["M"], ENS["dt"], ENS["members"]
    cases = [("ensemble_64x360x181", N, M, dt, ENS["coef"], mem, ens_steps, _ensemble_ics(gd, N, M, mem)),
             ("1440x721", 1440, 721, 60.0, 6.25e3, 1, w1440_steps, [gd.ic_balanced(1440, 721, seed=20242, nmax=80)])]
    for name, n, m, step_dt, coef, members, steps, ics in cases:
        row = {}
        for prec in ("mixed", "mixed_mfma", "mixed_bf16"):
            d = _cfg(gd, n, m, step_dt, coef, prec, members=members)
            for k, ic in enumerate(ics):
                d.set_state(*ic, member=k)
            el = _time_steps(d, steps, 12)
            r = {"days_per_s": round(members * steps * step_dt / 86400.0 / el, 4),
                 "ms_per_step": round(el / steps * 1e3, 4)}
            if prec != "mixed":
                us = d.time_kernel("filter_mm", 20)
                alg, exe = mm_flops(d, members)
                peak = 157.3 if prec == "mixed_mfma" else 2500.0
                mult = 1 if prec == "mixed_mfma" else 3  # hi.hi + hi.lo + lo.hi
                r["filter_mm_us"] = round(us, 2)
                r["filter_mm_tflops_algorithmic"] = round(alg / (us * 1e-6) / 1e12, 3)
                r["filter_mm_tflops_executed"] = round(mult * exe / (us * 1e-6) / 1e12, 3)
                r["mfma_peak_tflops"] = peak
                r["mfma_frac_executed"] = round(mult * exe / (us * 1e-6) / 1e12 / peak, 4)
            d.close()
            row[prec] = r
        out[name] = row
    return out


def other_workload(gd, args):
    base = {"n_gpus": 1, "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None,
            "dtype": "f32/f64", "roofline": None, "cpu_baseline": None}
    if args.workload == "ensemble":
        steps = min(args.steps, 1800)
        out, roof, tuned = ensemble_workload(gd, steps, args.warmup)
        mem, dt = ENS["members"], ENS["dt"]
        base["roofline"] = roof
        return dict(base, metric="member-days/sec, 64-member ensemble at 360x181 (mixed vs FP64)",
                    value=round(out["mixed"], 3), unit="member-days/s", steps=steps,
                    ms_per_step=round(mem * steps * dt / 86400.0 / out["mixed"] / steps * 1e3, 4),
                    scaling="weak", data="synthetic (seeded balanced state + per-member perturbations)",
                    config={"workload": "ensemble 64 x 360x181, dt=240 s, csp2 x6, diffusion 1e5, "
                                        f"filter cos(lat)/cos({FILTER_LAT0:g})",
                            "fp64_member_days_per_s": round(out["fp64"], 3),
                            "mixed_over_fp64": round(out["mixed"] / out["fp64"], 3), "autotuned": tuned})
    if args.workload == "1440":
        steps = min(args.steps, 2880)
        out, roof = w1440_workload(gd, steps, args.warmup)
        base["roofline"] = roof
        return dict(base, metric="model-days/sec at 1440x721 (mixed vs FP64)",
                    value=round(out["mixed"], 4), unit="model-days/s", steps=steps,
                    ms_per_step=round(60.0 / 86400.0 / out["mixed"] * 1e3, 4), scaling="weak",
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


def cpu_info():
    """Host CPU of this run: logical CPUs and the model name."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_worker(args):
    """One CPU process of the ensemble baseline: the -Ofast FP64 oracle, one
    thread, `cpu_worker_steps` steps of one member of the 64-member 360x181
    ensemble (state from the npz its parent wrote); prints the seconds."""
    from oracle import oracle as orc
    orc.use_timing_build()
    orc.set_threads(1)
    f = np.load(args.cpu_worker)
    N, M, dt = ENS["N"], ENS["M"], ENS["dt"]
    o = orc.Oracle(N, M, dt, split_scheme="csp2", subcycles=6, qcon_modified=True,
                   uv_adv_scheme="center_diff", use_diffusion=True, diffusion_order=2,
                   diffusion_coef=ENS["coef"], use_zonal_tend_filter=True, cutoff=list(f["cut"]))
    o.set_initial_state(f["u"], f["v"], f["gd"], f["ghs"])
    o.start()
    t0 = time.perf_counter()
    o.run(args.cpu_worker_steps)
    print(f"{time.perf_counter() - t0:.6f}", flush=True)


def cpu_ensemble_baseline(gd, nproc, steps):
    """P independent single-thread processes of the -Ofast oracle (the
    reference is serial), each stepping its own ensemble member: aggregate
    member-days/s.  Run before the process touches the GPU."""
    import subprocess
    import tempfile
    N, M = ENS["N"], ENS["M"]
    cut = gd.cutoff_from_lat0(N, M, FILTER_LAT0)
    procs, files = [], []
    tmp = tempfile.mkdtemp(prefix="gd_cpu_")
    for k in range(nproc):
        u, v, h, ghs = gd.ic_balanced(N, M, **ENS["ic"], pert_seed=5000 + k)
        path = os.path.join(tmp, f"m{k}.npz")
        np.savez(path, u=u, v=v, gd=h, ghs=ghs, cut=np.array(cut, np.int32))
        files.append(path)
    t0 = time.perf_counter()
    for path in files:
        procs.append(subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", path,
                                       "--cpu-worker-steps", str(steps)], stdout=subprocess.PIPE, text=True))
    secs = []
    for pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError("CPU baseline worker failed")
        secs.append(float(out.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    for path in files:
        os.remove(path)
    os.rmdir(tmp)
    return nproc * steps * ENS["dt"] / 86400.0 / max(secs), max(secs), wall


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
    ap.add_argument("--probe-grid", default="720", choices=["720", "360", "1440", "ens64"],
                    help="workload of the --pmc-probe steps: 720x361, the 360x181 or 1440x721 single run, "
                         "or the 64-member 360x181 ensemble")
    ap.add_argument("--mm-probe", type=int, default=0, metavar="K",
                    help="counter runs: K steps of the 64-member ensemble in the two matrix-core "
                         "filter modes, no timing output")
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the short ensemble / 1440x721 / serving measurements of the default line")
    ap.add_argument("--ens-steps", type=int, default=360)
    ap.add_argument("--w1440-steps", type=int, default=480)
    ap.add_argument("--serve-short", type=int, default=12)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="processes of the CPU ensemble baseline (0: min(16, CPUs))")
    ap.add_argument("--cpu-ens-steps", type=int, default=24)
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-worker-steps", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed, not built)"}))
        sys.exit(2)
    import __graft_entry__ as ge
    gd = ge.load_package()
    N, M, dt = WORKLOAD["num_lon"], WORKLOAD["num_lat"], WORKLOAD["time_step_size"]
    ic = gd.ic_balanced(N, M, **IC)
    if args.pmc_probe:
        pmc_probe(gd, ic, args.pmc_probe, args.probe_grid)
        return
    if args.mm_probe:
        # counter runs of the matrix-core filter: K steps of the 64-member
        # ensemble in each matrix-core mode (run with GD_NO_GRAPH=1 so each
        # kernel is its own dispatch)
        ics = _ensemble_ics(gd, ENS["N"], ENS["M"], ENS["members"])
        for prec in ("mixed_mfma", "mixed_bf16"):
            d = _cfg(gd, ENS["N"], ENS["M"], ENS["dt"], ENS["coef"], prec, members=ENS["members"])
            for k, x in enumerate(ics):
                d.set_state(*x, member=k)
            d.step(args.mm_probe)
            alg, exe = mm_flops(d, ENS["members"])
            d.close()
            print(f"mm probe {prec}: {args.mm_probe} steps, {alg} algorithmic / {exe} executed flops per "
                  f"filter launch (x3 products in mixed_bf16)", flush=True)
        return
    if args.workload != "cfg2":
        print(json.dumps(other_workload(gd, args)))
        return
    # CPU legs first, before this process touches the GPU (the ensemble leg
    # starts child processes)
    cpu = None
    if not args.no_cpu:
        cv, cel = cpu_baseline(ic, args.cpu_steps)
        nproc = args.cpu_procs or min(16, os.cpu_count() or 1)
        ev, esec, ewall = cpu_ensemble_baseline(gd, nproc, args.cpu_ens_steps)
        cpu = {"value": round(cv, 6), "unit": "model-days/s", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same 720x361 FP64 workload, "
                         f"oracle/gd_oracle.c built -O3 -ffast-math (the reference's -Ofast), "
                         f"single thread ({cel:.1f} s)",
               "host": cpu_info(),
               "ensemble": {"value": round(ev, 5), "unit": "member-days/s", "processes": nproc,
                            "cores": nproc,
                            "sample": f"{nproc} single-thread processes, each the first "
                                      f"{args.cpu_ens_steps} steps of its own member of the 64-member "
                                      f"360x181 ensemble (same -Ofast oracle; slowest {esec:.1f} s)"}}
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
    kname = KERNEL_OF_FORM[mixed["fused"]]
    roofline = {"bound": "hbm",
                "kernel": f"{kname}<float, fast pass> (second predictor" +
                          (": tendency + filter rows + epilogue)" if mixed["fused"] else ")"),
                "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": traffic_from_profile(kname),
                "bytes_per_launch": int(nbytes),
                "avg_launch_us": round(mixed["kern_us"], 3),
                "dominant_family_by_time": dom,
                "family_ms_per_step": {k: round(v[1] / 60.0, 4) for k, v in prof.items()}}
    extra = {}
    if args.extra:
        eo, eroof, etuned = ensemble_workload(gd, args.ens_steps, 36)
        extra["ensemble_64x360x181"] = {
            "member_days_per_s": round(eo["mixed"], 3), "fp64_member_days_per_s": round(eo["fp64"], 3),
            "mixed_over_fp64": round(eo["mixed"] / eo["fp64"], 3), "steps": args.ens_steps,
            "autotuned": etuned, "roofline": {k: eroof[k] for k in ("kernel", "achieved", "frac", "avg_launch_us", "traffic")}}
        if cpu is not None:
            extra["ensemble_64x360x181"]["over_cpu_processes"] = round(eo["mixed"] / cpu["ensemble"]["value"], 1)
        so, sroof = w360_workload(gd, 360, 30)
        extra["360x181"] = {
            "model_days_per_s": round(so["mixed"], 3), "fp64_model_days_per_s": round(so["fp64"], 3),
            "mixed_over_fp64": round(so["mixed"] / so["fp64"], 3), "steps": 360,
            "roofline": {k: sroof[k] for k in ("kernel", "achieved", "frac", "avg_launch_us", "traffic")}}
        wo, wroof = w1440_workload(gd, args.w1440_steps, 60)
        extra["1440x721"] = {
            "model_days_per_s": round(wo["mixed"], 4), "fp64_model_days_per_s": round(wo["fp64"], 4),
            "mixed_over_fp64": round(wo["mixed"] / wo["fp64"], 3), "steps": args.w1440_steps,
            "roofline": {k: wroof[k] for k in ("kernel", "achieved", "frac", "avg_launch_us", "traffic")}}
        extra["filter_implementations"] = filter_crossover(gd, 120, 240)
        sv = serving_workload(gd, argparse.Namespace(serve_requests=args.serve_short), {})
        extra["serving"] = {"model_days_per_s": sv["value"], "p50_s": sv["latency_s"]["p50"],
                            "p99_s": sv["latency_s"]["p99"],
                            "fp64_model_days_per_s": sv["config"]["fp64_model_days_per_s"],
                            "fp64_p50_s": sv["latency_s"]["fp64_p50"], "fp64_p99_s": sv["latency_s"]["fp64_p99"],
                            "requests": sv["config"]["requests"], "model_days": sv["config"]["model_days"]}
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
                   "parallelism": "single GPU, single run",
                   "other_workloads": extra or None},
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
