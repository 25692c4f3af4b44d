"""gamil-dycore_amd -- MI355X-native barotropic time step of gamil-dycore.

Python mirror of the C ABI in include/gamil_dycore.h (ctypes over the in-tree
libgamil_dycore.so).  The object model follows the reference's dycore_mod
facade (src/dycore_mod.F90:22-25): construct with the &dycore_params values
(dycore_init), hand over the initial state (the IC modules' writes into
state(1)), then step (the dycore_run loop body) and read the state back.

There is no CPU fallback: every compute call runs the HIP kernels, and
importing the module fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgamil_dycore.so")

GD_OK, GD_ERR_INVALID_ARG, GD_ERR_HIP, GD_ERR_NAN, GD_ERR_NOT_SUPPORTED = range(5)
SPLIT = {"none": 0, "csp1": 1, "csp2": 2, "isp": 3}
ADV = {"center_diff": 0, "upwind": 1, "weno": 2}
PASS = {"all": 0, "fast": 1, "slow": 2}
PRECISION = {"fp64": 0, "mixed": 1, "mixed_bf16": 2, "mixed_mfma": 3}


class DycoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[Error]: {msg} (code {code})")
        self.code = code


class _Config(C.Structure):
    _fields_ = [
        ("num_lon", C.c_int32), ("num_lat", C.c_int32), ("time_step_size", C.c_double),
        ("split_scheme", C.c_int32), ("subcycles", C.c_int32), ("qcon_modified", C.c_int32),
        ("uv_adv_scheme", C.c_int32), ("uv_adv_upwind_lon_beta", C.c_double),
        ("uv_adv_upwind_lat_beta", C.c_double), ("use_diffusion", C.c_int32),
        ("diffusion_order", C.c_int32), ("diffusion_coef", C.c_double),
        ("use_zonal_tend_filter", C.c_int32), ("n_cutoff", C.c_int32),
        ("cutoff", C.POINTER(C.c_int32)), ("precision", C.c_int32), ("device", C.c_int32),
    ]


class IOInfo(C.Structure):
    _fields_ = [
        ("desc", C.c_char_p), ("author", C.c_char_p), ("time_units", C.c_char_p),
        ("start_time", C.c_char_p), ("time_value", C.c_double), ("time_step_size", C.c_double),
        ("time_scheme", C.c_char_p), ("split_scheme", C.c_char_p), ("subcycles", C.c_int32),
        ("restart_time", C.c_char_p), ("elapsed_seconds", C.c_double),
    ]


class ServeRequest(C.Structure):
    _fields_ = [
        ("num_lon", C.c_int32), ("num_lat", C.c_int32), ("time_step_size", C.c_double),
        ("diffusion_coef", C.c_double), ("nsteps", C.c_int32), ("arrival_s", C.c_double),
        ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)), ("gd", C.POINTER(C.c_double)),
        ("ghs", C.POINTER(C.c_double)),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("arrival_s", C.c_double), ("start_s", C.c_double), ("finish_s", C.c_double),
        ("total_mass", C.c_double), ("total_energy", C.c_double), ("state_hash", C.c_uint64),
        ("status", C.c_int32), ("slot", C.c_int32),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("base", C.c_void_p), ("filter_lat0", C.c_double), ("hbm_budget_bytes", C.c_int64),
        ("max_slots", C.c_int32), ("chunk_steps", C.c_int32),
    ]


class ICParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("nmax", C.c_int32), ("slope", C.c_double),
        ("rms_height", C.c_double), ("mean_depth", C.c_double), ("n_mountains", C.c_int32),
        ("mountain_seed", C.c_uint64), ("pert_seed", C.c_uint64), ("pert_nmax", C.c_int32),
        ("pert_rms", C.c_double),
    ]


# every symbol declared in include/gamil_dycore.h
EXPORTS = (
    "gd_create", "gd_destroy", "gd_last_error", "gd_set_state", "gd_diag", "gd_step",
    "gd_get_state", "gd_get_diag_fields", "gd_space_operators", "gd_predict_correct",
    "gd_ordinary_diffusion", "gd_filter_rows", "gd_filter_table", "gd_ic_balanced",
    "gd_ic_rossby_haurwitz", "gd_ic_mountain_zonal_flow", "gd_profile_kernels",
    "gd_device_bytes", "gd_cutoff_from_lat0", "gd_time_kernel", "gd_history_write",
    "gd_restart_write", "gd_restart_read", "gd_io_last_error", "gd_ic_steady_geostrophic",
    "gd_ic_jet_zonal_flow", "gd_ic_shallow_water_waves", "gd_serve", "gd_serve_last_error",
    "gd_set_active_members", "gd_copy_member", "gd_fused_phases", "gd_filter_rows_noalloc",
    "gd_check_antisymmetry",
    "gd_autotune", "gd_server_create", "gd_server_submit", "gd_server_poll", "gd_server_destroy",
)

_lib = None


def lib():
    """Load libgamil_dycore.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    i64 = C.c_int64
    L.gd_create.argtypes = [C.POINTER(_Config), C.c_int32, C.POINTER(vp)]
    L.gd_destroy.argtypes = [vp]
    L.gd_last_error.restype = C.c_char_p
    L.gd_last_error.argtypes = [vp]
    L.gd_set_state.argtypes = [vp, C.c_int32, dp, dp, dp, dp, i64]
    L.gd_diag.argtypes = [vp, dp, dp]
    L.gd_step.argtypes = [vp, C.c_int32, dp, dp, dp, ip]
    L.gd_get_state.argtypes = [vp, C.c_int32, dp, dp, dp, i64]
    L.gd_get_diag_fields.argtypes = [vp, C.c_int32, dp, dp, i64]
    L.gd_space_operators.argtypes = [vp, C.c_int32, C.c_int32, dp, dp, dp, i64]
    L.gd_predict_correct.argtypes = [vp, C.c_double, C.c_int32, C.c_int32, dp]
    L.gd_ordinary_diffusion.argtypes = [vp, C.c_double]
    L.gd_filter_rows.argtypes = [vp, C.c_int32, dp, i64]
    L.gd_filter_table.argtypes = [vp, ip, ip, ip, ip]
    L.gd_filter_rows_noalloc.argtypes = [vp, C.c_int32, C.c_void_p, i64, vp]
    L.gd_autotune.argtypes = [vp, C.c_int32, C.c_char_p, ip, ip, dp]
    L.gd_ic_balanced.argtypes = [C.c_int32, C.c_int32, C.POINTER(ICParams), dp, dp, dp, dp, i64]
    L.gd_ic_rossby_haurwitz.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                        C.c_double, dp, dp, dp, dp, i64]
    L.gd_ic_mountain_zonal_flow.argtypes = [C.c_int32, C.c_int32, dp, dp, dp, dp, i64]
    L.gd_ic_steady_geostrophic.argtypes = [C.c_int32, C.c_int32, dp, dp, dp, dp, i64]
    L.gd_ic_jet_zonal_flow.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, dp, dp, dp, i64]
    L.gd_ic_shallow_water_waves.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, dp, dp, dp, i64, dp]
    L.gd_profile_kernels.argtypes = [vp, C.c_int32, C.c_int32, ip, C.POINTER(C.c_int64), dp,
                                     C.c_char_p, C.c_int32]
    L.gd_cutoff_from_lat0.argtypes = [C.c_int32, C.c_int32, C.c_double, ip, C.c_int32]
    L.gd_device_bytes.restype = C.c_int64
    L.gd_device_bytes.argtypes = [vp]
    L.gd_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, dp]
    L.gd_history_write.argtypes = [C.c_char_p, C.POINTER(IOInfo), C.c_int32, C.c_int32, dp, dp, dp,
                                   dp, dp, dp, i64, C.c_double, C.c_double]
    L.gd_restart_write.argtypes = [C.c_char_p, C.POINTER(IOInfo), C.c_int32, C.c_int32, dp, dp, dp,
                                   dp, i64]
    L.gd_restart_read.argtypes = [C.c_char_p, C.c_int32, C.c_int32, dp, dp, dp, dp, i64, C.c_char_p,
                                  C.c_int32, dp]
    L.gd_io_last_error.restype = C.c_char_p
    L.gd_io_last_error.argtypes = []
    L.gd_set_active_members.argtypes = [vp, C.c_int32]
    L.gd_copy_member.argtypes = [vp, C.c_int32, C.c_int32]
    L.gd_fused_phases.argtypes = [vp]
    L.gd_check_antisymmetry.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.gd_serve.argtypes = [C.POINTER(ServeConfig), C.POINTER(ServeRequest), C.c_int32,
                           C.POINTER(ServeResult), dp]
    L.gd_serve_last_error.restype = C.c_char_p
    L.gd_serve_last_error.argtypes = []
    L.gd_server_create.argtypes = [C.POINTER(ServeConfig), C.POINTER(vp)]
    L.gd_server_submit.argtypes = [vp, C.POINTER(ServeRequest), C.POINTER(C.c_int64)]
    L.gd_server_poll.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(ServeResult), ip]
    L.gd_server_destroy.argtypes = [vp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _check(code, h=None):
    if code != GD_OK:
        msg = lib().gd_last_error(h).decode(errors="replace")
        raise DycoreError(code, msg)


# ---------------------------------------------------------------------------
# initial conditions (host set-up, not on the time step)
# ---------------------------------------------------------------------------

def ic_balanced(num_lon, num_lat, seed=1, nmax=20, slope=-3.0, rms_height=150.0,
                mean_depth=8000.0, n_mountains=0, mountain_seed=0, pert_seed=0, pert_nmax=0,
                pert_rms=0.0):
    """Seeded balanced spherical-harmonic state: (u, v, gd, ghs) dense arrays."""
    N, M = int(num_lon), int(num_lat)
    u = np.zeros((M, N)); v = np.zeros((M - 1, N)); gd = np.zeros((M, N)); ghs = np.zeros((M, N))
    p = ICParams(seed, nmax, slope, rms_height, mean_depth, n_mountains, mountain_seed,
                 pert_seed, pert_nmax, pert_rms)
    _check(lib().gd_ic_balanced(N, M, C.byref(p), _dp(u), _dp(v), _dp(gd), _dp(ghs), N))
    return u, v, gd, ghs


def ic_rossby_haurwitz(num_lon, num_lat, R=4.0, omg=7.848e-6, gd0=8.0e3 * 9.80616):
    N, M = int(num_lon), int(num_lat)
    u = np.zeros((M, N)); v = np.zeros((M - 1, N)); gd = np.zeros((M, N)); ghs = np.zeros((M, N))
    _check(lib().gd_ic_rossby_haurwitz(N, M, R, omg, gd0, _dp(u), _dp(v), _dp(gd), _dp(ghs), N))
    return u, v, gd, ghs


def ic_mountain_zonal_flow(num_lon, num_lat):
    N, M = int(num_lon), int(num_lat)
    u = np.zeros((M, N)); v = np.zeros((M - 1, N)); gd = np.zeros((M, N)); ghs = np.zeros((M, N))
    _check(lib().gd_ic_mountain_zonal_flow(N, M, _dp(u), _dp(v), _dp(gd), _dp(ghs), N))
    return u, v, gd, ghs


def _ic_arrays(N, M):
    return np.zeros((M, N)), np.zeros((M - 1, N)), np.zeros((M, N)), np.zeros((M, N))


def ic_steady_geostrophic(num_lon, num_lat):
    """Williamson test 2 (steady_geostrophic_flow_test_mod.F90:20-62)."""
    N, M = int(num_lon), int(num_lat)
    u, v, gd, ghs = _ic_arrays(N, M)
    _check(lib().gd_ic_steady_geostrophic(N, M, _dp(u), _dp(v), _dp(gd), _dp(ghs), N))
    return u, v, gd, ghs


def ic_jet_zonal_flow(num_lon, num_lat, perturb=True):
    """Galewsky barotropic jet (jet_zonal_flow_test_mod.F90:29-105)."""
    N, M = int(num_lon), int(num_lat)
    u, v, gd, ghs = _ic_arrays(N, M)
    _check(lib().gd_ic_jet_zonal_flow(N, M, int(bool(perturb)), _dp(u), _dp(v), _dp(gd), _dp(ghs), N))
    return u, v, gd, ghs


def ic_shallow_water_waves(num_lon, num_lat, wave="rossby"):
    """Shamir-Paldor analytic wave at time 0 (shallow_water_waves_test_mod.F90):
    (u, v, gd, ghs, phase_speed [rad/s]); wave 'rossby', 'eig' or 'wig'."""
    N, M = int(num_lon), int(num_lat)
    u, v, gd, ghs = _ic_arrays(N, M)
    c = C.c_double(0.0)
    flag = {"rossby": 0, "eig": 1, "wig": -1}[wave]
    _check(lib().gd_ic_shallow_water_waves(N, M, flag, _dp(u), _dp(v), _dp(gd), _dp(ghs), N, C.byref(c)))
    return u, v, gd, ghs, c.value


def cutoff_from_lat0(num_lon, num_lat, lat0_deg):
    """Polar filter table whose kept wavenumber falls toward the poles as
    cos(lat)/cos(lat0) (gd_cutoff_from_lat0 in the C ABI)."""
    cap = int(num_lat)
    out = np.zeros(cap, np.int32)
    n = lib().gd_cutoff_from_lat0(int(num_lon), int(num_lat), float(lat0_deg),
                                  out.ctypes.data_as(C.POINTER(C.c_int32)), cap)
    if n < 0:
        raise DycoreError(GD_ERR_INVALID_ARG, "bad arguments to cutoff_from_lat0")
    return [int(c) for c in out[:n]]


# ---------------------------------------------------------------------------
# history and restart files (NetCDF classic, host only)
# ---------------------------------------------------------------------------

def _b(x):
    return None if x is None else str(x).encode()


def _info(desc="", author="N/A", time_units="days", start_time="0001-01-01T00_00_00",
          time_value=0.0, time_step_size=0.0, time_scheme="predict_correct", split_scheme="",
          subcycles=0, restart_time="", elapsed_seconds=0.0):
    return IOInfo(_b(desc), _b(author), _b(time_units), _b(start_time), float(time_value),
                  float(time_step_size), _b(time_scheme), _b(split_scheme), int(subcycles),
                  _b(restart_time), float(elapsed_seconds))


def _io_check(code):
    if code != GD_OK:
        raise DycoreError(code, lib().gd_io_last_error().decode(errors="replace"))


def _dense(a, rows, N):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(rows, N))


def write_history(path, u, v, gd, ghs, div, vor, total_energy, total_mass, **info):
    """history_write_state (history_mod.F90:93-124) of C-grid fields into the
    NetCDF file `path`; `info` holds the gd_io_info metadata (time_value,
    time_units, start_time, desc, time_step_size, split_scheme, subcycles ...)."""
    M, N = np.shape(gd)
    a = [_dense(u, M, N), _dense(v, M - 1, N), _dense(gd, M, N), _dense(ghs, M, N),
         _dense(div, M, N), _dense(vor, M - 1, N)]
    inf = _info(**info)
    _io_check(lib().gd_history_write(_b(path), C.byref(inf), N, M, *[_dp(x) for x in a], N,
                                     float(total_energy), float(total_mass)))


def write_restart(path, u, v, gd, ghs, **info):
    """restart_write (restart_mod.F90:57-75) into the NetCDF file `path`."""
    M, N = np.shape(gd)
    a = [_dense(u, M, N), _dense(v, M - 1, N), _dense(gd, M, N), _dense(ghs, M, N)]
    inf = _info(**info)
    _io_check(lib().gd_restart_write(_b(path), C.byref(inf), N, M, *[_dp(x) for x in a], N))


def read_restart(path, num_lon, num_lat):
    """restart_read (restart_mod.F90:37-55): (u, v, gd, ghs, restart_time, elapsed_seconds)."""
    N, M = int(num_lon), int(num_lat)
    u = np.zeros((M, N)); v = np.zeros((M - 1, N)); gd = np.zeros((M, N)); ghs = np.zeros((M, N))
    rt = C.create_string_buffer(64)
    el = C.c_double(0.0)
    _io_check(lib().gd_restart_read(_b(path), N, M, _dp(u), _dp(v), _dp(gd), _dp(ghs), N, rt, 64,
                                    C.byref(el)))
    return u, v, gd, ghs, rt.value.decode(), el.value


# ---------------------------------------------------------------------------
# serving: continuous batching of independent runs (gd_serve)
# ---------------------------------------------------------------------------

def _serve_config(precision, split_scheme, subcycles, qcon_modified, uv_adv_scheme, use_diffusion,
                  diffusion_order, use_zonal_tend_filter, filter_lat0, hbm_budget_bytes, max_slots,
                  chunk_steps, device):
    cut = np.zeros(1, np.int32)
    base = _Config(8, 5, 1.0, SPLIT[split_scheme], int(subcycles), int(bool(qcon_modified)),
                   ADV[uv_adv_scheme], 0.0, 0.5, int(bool(use_diffusion)), int(diffusion_order), 0.0,
                   int(bool(use_zonal_tend_filter)), 0, cut.ctypes.data_as(C.POINTER(C.c_int32)),
                   PRECISION[precision], int(device))
    cfg = ServeConfig(C.cast(C.pointer(base), C.c_void_p), float(filter_lat0), int(hbm_budget_bytes),
                      int(max_slots), int(chunk_steps))
    return cfg, (cut, base)


class Server:
    """Online continuous batching (gd_server_*): submit runs while others are
    on the GPU, poll for their results.  Options as for serve()."""

    def __init__(self, precision="mixed", split_scheme="csp2", subcycles=6, qcon_modified=True,
                 uv_adv_scheme="center_diff", use_diffusion=True, diffusion_order=2,
                 use_zonal_tend_filter=True, filter_lat0=75.0, hbm_budget_bytes=64 << 30, max_slots=64,
                 chunk_steps=60, device=0):
        self._cfg, self._keep = _serve_config(precision, split_scheme, subcycles, qcon_modified,
                                              uv_adv_scheme, use_diffusion, diffusion_order,
                                              use_zonal_tend_filter, filter_lat0, hbm_budget_bytes,
                                              max_slots, chunk_steps, device)
        self.s = C.c_void_p()
        code = lib().gd_server_create(C.byref(self._cfg), C.byref(self.s))
        if code != GD_OK:
            raise DycoreError(code, lib().gd_serve_last_error().decode(errors="replace"))

    def submit(self, num_lon, num_lat, time_step_size, nsteps, state, diffusion_coef=0.0):
        """Submit one run (its state is copied); returns its id."""
        N, M = int(num_lon), int(num_lat)
        u, v, g, hs = (_dense(a, r, N) for a, r in zip(state, (M, M - 1, M, M)))
        q = ServeRequest(N, M, float(time_step_size), float(diffusion_coef), int(nsteps), 0.0,
                         _dp(u), _dp(v), _dp(g), _dp(hs))
        rid = C.c_int64(-1)
        code = lib().gd_server_submit(self.s, C.byref(q), C.byref(rid))
        if code != GD_OK:
            raise DycoreError(code, lib().gd_serve_last_error().decode(errors="replace"))
        return rid.value

    def poll(self, rid, wait=False):
        """The run's result dict when it is done, else None."""
        r = ServeResult()
        done = C.c_int32(0)
        code = lib().gd_server_poll(self.s, int(rid), int(bool(wait)), C.byref(r), C.byref(done))
        if code != GD_OK:
            raise DycoreError(code, lib().gd_serve_last_error().decode(errors="replace"))
        return {f: getattr(r, f) for f, _ in ServeResult._fields_} if done.value else None

    def close(self):
        """Wait for every submitted run, then release the pools."""
        if getattr(self, "s", None):
            code = lib().gd_server_destroy(self.s)
            self.s = None
            if code != GD_OK:
                raise DycoreError(code, lib().gd_serve_last_error().decode(errors="replace"))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve(requests, precision="mixed", split_scheme="csp2", subcycles=6, qcon_modified=True,
          uv_adv_scheme="center_diff", use_diffusion=True, diffusion_order=2,
          use_zonal_tend_filter=True, filter_lat0=75.0, hbm_budget_bytes=64 << 30, max_slots=64,
          chunk_steps=60, device=0):
    """Serve a stream of runs.  Each request is a dict with num_lon, num_lat,
    time_step_size, nsteps, arrival_s, diffusion_coef and state = (u, v, gd,
    ghs) dense arrays.  Returns (results, wall_seconds); results[k] is a dict
    for requests[k] (arrival_s, start_s, finish_s, total_mass, total_energy,
    state_hash, status, slot)."""
    L = lib()
    keep = []
    reqs = (ServeRequest * max(1, len(requests)))()
    for k, q in enumerate(requests):
        N, M = int(q["num_lon"]), int(q["num_lat"])
        u, v, g, hs = (_dense(a, r, N) for a, r in zip(q["state"], (M, M - 1, M, M)))
        keep += [u, v, g, hs]
        reqs[k] = ServeRequest(N, M, float(q["time_step_size"]), float(q.get("diffusion_coef", 0.0)),
                               int(q["nsteps"]), float(q["arrival_s"]), _dp(u), _dp(v), _dp(g), _dp(hs))
    cfg, keep_cfg = _serve_config(precision, split_scheme, subcycles, qcon_modified, uv_adv_scheme,
                                  use_diffusion, diffusion_order, use_zonal_tend_filter, filter_lat0,
                                  hbm_budget_bytes, max_slots, chunk_steps, device)
    res = (ServeResult * max(1, len(requests)))()
    wall = C.c_double(0.0)
    code = L.gd_serve(C.byref(cfg), reqs, len(requests), res, C.byref(wall))
    if code != GD_OK:
        raise DycoreError(code, L.gd_serve_last_error().decode(errors="replace"))
    out = [{f: getattr(res[k], f) for f, _ in ServeResult._fields_} for k in range(len(requests))]
    return out, wall.value


# ---------------------------------------------------------------------------
# the dycore
# ---------------------------------------------------------------------------

class Dycore:
    """A device-resident ensemble of `members` runs of one &dycore_params config."""

    def __init__(self, num_lon, num_lat, time_step_size, split_scheme="csp2", subcycles=4,
                 qcon_modified=True, uv_adv_scheme="center_diff", uv_adv_upwind_lon_beta=0.0,
                 uv_adv_upwind_lat_beta=0.5, use_diffusion=False, diffusion_order=2,
                 diffusion_coef=0.0, use_zonal_tend_filter=True,
                 zonal_tend_filter_cutoff_wavenumber=(), precision="fp64", members=1, device=0):
        L = lib()
        self.N, self.M, self.members = int(num_lon), int(num_lat), int(members)
        self.dt = float(time_step_size)
        cut = list(zonal_tend_filter_cutoff_wavenumber)
        self._cut = np.ascontiguousarray(np.asarray(cut or [0], dtype=np.int32))
        cfg = _Config(self.N, self.M, self.dt, SPLIT[split_scheme], int(subcycles),
                      int(bool(qcon_modified)), ADV[uv_adv_scheme], float(uv_adv_upwind_lon_beta),
                      float(uv_adv_upwind_lat_beta), int(bool(use_diffusion)), int(diffusion_order),
                      float(diffusion_coef), int(bool(use_zonal_tend_filter)), len(cut),
                      self._cut.ctypes.data_as(C.POINTER(C.c_int32)), PRECISION[precision],
                      int(device))
        h = C.c_void_p()
        _check(L.gd_create(C.byref(cfg), self.members, C.byref(h)))
        self.h = h
        self._ghs = {}

    def close(self):
        if getattr(self, "h", None):
            lib().gd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state hand-over ------------------------------------------------
    def set_state(self, u, v, gd, ghs, member=0):
        f = lambda a, r: np.ascontiguousarray(a, dtype=np.float64).reshape(r, self.N)
        u, v, gd, ghs = f(u, self.M), f(v, self.M - 1), f(gd, self.M), f(ghs, self.M)
        _check(lib().gd_set_state(self.h, member, _dp(u), _dp(v), _dp(gd), _dp(ghs), self.N), self.h)
        self._ghs[member] = ghs.copy()

    def get_state(self, member=0):
        u = np.zeros((self.M, self.N)); v = np.zeros((self.M - 1, self.N)); gd = np.zeros((self.M, self.N))
        _check(lib().gd_get_state(self.h, member, _dp(u), _dp(v), _dp(gd), self.N), self.h)
        return u, v, gd

    def diag_fields(self, member=0):
        div = np.zeros((self.M, self.N)); vor = np.zeros((self.M - 1, self.N))
        _check(lib().gd_get_diag_fields(self.h, member, _dp(div), _dp(vor), self.N), self.h)
        return div, vor

    def diag(self):
        tm = np.zeros(self.members); te = np.zeros(self.members)
        _check(lib().gd_diag(self.h, _dp(tm), _dp(te)), self.h)
        return tm, te

    # -- time stepping ----------------------------------------------------
    def step(self, nsteps, raise_on_nan=True):
        n = int(nsteps)
        mass = np.zeros((n, self.members)); energy = np.zeros((n, self.members))
        beta = np.zeros((n, self.members)); ns = C.c_int32(0)
        code = lib().gd_step(self.h, n, _dp(mass), _dp(energy), _dp(beta), C.byref(ns))
        if code == GD_ERR_NAN and not raise_on_nan:
            return mass, energy, beta, ns.value
        _check(code, self.h)
        return mass, energy, beta, ns.value

    def set_active_members(self, n_active):
        """Step only members 0..n_active-1 from now on."""
        _check(lib().gd_set_active_members(self.h, int(n_active)), self.h)

    def copy_member(self, dst, src):
        _check(lib().gd_copy_member(self.h, int(dst), int(src)), self.h)
        if src in self._ghs:
            self._ghs[dst] = self._ghs[src]

    # -- operator seams ----------------------------------------------------
    def space_operators(self, pass_="all", member=0):
        du = np.zeros((self.M, self.N)); dv = np.zeros((self.M - 1, self.N)); dgd = np.zeros((self.M, self.N))
        _check(lib().gd_space_operators(self.h, member, PASS[pass_], _dp(du), _dp(dv), _dp(dgd), self.N), self.h)
        return du, dv, dgd

    def check_antisymmetry(self, member=0):
        """check_antisymmetry (dycore_mod.F90:794-851) of the member's current
        state on the device.  Returns (terms, magnitudes, sums): the ten inner
        products u_adv_lon, u_adv_lat, fv, u_pgf, v_adv_lon, v_adv_lat, fu,
        v_pgf, mass_div_lon, mass_div_lat; the same sums of magnitudes; and
        the four sums the reference prints (advection, u_pgf + mass_div_lon,
        fv - fu, v_pgf + mass_div_lat)."""
        terms, mags, sums = np.zeros(10), np.zeros(10), np.zeros(4)
        _check(lib().gd_check_antisymmetry(self.h, member, _dp(terms), _dp(mags), _dp(sums)), self.h)
        return terms, mags, sums

    def predict_correct(self, dt, pass_="all", qcon_modified=True):
        beta = np.zeros(self.members)
        _check(lib().gd_predict_correct(self.h, float(dt), PASS[pass_], int(bool(qcon_modified)), _dp(beta)), self.h)
        return beta

    def ordinary_diffusion(self, dt):
        _check(lib().gd_ordinary_diffusion(self.h, float(dt)), self.h)

    def filter_rows(self, x, half=False):
        y = np.ascontiguousarray(x, dtype=np.float64).copy()
        _check(lib().gd_filter_rows(self.h, int(bool(half)), _dp(y), self.N), self.h)
        return y

    def filter_rows_noalloc(self, x_ptr, half=False, ld=None, stream=None):
        """Filter the flagged rows of a float64 array in device memory at
        address x_ptr in place (gd_filter_rows_noalloc: no allocation; with
        stream=None the call waits for the result)."""
        _check(lib().gd_filter_rows_noalloc(self.h, int(bool(half)), C.c_void_p(int(x_ptr)),
                                            int(self.N if ld is None else ld), stream), self.h)

    def filter_table(self):
        ff = np.zeros(self.M, np.int32); hf = np.zeros(self.M - 1, np.int32)
        fc = np.zeros(self.M, np.int32); hc = np.zeros(self.M - 1, np.int32)
        ipp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _check(lib().gd_filter_table(self.h, ipp(ff), ipp(hf), ipp(fc), ipp(hc)), self.h)
        return ff, hf, fc, hc

    def profile_kernels(self, nsteps, max_families=16):
        nf = C.c_int32(0)
        launches = np.zeros(max_families, np.int64); ms = np.zeros(max_families)
        names = C.create_string_buffer(1024)
        _check(lib().gd_profile_kernels(self.h, int(nsteps), max_families, C.byref(nf),
                                        launches.ctypes.data_as(C.POINTER(C.c_int64)), _dp(ms),
                                        names, 1024), self.h)
        fam = names.value.decode().split(",")[: nf.value]
        return {f: (int(launches[k]), float(ms[k])) for k, f in enumerate(fam)}

    # -- output ------------------------------------------------------------
    def write_history(self, path, member=0, **info):
        """Write the member's current state as one history file (history_write_state)."""
        u, v, gd = self.get_state(member)
        div, vor = self.diag_fields(member)
        tm, te = self.diag()
        write_history(path, u, v, gd, self._ghs[member], div, vor, te[member], tm[member], **info)

    def write_restart(self, path, member=0, **info):
        """Write the member's current state as one restart file (restart_write)."""
        u, v, gd = self.get_state(member)
        write_restart(path, u, v, gd, self._ghs[member], **info)

    def fused_phases(self):
        """Form of the tendency phase launches: 0 a tendency launch then a
        filter launch, 1 one fused launch (k_phase), 2 one fused launch with
        a workgroup per row (k_wphase).  Nonzero means fused."""
        return int(lib().gd_fused_phases(self.h))

    def autotune(self, steps=40, cache=None):
        """Pick the fastest launch shapes for the current active-member count
        (gd_autotune; results stay bit for bit those of the defaults).
        Returns {'fused', 'form', 'upd_wpm', 'ms_per_step', 'cache'}: 'form' as
        fused_phases() reports it, 'fused' whether that form is fused."""
        f, w, ms = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        _check(lib().gd_autotune(self.h, int(steps), None if cache is None else str(cache).encode(),
                                 C.byref(f), C.byref(w), C.byref(ms)), self.h)
        return {"fused": bool(f.value), "form": int(f.value), "upd_wpm": int(w.value),
                "ms_per_step": float(ms.value), "cache": cache}

    def device_bytes(self):
        return int(lib().gd_device_bytes(self.h))

    def time_kernel(self, kernel, reps=200):
        """Microseconds per launch of one kernel replayed back to back:
        kernel 'tend' (fast pass, second predictor), 'filter', 'update',
        'phase' (that phase as one fused launch; fused handles only) or
        'filter_mm' (the matrix-core projection of the momentum rows; the
        mixed_mfma and mixed_bf16 modes only)."""
        us = C.c_double(0.0)
        kid = {"tend": 0, "filter": 1, "update": 2, "phase": 3, "filter_mm": 4}[kernel]
        _check(lib().gd_time_kernel(self.h, kid, int(reps), C.byref(us)), self.h)
        return us.value
