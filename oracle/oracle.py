"""ctypes front end of the FP64 CPU oracle (oracle/gd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path
(gamil-dycore_amd) never imports it.

The class mirrors the reference's module-global model (state(-1:2),
tend(-2:2), old/new time indices) so tests can drive one primitive at a time:
space_operators (dycore_mod.F90:184), update_state (600), predict_correct
(754), ordinary_diffusion (diffusion_mod.F90:74), diag_run (diag_mod.F90:42).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GD_ORACLE_LIB points the tests at another build of the restatement, e.g. a
# scratch build with a deliberate slip, to show that the pins catch it
_LIB = os.environ.get("GD_ORACLE_LIB") or os.path.join(_HERE, "_build", "libgd_oracle.so")
_LIB_FAST = os.path.join(_HERE, "_build", "libgd_oracle_fast.so")
_REF_FFT = os.path.join(_HERE, "_ref", "libfftpack_ref.so")

ALL_PASS, FAST_PASS, SLOW_PASS = 0, 1, 2
U, V, GD, IAP_U, IAP_V, IAP_GD = range(6)
(T_U_ADV_LON, T_U_ADV_LAT, T_V_ADV_LON, T_V_ADV_LAT, T_FU, T_FV, T_U_PGF,
 T_V_PGF, T_MASS_DIV_LON, T_MASS_DIV_LAT, T_DU, T_DV, T_DGD) = range(13)
_HALF_STATE = (V, IAP_V)
_HALF_TEND = (T_V_ADV_LON, T_V_ADV_LAT, T_FU, T_V_PGF, T_DV)
SPLIT = {"none": 0, "csp1": 1, "csp2": 2, "isp": 3}
ADV = {"center_diff": 0, "upwind": 1}


class _Params(C.Structure):
    _fields_ = [
        ("num_lon", C.c_int), ("num_lat", C.c_int), ("dt", C.c_double),
        ("split_scheme", C.c_int), ("subcycles", C.c_int),
        ("qcon_modified", C.c_int), ("uv_adv_scheme", C.c_int),
        ("upwind_lon_beta", C.c_double), ("upwind_lat_beta", C.c_double),
        ("use_diffusion", C.c_int), ("diffusion_order", C.c_int),
        ("diffusion_coef", C.c_double), ("use_zonal_tend_filter", C.c_int),
        ("n_cutoff", C.c_int), ("cutoff", C.POINTER(C.c_int)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle not built: {_LIB} (run make -C oracle)")
        L = C.CDLL(_LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.og_create.restype = vp
        L.og_create.argtypes = [C.POINTER(_Params)]
        L.og_destroy.argtypes = [vp]
        L.og_set_threads.argtypes = [C.c_int]
        L.og_old.argtypes = [vp]
        L.og_new.argtypes = [vp]
        L.og_get_coef.argtypes = [vp, C.c_int, dp]
        L.og_get_filter.argtypes = [vp, ip, ip, ip, ip]
        L.og_get_field.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.og_set_field.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.og_set_ghs.argtypes = [vp, dp]
        L.og_set_initial_state.argtypes = [vp, dp, dp, dp, dp]
        L.og_start.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.og_reset_cos_lat_at_poles.argtypes = [vp]
        L.og_iap_transform.argtypes = [vp, C.c_int]
        L.og_space_operators.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.og_update_state.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int]
        L.og_inner_product_tt.restype = C.c_double
        L.og_inner_product_tt.argtypes = [vp, C.c_int, C.c_int]
        L.og_inner_product_ts.restype = C.c_double
        L.og_inner_product_ts.argtypes = [vp, C.c_int, C.c_int]
        L.og_predict_correct.restype = C.c_double
        L.og_predict_correct.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.og_time_integrate.argtypes = [vp, C.POINTER(C.c_double)]
        L.og_ordinary_diffusion.argtypes = [vp, C.c_double, C.c_int]
        L.og_diag.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.c_void_p, C.c_void_p]
        L.og_filter_full.argtypes = [vp, C.c_int, dp]
        L.og_filter_half.argtypes = [vp, C.c_int, dp]
        L.og_check_antisymmetry.argtypes = [vp, C.c_int, C.c_int, dp]
        L.og_run.restype = C.c_int
        L.og_run.argtypes = [vp, C.c_int, dp, dp, dp]
        _lib = L
    return _lib


def use_timing_build():
    """Use the -Ofast build (the reference's optimisation level) instead of
    the strict parity build: for CPU timing only.  Call before first use."""
    global _LIB
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB = _LIB_FAST


def set_threads(n):
    lib().og_set_threads(int(n))


class Oracle:
    """One reference model instance (params_mod namelist values as kwargs)."""

    def __init__(self, num_lon, num_lat, dt, split_scheme="csp2", subcycles=4,
                 qcon_modified=True, uv_adv_scheme="center_diff",
                 upwind_lon_beta=0.0, upwind_lat_beta=0.5, use_diffusion=False,
                 diffusion_order=2, diffusion_coef=0.0, use_zonal_tend_filter=True,
                 cutoff=()):
        L = lib()
        self.N, self.M = int(num_lon), int(num_lat)
        self._cut = np.ascontiguousarray(np.asarray(list(cutoff) or [0], dtype=np.int32))
        p = _Params(self.N, self.M, float(dt),
                    SPLIT[split_scheme] if isinstance(split_scheme, str) else int(split_scheme),
                    int(subcycles), int(bool(qcon_modified)),
                    ADV[uv_adv_scheme] if isinstance(uv_adv_scheme, str) else int(uv_adv_scheme),
                    float(upwind_lon_beta), float(upwind_lat_beta), int(bool(use_diffusion)),
                    int(diffusion_order), float(diffusion_coef),
                    int(bool(use_zonal_tend_filter)), len(list(cutoff)),
                    self._cut.ctypes.data_as(C.POINTER(C.c_int)))
        self.h = L.og_create(C.byref(p))
        self.dt = float(dt)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().og_destroy(h)
            self.h = None

    # -- indices -------------------------------------------------------
    @property
    def old(self):
        return lib().og_old(self.h)

    @property
    def new(self):
        return lib().og_new(self.h)

    # -- data access ---------------------------------------------------
    def coef(self, which):
        names = ["full_cos", "half_cos", "full_sin", "half_sin", "full_f", "full_c",
                 "full_dlon", "full_dlat", "half_dlon", "half_dlat", "full_lat", "half_lat"]
        k = names.index(which)
        half = which.startswith("half")
        out = np.zeros(self.M - 1 if half else self.M)
        lib().og_get_coef(self.h, k, out)
        return out

    def filter_table(self):
        ff = np.zeros(self.M, np.int32); hf = np.zeros(self.M - 1, np.int32)
        fc = np.zeros(self.M, np.int32); hc = np.zeros(self.M - 1, np.int32)
        lib().og_get_filter(self.h, ff, hf, fc, hc)
        return ff, hf, fc, hc

    def _rows(self, kind, field):
        half = field in (_HALF_STATE if kind == 0 else _HALF_TEND)
        return self.M - 1 if half else self.M

    def state(self, idx, field):
        out = np.zeros((self._rows(0, field), self.N))
        lib().og_get_field(self.h, 0, idx, field, out)
        return out

    def set_state(self, idx, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self._rows(0, field), self.N)
        lib().og_set_field(self.h, 0, idx, field, a)

    def tend(self, idx, field):
        out = np.zeros((self._rows(1, field), self.N))
        lib().og_get_field(self.h, 1, idx, field, out)
        return out

    def set_ghs(self, ghs):
        lib().og_set_ghs(self.h, np.ascontiguousarray(ghs, dtype=np.float64))

    # -- reference entry points ----------------------------------------
    def set_initial_state(self, u, v, gd, ghs):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        lib().og_set_initial_state(self.h, f(u), f(v), f(gd), f(ghs))

    def start(self):
        tm, te = C.c_double(), C.c_double()
        lib().og_start(self.h, C.byref(tm), C.byref(te))
        return tm.value, te.value

    def reset_cos_lat_at_poles(self):
        lib().og_reset_cos_lat_at_poles(self.h)

    def iap_transform(self, s):
        lib().og_iap_transform(self.h, s)

    def space_operators(self, s, t, pass_):
        lib().og_space_operators(self.h, s, t, pass_)

    def update_state(self, dt, t, s_old, s_new):
        lib().og_update_state(self.h, dt, t, s_old, s_new)

    def inner_product(self, t1, t2):
        return lib().og_inner_product_tt(self.h, t1, t2)

    def predict_correct(self, dt, s_old, s_new, pass_, qcon=True):
        return lib().og_predict_correct(self.h, dt, s_old, s_new, pass_, int(bool(qcon)))

    def time_integrate(self):
        b = C.c_double()
        lib().og_time_integrate(self.h, C.byref(b))
        return b.value

    def ordinary_diffusion(self, dt, s):
        lib().og_ordinary_diffusion(self.h, dt, s)

    def diag(self, s, with_fields=False):
        tm, te = C.c_double(), C.c_double()
        if with_fields:
            div = np.zeros((self.M, self.N)); vor = np.zeros((self.M - 1, self.N))
            lib().og_diag(self.h, s, C.byref(tm), C.byref(te),
                          div.ctypes.data_as(C.c_void_p), vor.ctypes.data_as(C.c_void_p))
            return tm.value, te.value, div, vor
        lib().og_diag(self.h, s, C.byref(tm), C.byref(te), None, None)
        return tm.value, te.value

    def filter_full(self, j, x):
        y = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().og_filter_full(self.h, j, y)
        return y

    def filter_half(self, j, x):
        y = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().og_filter_half(self.h, j, y)
        return y

    def check_antisymmetry(self, t, s):
        out = np.zeros(4)
        lib().og_check_antisymmetry(self.h, t, s, out)
        return out

    def run(self, nsteps):
        mass = np.zeros(nsteps); energy = np.zeros(nsteps); beta = np.zeros(nsteps)
        nan_step = lib().og_run(self.h, int(nsteps), mass, energy, beta)
        return mass, energy, beta, nan_step


class RefFFTPACK:
    """The reference's own FFTPACK 5.1 (lib/fft), compiled into oracle/_ref.

    rfft1i/rfft1f/rfft1b with the call pattern of filter_mod.F90:62-68,119,124.
    """

    def __init__(self, n):
        if not os.path.exists(_REF_FFT):
            raise FileNotFoundError(_REF_FFT)
        self.L = C.CDLL(_REF_FFT)
        self.n = int(n)
        lensav = n + int(np.log(float(n)) / np.log(2.0)) + 4
        self.wsave = np.zeros(lensav)
        self.work = np.zeros(n)
        ier = C.c_int(0)
        self.L.rfft1i_(C.byref(C.c_int(n)), self.wsave.ctypes.data_as(C.c_void_p),
                       C.byref(C.c_int(lensav)), C.byref(ier))
        assert ier.value == 0

    def _call(self, fn, x):
        r = np.ascontiguousarray(x, dtype=np.float64).copy()
        ier = C.c_int(0)
        n = self.n
        fn(C.byref(C.c_int(n)), C.byref(C.c_int(1)), r.ctypes.data_as(C.c_void_p),
           C.byref(C.c_int(n)), self.wsave.ctypes.data_as(C.c_void_p),
           C.byref(C.c_int(self.wsave.size)), self.work.ctypes.data_as(C.c_void_p),
           C.byref(C.c_int(n)), C.byref(ier))
        assert ier.value == 0
        return r

    def forward(self, x):
        return self._call(self.L.rfft1f_, x)

    def backward(self, c):
        return self._call(self.L.rfft1b_, c)

    def filter(self, x, mask):
        """filter_mod.F90:119-124: forward, multiply by the 0/1 factor, backward."""
        return self.backward(self.forward(x) * mask)


def ref_fftpack_available():
    return os.path.exists(_REF_FFT)
