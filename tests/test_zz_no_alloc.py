"""gd_filter_rows_noalloc (the allocation-free variant of gd_filter_rows: rows
in device memory, staged in the library's fixed scratch array) against
gd_filter_rows, bit for bit, on two calls in a row with the second on a
smaller grid, on full and half rows, with a padded leading dimension and on
a caller's stream; plus the size limit of the scratch array.  Device memory
and the stream come from the HIP runtime directly (ctypes)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")

    def ok(self, code, what):
        assert code == 0, f"{what} failed ({code})"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p

    def free(self, p):
        self.ok(self.L.hipFree(p), "hipFree")

    def put(self, p, a):
        a = np.ascontiguousarray(a)
        self.ok(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D), "hipMemcpy")

    def get(self, p, shape):
        a = np.empty(shape, np.float64)
        self.ok(self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), D2H), "hipMemcpy")
        return a

    def stream(self):
        s = C.c_void_p()
        self.ok(self.L.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s

    def sync(self, s):
        self.ok(self.L.hipStreamSynchronize(s), "hipStreamSynchronize")

    def destroy(self, s):
        self.ok(self.L.hipStreamDestroy(s), "hipStreamDestroy")


def _check(gd, hip, N, M, lat0, seed, ld=None, stream=None):
    rng = np.random.default_rng(seed)
    cut = gd.cutoff_from_lat0(N, M, lat0)
    ld = N if ld is None else ld
    with gd.Dycore(N, M, 600.0, zonal_tend_filter_cutoff_wavenumber=cut) as d:
        for half, rows in ((False, M), (True, M - 1)):
            print(f"no-alloc filter rows N={N} M={M} half={half} ld={ld} stream={stream is not None}",
                  flush=True)
            x = rng.standard_normal((rows, N))
            ref = d.filter_rows(x, half=half)
            assert np.abs(ref - x).max() > 1e-3  # flagged rows were filtered
            buf = np.full((rows, ld), -7.25)
            buf[:, :N] = x
            p = hip.alloc(buf.nbytes)
            try:
                hip.put(p, buf)
                if stream is None:
                    d.filter_rows_noalloc(p.value, half=half, ld=ld)
                else:
                    d.filter_rows_noalloc(p.value, half=half, ld=ld, stream=stream.value)
                    hip.sync(stream)
                got = hip.get(p, buf.shape)
            finally:
                hip.free(p)
            np.testing.assert_array_equal(got[:, :N], ref)
            assert np.all(got[:, N:] == -7.25)


def test_filter_rows_noalloc_matches_filter_rows(gd):
    hip = Hip()
    _check(gd, hip, 720, 361, 75.0, 1)
    _check(gd, hip, 180, 91, 60.0, 2, ld=200)  # second call: a smaller input
    s = hip.stream()
    try:
        _check(gd, hip, 360, 181, 70.0, 3, stream=s)
    finally:
        hip.destroy(s)


def test_filter_rows_noalloc_rejects_rows_beyond_scratch(gd):
    hip = Hip()
    N, M = 4096, 361
    cut = gd.cutoff_from_lat0(N, M, 50.0)  # 160 flagged rows x 4096 x 8 B > 4 MiB
    print("no-alloc filter rows: size limit", flush=True)
    with gd.Dycore(N, M, 60.0, zonal_tend_filter_cutoff_wavenumber=cut) as d:
        zero = np.zeros((M, N))
        p = hip.alloc(zero.nbytes)
        try:
            hip.put(p, zero)
            with pytest.raises(gd.DycoreError, match="scratch") as e:
                d.filter_rows_noalloc(p.value)
            assert e.value.code == gd.GD_ERR_INVALID_ARG
            assert np.all(hip.get(p, zero.shape) == 0.0)
        finally:
            hip.free(p)
