"""adf_resample_f32_noalloc, the entry point that allocates nothing, against
adf_resample_f32 on the same inputs, bit for bit: two calls in a row, the
second on a smaller input (the shared scratch keeps the first call's
contents), with and without denoise; a size beyond the scratch is refused.
Device buffers for its inputs and outputs are the caller's, allocated here
through the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

from adfdtd import _lib as L
from adfdtd.solver import context

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    return h


class DevArray:
    def __init__(self, hip, host):
        self.hip, self.nbytes = hip, host.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.nbytes) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


CASES = [  # (n_in, ncol, n_out, t_in0, dt_in, t_out0, dt_out, amp, denoise)
    (1200, 200, 70, -3.1e-4, 5.77e-6, 0.0, 1e-4, 1.0, 0),
    (301, 37, 150, 0.0, 1e-3, -0.01, 2.1e-3, 0.73, 1),
]


def test_noalloc_equals_allocating_entry_point_twice_smaller_second():
    lib, hip = L.load(), _hip()
    ctx = context(0)
    rng = np.random.default_rng(21)
    for (n_in, ncol, n_out, t_in0, dt_in, t_out0, dt_out, amp, den) in CASES:
        x = rng.standard_normal((ncol, n_in)).astype(np.float32)
        x[::3] *= 1e-6  # traces with samples below the denoise threshold
        ref = np.empty((ncol, n_out), np.float32)
        L.check(lib.adf_resample_f32(ctx, x.ctypes.data, n_in, ncol, t_in0, dt_in, t_out0, dt_out, n_out, amp, den,
                                     ref.ctypes.data), "adf_resample_f32")
        din = DevArray(hip, x)
        dout = DevArray(hip, np.full((ncol, n_out), np.nan, np.float32))
        try:
            L.check(lib.adf_resample_f32_noalloc(ctx, din.p, n_in, ncol, t_in0, dt_in, t_out0, dt_out, n_out, amp,
                                                 den, dout.p), "adf_resample_f32_noalloc")
            L.check(lib.adf_ctx_synchronize(ctx), "adf_ctx_synchronize")
            got = dout.get(ref)
        finally:
            din.free()
            dout.free()
        assert np.array_equal(got, ref), (n_in, ncol)
        assert np.isfinite(got).all() and np.abs(got).max() > 0


def test_noalloc_refuses_sizes_beyond_its_scratch():
    lib, hip = L.load(), _hip()
    ctx = context(0)
    one = DevArray(hip, np.zeros(16, np.float32))
    try:
        rc = lib.adf_resample_f32_noalloc(ctx, one.p, 5000, 1000, 0.0, 1.0, 0.0, 1.0, 4, 1.0, 0, one.p)
        assert rc == -1 and b"scratch" in lib.adf_last_error()  # ADF_EINVAL
    finally:
        one.free()
