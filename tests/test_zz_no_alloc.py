"""mcp_mlp_forward_noalloc (include/mcpdlp_mlp.h): the allocation-free forward
pass of the warm-start network, called through the C-ABI with caller-owned
device arrays, against the allocating mcp_mlp_forward on the same network.

Bars: bit-identical outputs (same kernels, same BF16 rounding of the input)
on two calls in a row, the second with a smaller input, and the caller's
output array untouched past the n x out values of each call; a batch past the
scratch maximum is refused with an error before anything runs."""
import ctypes as C
import os

import numpy as np
import pytest

import mcpdlp as mp

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    path = "/opt/rocm/lib/libamdhip64.so"
    lib = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipSetDevice.argtypes = [C.c_int]
    return lib


class DevArray:
    """A float32 device array owned by the test (the caller side allocates)."""

    def __init__(self, lib, count):
        self.lib, self.count, self.p = lib, count, C.c_void_p()
        assert lib.hipMalloc(C.byref(self.p), 4 * count) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size <= self.count
        assert self.lib.hipMemcpy(self.p, a.ctypes.data, 4 * a.size, H2D) == 0

    def get(self, count=None):
        out = np.empty(self.count if count is None else count, np.float32)
        assert self.lib.hipMemcpy(out.ctypes.data, self.p, 4 * out.size, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.lib.hipFree(self.p)
            self.p = C.c_void_p()


def test_noalloc_forward_matches_forward_twice():
    assert mp.device_count() >= 1, "no HIP device visible"
    lib = mp.engine_lib()
    lib.mcp_mlp_forward_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.mcp_mlp_forward_noalloc.restype = C.c_int
    lib.mcp_mlp_last_error.restype = C.c_char_p
    rt = hip()
    assert rt.hipSetDevice(0) == 0
    net = mp.MLP(in_dim=256, hidden=256, layers=3, out_dim=64, capacity=128, seed=5)
    rng = np.random.default_rng(17)
    p = net.params()
    net.set_params((rng.standard_normal(p.size) * 0.08).astype(np.float32))
    n_max = 300
    x_dev, y_dev = DevArray(rt, n_max * 256), DevArray(rt, n_max * 64)
    try:
        for n in (300, 70):  # the second call is smaller: stale scratch rows must not leak in
            print("noalloc forward, n =", n, flush=True)
            x = rng.standard_normal((n, 256)).astype(np.float32)
            want = net.forward(x)  # allocating entry point (host arrays, chunks of the capacity)
            x_dev.put(x)
            sentinel = np.full(n_max * 64, 12345.0, np.float32)
            y_dev.put(sentinel)
            rc = lib.mcp_mlp_forward_noalloc(net.h, x_dev.p, n, y_dev.p)
            assert rc == 0, lib.mcp_mlp_last_error()
            got = y_dev.get()
            assert np.array_equal(got[: n * 64].reshape(n, 64), want)
            assert np.array_equal(got[n * 64:], sentinel[n * 64:])  # nothing written past n x out
        # past the scratch maximum (8192 samples at width 256): refused before any launch
        rc = lib.mcp_mlp_forward_noalloc(net.h, x_dev.p, 8193, y_dev.p)
        assert rc == 1 and b"maximum" in lib.mcp_mlp_last_error()
    finally:
        x_dev.free()
        y_dev.free()
        net.close()
