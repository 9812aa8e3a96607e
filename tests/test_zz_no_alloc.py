"""trt_convergence_noalloc (no counterpart in the reference; for callers that cannot
allocate) through the C ABI against trt_convergence on the same values: two calls in a
row, the second on a smaller input, every norm type, first and later iterations."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """Device memory for the test's inputs and outputs, from the HIP runtime the library
    itself is linked against (already loaded with it)."""

    def __init__(self):
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0          # host to device
        return p

    def get(self, p, n):
        a = np.empty(n)
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0          # device to host
        return a

    def free(self, p):
        assert self.L.hipFree(p) == 0


def test_noalloc_convergence_equals_context_entry_point(trt):
    L = trt.native.load()
    hip = _Hip()
    rng = np.random.default_rng(11)
    out = hip.put(np.full(3, -7.0))
    for nx, ny in ((64, 48), (9, 7)):                   # the second call on a smaller input
        prob = trt.problem.small(nx=nx, ny=ny, order=4, ng=2)
        ctx = trt.native.Context(prob)
        x, xo, xo2 = (rng.uniform(0.5, 2.0, (nx, ny)) for _ in range(3))
        xo[3, 2] = -1.5                                 # a sign change for norm type 2
        dev = [hip.put(a.ravel(order="F")) for a in (x, xo, xo2)]
        for name, a in zip(("Temp", "Temp_RTold", "Temp_RTold2"), (x, xo, xo2)):
            ctx.push(name, a)
        for ct in (1, 2, 3):
            for its in (1, 4):
                for eps in (1e-2, 1e3):
                    rc = L.trt_convergence_noalloc(dev[0], dev[1], dev[2], nx * ny, ct, eps, its, out, None)
                    assert rc == 0, L.trt_last_error(None).decode()
                    assert hip.L.hipDeviceSynchronize() == 0
                    norm, rho, conv = hip.get(out, 3)
                    ref = ctx.convergence("Temp", "Temp_RTold", "Temp_RTold2", ct, eps, its)
                    assert (bool(conv), norm, rho) == ref, (nx, ct, its, eps)
        for p in dev:
            hip.free(p)
        ctx.close()
    # argument errors come back through trt_last_error(NULL), with nothing launched
    rc = L.trt_convergence_noalloc(out, out, out, 3, 4, 1e-2, 1, out, None)
    assert rc != 0 and b"conv_type" in L.trt_last_error(None)
    rc = L.trt_convergence_noalloc(out, out, out, 1 << 31, 1, 1e-2, 1, out, None)
    assert rc != 0 and b"2147483647" in L.trt_last_error(None)
    rc = L.trt_convergence_noalloc(out, out, out, 0, 1, 1e-2, 1, out, None)
    assert rc != 0 and b"empty" in L.trt_last_error(None)
    hip.free(out)
