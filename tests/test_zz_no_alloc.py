"""tpls_get_difference_noalloc (the entry point that allocates nothing) against
tpls_get_difference (host arrays, jacobi_iteration_allflux.f90:285-311) on
the same seeded arrays: both run the same reduction kernels with the same
block count, so the sums are identical.  Two calls in a row, the second on
smaller arrays, so the second call runs over the scratch the first one left.

The device buffers the caller passes are allocated here with the HIP runtime
the library itself is linked against (loaded by soname through ctypes), so
no second runtime enters the process.
"""
import ctypes as C

import numpy as np
import pytest

from tests.fields import rand, shapes

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from tpls_amd import _lib
    _lib.lib()                              # loads libtpls_hip.so and its HIP runtime
    h = C.CDLL("libamdhip64.so.7")          # the same, already loaded, runtime
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipDeviceSynchronize.restype = C.c_int
    return h


def _fn():
    from tpls_amd import _lib
    f = _lib.lib().tpls_get_difference_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 6 + [C.c_void_p, C.c_void_p]
    return f


class _Dev:
    """A device buffer (freed on exit), filled from / read into numpy."""

    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a.ravel(order="F"))
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(self.p)


def test_noalloc_difference_equals_the_allocating_entry_point():
    from tpls_amd import _lib
    hip, f = _hip(), _fn()
    for (maxl, maxm, maxn, kind), seed in zip([(65, 33, 41, "uv"), (21, 9, 17, "w")], (1, 2)):
        S = shapes(maxl, maxm, maxn)
        r = np.random.default_rng(seed)
        a, b = rand(r, S[kind]), rand(r, S[kind])
        ex, ey, ez = maxl - 1, maxm - 1, S[kind][2] - 1
        want = _lib.call("get_difference", a, b, ex, ey, ez, 1, 1, 0)[0]
        with _Dev(hip, a.nbytes) as da, _Dev(hip, b.nbytes) as db, _Dev(hip, 8) as out:
            da.put(a)
            db.put(b)
            out.put(np.full(1, -1.0))
            rc = f(da.p, db.p, ex, ey, ez, 1, 1, 0, out.p, None)
            assert rc == 0, _lib.lib().tpls_last_error()
            assert hip.hipDeviceSynchronize() == 0
            got = out.get(1)[0]
        print(f"[noalloc] {maxl}x{maxm}x{maxn} {kind}: {got!r} vs {want!r}", flush=True)
        assert got == want


def test_noalloc_rejects_host_memory_and_other_decompositions():
    from tpls_amd import _lib
    hip, f = _hip(), _fn()
    a = np.zeros((10, 6, 5), order="F")
    with _Dev(hip, a.nbytes) as d, _Dev(hip, 8) as out:
        assert f(a.ctypes.data, a.ctypes.data, 8, 4, 4, 1, 1, 0, out.p, None) == -2
        assert b"device memory" in _lib.lib().tpls_last_error()
        assert f(d.p, d.p, 8, 4, 4, 2, 1, 0, out.p, None) == -3
        assert f(None, d.p, 8, 4, 4, 1, 1, 0, out.p, None) == -2
