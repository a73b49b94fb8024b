"""The device mirror's coherence protocol (tpls_amd/csrc/mirror.cpp), on the
CPU: mirror.cpp is built with g++ beside a backend whose "device" copy is
host memory (tests/native/mirror_host_backend.cpp), so a test can play the
kernel's part by writing the device copy directly.

What must hold, whatever the access pattern:
- after to_device, the device copy equals the host array;
- after written, every host read sees the device copy;
- only the granules the host touched cross the link (counted in bytes);
- a fault outside any attached array still reaches the previous handler.
"""
import ctypes as C
import os
import subprocess
import sys
import textwrap
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tpls_amd", "csrc", "mirror.cpp"),
       os.path.join(ROOT, "tests", "native", "mirror_host_backend.cpp")]
PAGE = os.sysconf("SC_PAGESIZE")


def build(out_dir) -> str:
    so = os.path.join(out_dir, "libmirror_host.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-shared", "-fPIC",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "tpls_amd", "csrc"),
                    *SRC, "-o", so, "-lpthread"], check=True)
    return so


@pytest.fixture(scope="module")
def so_path(tmp_path_factory):
    return build(str(tmp_path_factory.mktemp("mirror")))


@pytest.fixture(scope="module")
def M(so_path):
    L = C.CDLL(so_path)
    L.tpls_mirror_attach.argtypes = [C.c_void_p, C.c_longlong]
    L.tpls_mirror_detach.argtypes = [C.c_void_p]
    L.tpls_mirror_set_granule.argtypes = [C.c_longlong]
    L.tpls_mirror_counters.argtypes = [C.POINTER(C.c_longlong)]
    L.mh_device_of.argtypes = [C.c_void_p, C.c_longlong]
    L.mh_device_of.restype = C.c_void_p
    for f in ("mh_to_device", "mh_written", "mh_host_current"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_longlong]
    L.mh_host_access.argtypes = [C.c_void_p, C.c_longlong, C.c_int]
    L.mh_error.restype = C.c_char_p
    return L


def counters(M):
    c = (C.c_longlong * 6)()
    assert M.tpls_mirror_counters(c) == 0
    return {"up": c[0], "down": c[1], "faults": c[2], "uploads": c[3], "downloads": c[4], "attached": c[5]}


def delta(a, b):
    return {k: b[k] - a[k] for k in a}


def unaligned(n, seed, offset=24):
    """n doubles starting offset bytes past a page boundary, so the array has
    partial pages at both ends."""
    raw = np.zeros(n * 8 + 2 * PAGE + offset, np.uint8)
    start = (-raw.ctypes.data) % PAGE + offset
    a = raw[start:start + n * 8].view(np.float64)
    a[:] = np.random.default_rng(seed).standard_normal(n)
    return raw, a


def dev(M, a):
    """The device copy of attached array a, as a numpy view (the kernel's side)."""
    p = M.mh_device_of(a.ctypes.data, a.nbytes)
    assert p, "not attached"
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(a.size,))


@pytest.fixture()
def arr(M):
    """A 40-granule array with 1-page granules, detached afterwards."""
    assert M.tpls_mirror_set_granule(PAGE) == 0
    raw, a = unaligned(40 * PAGE // 8 + 100, 1)
    assert M.tpls_mirror_attach(a.ctypes.data, a.nbytes) == 0
    yield a
    M.tpls_mirror_detach(a.ctypes.data)
    M.tpls_mirror_set_granule(1 << 20)
    del raw


def test_attach_uploads_once_and_reads_are_free(M, arr):
    want = arr.copy()
    assert np.array_equal(dev(M, arr), want)
    c0 = counters(M)
    s = float(arr.sum())                  # reads of a CLEAN array: no fault
    assert s == float(want.sum())
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    d = delta(c0, counters(M))
    assert d["faults"] == 0
    # nothing changed: nothing crosses, not even the unprotected end pages
    assert d["up"] == 0 and d["down"] == 0
    # the end pages are never protected: a write there is found by comparison
    arr[0] = 5.0
    arr[-1] = -5.0
    assert counters(M)["faults"] == c0["faults"]
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    d = delta(c0, counters(M))
    assert 0 < d["up"] < 2 * PAGE
    assert dev(M, arr)[0] == 5.0 and dev(M, arr)[-1] == -5.0


def test_host_write_uploads_only_the_touched_granule(M, arr):
    c0 = counters(M)
    arr[arr.size // 2] = 42.0
    assert counters(M)["faults"] - c0["faults"] == 1
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    d = delta(c0, counters(M))
    assert PAGE <= d["up"] < 3 * PAGE, d
    assert np.array_equal(dev(M, arr), arr)
    # a second to_device has nothing to move but the end pages
    c1 = counters(M)
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    assert delta(c1, counters(M))["up"] < 2 * PAGE


def test_device_write_is_fetched_where_the_host_reads(M, arr):
    D = dev(M, arr)
    D[:] = np.arange(arr.size, dtype=np.float64)   # the "kernel"
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    assert M.mh_host_current(arr.ctypes.data, arr.nbytes) == 0
    c0 = counters(M)
    i = 10 * PAGE // 8
    assert arr[i] == float(i)                       # one granule fetched
    d = delta(c0, counters(M))
    assert d["faults"] == 1 and d["down"] == PAGE, d
    assert np.array_equal(arr, np.arange(arr.size, dtype=np.float64))   # the rest
    assert M.mh_host_current(arr.ctypes.data, arr.nbytes) == 1
    assert delta(c0, counters(M))["down"] <= arr.nbytes


def test_host_write_into_stale_data_keeps_the_device_values(M, arr):
    D = dev(M, arr)
    D[:] = 7.0
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    arr[5 * PAGE // 8 + 3] += 1.0                   # read-modify-write of device data
    assert arr[5 * PAGE // 8 + 3] == 8.0
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    want = np.full(arr.size, 7.0)
    want[5 * PAGE // 8 + 3] = 8.0
    assert np.array_equal(D, want) and np.array_equal(arr, want)


def test_whole_array_host_statements(M, arr):
    """u3_old = u3; u3 = 0; boundary-plane updates: every pattern main's own
    statements use, between device writes."""
    D = dev(M, arr)
    D[:] = 3.0
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    old = arr.copy()                                # reads every granule
    assert np.all(old == 3.0)
    arr[:] = 0.0                                    # writes every granule
    arr[::97] = 1.5                                 # a strided plane
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    want = np.zeros(arr.size)
    want[::97] = 1.5
    assert np.array_equal(D, want)


def test_threads_fault_on_their_own_granules(M, arr):
    D = dev(M, arr)
    D[:] = np.arange(arr.size, dtype=np.float64)
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    chunk = arr.size // 8

    def work(t):
        s = slice(t * chunk, (t + 1) * chunk if t < 7 else arr.size)
        arr[s] += 1.0
    with ThreadPoolExecutor(8) as ex:
        list(ex.map(work, range(8)))
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    want = np.arange(arr.size, dtype=np.float64) + 1.0
    assert np.array_equal(arr, want) and np.array_equal(D, want)


def test_host_access_and_detach(M, arr):
    D = dev(M, arr)
    D[:] = -2.0
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    # a copy engine is about to write the middle third: made current and writable
    lo, hi = arr.size // 3, 2 * arr.size // 3
    sub = arr[lo:hi]
    assert M.mh_host_access(sub.ctypes.data, sub.nbytes, 1) == 0
    c0 = counters(M)
    C.memset(sub.ctypes.data, 0, sub.nbytes)        # not through the fault path
    assert counters(M)["faults"] == c0["faults"]
    assert M.mh_to_device(arr.ctypes.data, arr.nbytes) == 0
    assert np.all(D[lo:hi] == 0.0)
    D[:] = 5.0
    assert M.mh_written(arr.ctypes.data, arr.nbytes) == 0
    p = arr.ctypes.data
    assert M.tpls_mirror_detach(p) == 0              # fetches the STALE granules
    c1 = counters(M)
    assert np.all(arr == 5.0)
    arr[:] = 1.0                                     # plain memory again
    assert counters(M)["faults"] == c1["faults"]
    assert M.tpls_mirror_attach(p, arr.nbytes) == 0  # for the fixture's detach


def test_attach_rules(M):
    raw, a = unaligned(4096, 2)
    assert M.tpls_mirror_attach(a.ctypes.data, a.nbytes) == 0
    assert M.tpls_mirror_attach(a.ctypes.data, a.nbytes) == 0            # same bounds: no-op
    assert M.tpls_mirror_attach(a[8:].ctypes.data, 800) != 0             # overlap
    assert b"overlaps" in M.mh_error()
    assert M.mh_device_of(a.ctypes.data, a.nbytes - 8) is None           # bounds must match
    assert M.tpls_mirror_detach(a[1:].ctypes.data) != 0
    assert M.tpls_mirror_detach(a.ctypes.data) == 0
    assert M.tpls_mirror_attach(None, 8) != 0
    # an array with no whole page is never protected, and always synced
    small = np.arange(16, dtype=np.float64)
    assert M.tpls_mirror_attach(small.ctypes.data, small.nbytes) == 0
    c0 = counters(M)
    small[3] = 9.0
    assert counters(M)["faults"] == c0["faults"]
    assert M.mh_to_device(small.ctypes.data, small.nbytes) == 0
    assert dev(M, small)[3] == 9.0
    dev(M, small)[4] = -1.0
    assert M.mh_written(small.ctypes.data, small.nbytes) == 0
    assert small[4] == -1.0
    assert M.tpls_mirror_detach(small.ctypes.data) == 0
    del raw


def test_foreign_faults_reach_the_previous_handler(so_path):
    """A segmentation fault outside every attached array must still end the
    process through the handler that was there before (Python's faulthandler
    here), not be swallowed or retried forever."""
    code = textwrap.dedent(f"""
        import ctypes as C, faulthandler, numpy as np
        faulthandler.enable()
        L = C.CDLL({so_path!r})
        L.tpls_mirror_attach.argtypes = [C.c_void_p, C.c_longlong]
        a = np.zeros(1 << 16)
        assert L.tpls_mirror_attach(a.ctypes.data, a.nbytes) == 0
        a[100] = 1.0
        print("resolved", flush=True)
        C.string_at(16)   # an unmapped address
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert "resolved" in r.stdout
    assert r.returncode == -11, (r.returncode, r.stderr[-500:])
    assert "Fatal Python error: Segmentation fault" in r.stderr
