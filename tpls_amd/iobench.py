"""Throughput of the data-file format (include/tpls_io.h) on the host.

    python -m tpls_amd.iobench [--grid 512x128x256] [--dir DIR] [--fortran PATH] [--case]

For one pressure-grid field of the grid ((nx+2) x (ny+2) x (nz+2) values,
the size of initial_pressure.dat) it times the library's save_data and
load_data_into_array, and, given the Fortran list-directed driver
(tests/fortran/drive_io), the same write and read done with the statements
of io.F90:60-72 and io.F90:98-109 by the Fortran runtime on one thread.
--case also times, on the GPU, tpls_case_load of the five initial arrays
from host memory, tpls_case_load_config of the grid (six files into HBM)
and tpls_case_save_field of the pressure (HBM to a file).  Prints one JSON
line: seconds, values/s and text MB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import tempfile
import time

import numpy as np


def _grid(s: str):
    nx, ny, nz = (int(v) for v in s.lower().split("x"))
    return nx, ny, nz


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--grid", default="512x128x256")
    ap.add_argument("--dir", default=None, help="where the files go (default: a temporary directory)")
    ap.add_argument("--fortran", default=None, help="tests/fortran/drive_io, to time the Fortran runtime")
    ap.add_argument("--case", action="store_true", help="time tpls_case_load_config of the grid on the GPU")
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    from . import io

    nx, ny, nz = _grid(a.grid)
    maxl, maxm, maxn = nx + 1, ny + 1, nz + 1
    shape = (maxl + 1, maxm + 1, maxn + 1)
    n = int(np.prod(shape))
    rng = np.random.default_rng(0)
    data = np.asfortranarray(rng.standard_normal(shape))
    out = {"grid": a.grid, "values": n, "threads": os.cpu_count()}
    with tempfile.TemporaryDirectory(dir=a.dir) as d:
        f = os.path.join(d, "initial_pressure.dat")
        t = []
        for _ in range(a.reps):
            t0 = time.perf_counter()
            io.save_data(f, "Initial pressure", data, 0, maxl, 0, maxm, 0, maxn)
            t.append(time.perf_counter() - t0)
        mb = os.path.getsize(f) / 1e6
        out["text_MB"] = round(mb, 1)
        out["save_s"] = round(min(t), 3)
        back = np.zeros_like(data)
        t = []
        for _ in range(a.reps):
            t0 = time.perf_counter()
            ierr = io.load_data_into_array(f, back, 0, maxl, 0, maxm, 0, maxn)
            t.append(time.perf_counter() - t0)
        assert ierr == 0 and np.array_equal(back.view(np.int64), data.view(np.int64))
        out["load_s"] = round(min(t), 3)
        out["save_Mvalues_s"] = round(n / out["save_s"] / 1e6, 1)
        out["load_Mvalues_s"] = round(n / out["load_s"] / 1e6, 1)
        out["save_MB_s"] = round(mb / out["save_s"], 1)
        out["load_MB_s"] = round(mb / out["load_s"], 1)
        if a.fortran:
            raw, fdat, fraw = (os.path.join(d, x) for x in ("x.raw", "fortran.dat", "fortran.raw"))
            data.ravel(order="F").tofile(raw)
            t0 = time.perf_counter()
            subprocess.run([a.fortran, "write", raw, fdat, "0", str(maxl), "0", str(maxm), "0", str(maxn)],
                           check=True)
            out["fortran_save_s"] = round(time.perf_counter() - t0, 3)
            t0 = time.perf_counter()
            subprocess.run([a.fortran, "read", f, fraw], check=True)
            out["fortran_load_s"] = round(time.perf_counter() - t0, 3)
            v = np.fromfile(fraw, dtype=np.float64, offset=24)
            assert np.array_equal(v.view(np.int64), data.ravel(order="F").view(np.int64))
            out["fortran_save_Mvalues_s"] = round(n / out["fortran_save_s"] / 1e6, 2)
            out["fortran_load_Mvalues_s"] = round(n / out["fortran_load_s"] / 1e6, 2)
        if a.case:
            from .case import Case
            from .config import CaseParams
            from .initial import make_initial
            p = CaseParams(nx=nx, ny=ny, nz=nz)
            fields = make_initial(p)
            io.write_initial_conditions(d, p, fields)
            with Case(p) as c:
                c.load(*fields)  # first use: staging buffers
                c.sync()
                t0 = time.perf_counter()
                c.load(*fields)
                c.sync()
                out["case_load_arrays_s"] = round(time.perf_counter() - t0, 3)
                t0 = time.perf_counter()
                c.load_config(d)
                c.sync()
                out["case_load_config_s"] = round(time.perf_counter() - t0, 3)
                t0 = time.perf_counter()
                c.save_field("pres", os.path.join(d, "pres.dat"), "pressure")
                out["case_save_field_s"] = round(time.perf_counter() - t0, 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
