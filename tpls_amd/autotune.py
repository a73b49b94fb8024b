"""Autotuner of the Helmholtz sweeps' launch shapes.

The sweeps have three run-time launch parameters that do not change any
computed value: the rows per 64-wide tile of the Jacobi and even-parity SOR
sweeps, the rows per tile of the odd-parity (column) SOR sweep, and how many
levels that column sweep loads ahead.  tune() times every candidate on the
actual grid and precision with the library's HIP-event kernel timing and
returns the fastest; tuned() looks the answer up in tuning.json (results of
earlier tuning runs on the GPU, keyed by grid and precision) and tunes on a
miss.

    python -m tpls_amd.autotune 512x128x256 64x16x32 ... [--precision fp32]

prints the tuned shapes as JSON lines for tuning.json.
"""
from __future__ import annotations

import argparse
import json
import os
from dataclasses import replace

ROWS = (1, 2, 4, 8)
ODD = ((1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8))
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
_cache: dict = {}


def key(p, precision: str) -> str:
    return f"{p.nx}x{p.ny}x{p.nz}/{precision}"


def tune(p, precision: str = "fp64", steps: int = 1, report: dict | None = None):
    """(rows, rows_odd, prefetch_odd) with the least kernel time on this grid."""
    from .case import Case
    from .initial import make_initial
    q = replace(p, ksp_pc="mg") if p.nx % 2 == 0 else p
    with Case(q, precision=precision, diagnostics=False) as c:
        c.load(*make_initial(q))
        c.step(1)
        times = {}
        for r in ROWS:
            c.set_tiling(r, 4, 4)
            tot = 0.0
            for cls in ("jacobi", "sor_even"):
                c.time_kernels(cls)
                c.step(2 * steps)        # even iteration_time sweeps run every other outer iteration
                ms, n = c.kernel_time()
                tot += ms / max(n, 1)
            c.time_kernels(None)
            times[("rows", r)] = tot
        best_rows = min(ROWS, key=lambda r: times[("rows", r)])
        for ro, d in ODD:
            c.set_tiling(best_rows, ro, d)
            c.time_kernels("sor_odd")
            c.step(steps)
            ms, n = c.kernel_time()
            c.time_kernels(None)
            times[("odd", ro, d)] = ms / max(n, 1)
        best_odd = min(ODD, key=lambda t: times[("odd",) + t])
    if report is not None:
        report.update({"/".join(map(str, k)): round(v * 1e3, 2) for k, v in times.items()})
    return (best_rows,) + best_odd


def tuned(p, precision: str = "fp64"):
    k = key(p, precision)
    if k in _cache:
        return _cache[k]
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            table = json.load(f)
        if k in table:
            _cache[k] = tuple(table[k]["tiles"])
            return _cache[k]
    _cache[k] = tune(p, precision)
    return _cache[k]


def main():
    from .config import CHANNEL_512x128x256, CaseParams
    ap = argparse.ArgumentParser()
    ap.add_argument("grids", nargs="+")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    a = ap.parse_args()
    for gs in a.grids:
        nx, ny, nz = (int(v) for v in gs.split("x"))
        p = replace(CHANNEL_512x128x256, nx=nx, ny=ny, nz=nz) if nx >= 512 else CaseParams(nx=nx, ny=ny, nz=nz)
        rep: dict = {}
        best = tune(p, a.precision, report=rep)
        print(json.dumps({key(p, a.precision): {"tiles": list(best), "us_per_launch": rep}}), flush=True)


if __name__ == "__main__":
    main()
