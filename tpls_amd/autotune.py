"""Autotuner of the Helmholtz sweeps' launch shapes.

The sweeps have run-time launch parameters that do not change any computed
value: the rows per 64-wide tile of the Jacobi and even-parity SOR sweeps,
the rows per tile of the odd-parity (column) SOR sweep and how many levels
that column sweep loads ahead; with the fused SOR pass (both colour sweeps
of an iteration in one kernel, sor_fused) instead, the rows per tile of the
Jacobi sweep and of the fused pass for each parity.  For the Jacobi sweep
also whether to stage its planes through LDS, and the staged tile's rows.  tune()
times every candidate on the actual grid and precision with the library's
HIP-event kernel timing and returns the fastest; tuned() looks the answer
up in tuning.json (results of earlier tuning runs on the GPU, keyed by grid
and precision) and tunes on a miss.

    python -m tpls_amd.autotune 512x128x256 64x16x32 ... [--precision fp32]

prints the tuned shapes as JSON lines for tuning.json.  A shape is
(rows, rows_odd, prefetch_odd, rows_rb, rows_rb_odd, rows_jst).
"""
from __future__ import annotations

import argparse
import json
import os
from dataclasses import replace

ROWS = (1, 2, 4, 8)
ODD = ((1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8))
RB = (2, 4, 8)
JST = (2, 4, 8)
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
_cache: dict = {}


def key(p, precision: str, fused: bool = False) -> str:
    return f"{p.nx}x{p.ny}x{p.nz}/{precision}" + ("/fused" if fused else "")


def _per_launch(c, cls: str, steps: int) -> float:
    c.time_kernels(cls)
    c.step(steps)
    ms, n = c.kernel_time()
    c.time_kernels(None)
    return ms / max(n, 1)


def _per_launch_shape(c, rows: int, cls: str, steps: int) -> float:
    """Per-launch time of the unstaged Jacobi sweep with `rows`."""
    c.set_tiling(rows, 4, 4, rows_jst=-1)
    return _per_launch(c, cls, steps)


def tune(p, precision: str = "fp64", steps: int = 1, report: dict | None = None, fused: bool = False):
    """The launch shape with the least kernel time on this grid."""
    from .case import Case
    from .initial import make_initial
    q = replace(p, ksp_pc="mg") if p.nx % 2 == 0 else p
    times = {}
    split = not fused
    with Case(q, precision=precision, diagnostics=False, sor_fused=fused) as c:
        c.load(*make_initial(q))
        c.step(1)
        for r in ROWS:
            c.set_tiling(r, 4, 4, rows_jst=-1)
            # even iteration_time sweeps run every other outer iteration
            times[("rows", r)] = _per_launch(c, "jacobi", 2 * steps)
            if split:
                times[("rows", r)] += _per_launch(c, "sor_even", 2 * steps)
        best_rows = min(ROWS, key=lambda r: times[("rows", r)])
        # the staged Jacobi sweep against the best unstaged one
        for r in JST:
            c.set_tiling(best_rows, 4, 4, 4, 8, r)
            times[("jst", r)] = _per_launch(c, "jacobi", 2 * steps)
        times[("jst", -1)] = _per_launch_shape(c, best_rows, "jacobi", 2 * steps)
        best_jst = min((-1,) + JST, key=lambda r: times[("jst", r)])
        if split:
            for ro, d in ODD:
                c.set_tiling(best_rows, ro, d, 4, 8, best_jst)
                times[("odd", ro, d)] = _per_launch(c, "sor_odd", steps)
            best = (best_rows,) + min(ODD, key=lambda t: times[("odd",) + t]) + (4, 8, best_jst)
        else:
            for rb in RB:
                c.set_tiling(best_rows, 2, 8, rb, 8, best_jst)
                times[("rb", rb)] = _per_launch(c, "sor_even", 2 * steps)
            for rb in RB:
                c.set_tiling(best_rows, 2, 8, 4, rb, best_jst)
                times[("rb_odd", rb)] = _per_launch(c, "sor_odd", steps)
            best = (best_rows, 2, 8, min(RB, key=lambda r: times[("rb", r)]),
                    min(RB, key=lambda r: times[("rb_odd", r)]), best_jst)
    if report is not None:
        report.update({"/".join(map(str, k)): round(v * 1e3, 2) for k, v in times.items()})
    return best


def tuned(p, precision: str = "fp64", fused: bool = False):
    k = key(p, precision, fused)
    if k in _cache:
        return _cache[k]
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            table = json.load(f)
        if k in table:
            _cache[k] = tuple(table[k]["tiles"])
            return _cache[k]
    _cache[k] = tune(p, precision, fused=fused)
    return _cache[k]


def main():
    from .config import CHANNEL_512x128x256, CaseParams
    ap = argparse.ArgumentParser()
    ap.add_argument("grids", nargs="+")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--fused", action="store_true", help="tune the fused SOR pass instead")
    a = ap.parse_args()
    for gs in a.grids:
        nx, ny, nz = (int(v) for v in gs.split("x"))
        p = replace(CHANNEL_512x128x256, nx=nx, ny=ny, nz=nz) if nx >= 512 else CaseParams(nx=nx, ny=ny, nz=nz)
        rep: dict = {}
        best = tune(p, a.precision, report=rep, fused=a.fused)
        print(json.dumps({key(p, a.precision, a.fused): {"tiles": list(best), "us_per_launch": rep}}), flush=True)


if __name__ == "__main__":
    main()
