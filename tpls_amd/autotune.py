"""Autotuner of the Helmholtz sweeps' launch shapes.

The sweeps have run-time launch parameters that do not change any computed
value: the rows per 64-wide tile of the Jacobi and even-parity SOR sweeps,
the rows per tile of the odd-parity (column) SOR sweep and how many levels
that column sweep loads ahead; for the iterations that run the fused SOR
pass instead (both colour sweeps in one kernel, sor_fused), the rows per
tile of the fused pass.  For the Jacobi sweep
also whether to stage its planes through LDS, and the staged tile's rows.  tune()
times every candidate on the actual grid and precision with the library's
HIP-event kernel timing and returns the fastest; tuned() looks the answer
up in tuning.json (results of earlier tuning runs on the GPU, keyed by grid
and precision) and tunes on a miss.

    python -m tpls_amd.autotune 512x128x256 64x16x32 ... [--precision fp32]

prints the tuned shapes as JSON lines for tuning.json.  A shape is
(rows, rows_odd, prefetch_odd, rows_rb, rows_rb_odd, rows_jst, sor_fused),
sor_fused with bit 2 when the even iterations' Jacobi step and first colour
sweep run as one pass.
"""
from __future__ import annotations

import argparse
import json
import os
from dataclasses import replace

ROWS = (1, 2, 4, 8)
ODD = ((1, 4), (2, 4), (4, 4), (1, 8), (2, 8), (4, 8))
RB = (2, 4, 8)
JST = (2, 4, 8, 16)
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
_cache: dict = {}


MODES = {0: "", -1: "/split", 3: "/fused"}  # sor_fused codes and their table suffixes


def key(p, precision: str, mode: int = 0) -> str:
    return f"{p.nx}x{p.ny}x{p.nz}/{precision}" + MODES.get(mode, f"/fused{mode}")


def _per_launch(c, cls: str, steps: int) -> float:
    c.time_kernels(cls)
    c.step(steps)
    ms, n = c.kernel_time()
    c.time_kernels(None)
    return ms / max(n, 1)


def _per_step(c, steps: int) -> float:
    """Milliseconds per step over `steps` steps, after one untimed step."""
    import time
    c.step(1)
    c.sync()
    t0 = time.perf_counter()
    c.step(steps)
    c.sync()
    return (time.perf_counter() - t0) / steps * 1e3


def tune(p, precision: str = "fp64", steps: int = 1, report: dict | None = None, mode: int = 0):
    """The launch shape with the least kernel time on this grid.  mode: the
    SOR schedule asked for (a tpls_params.sor_fused code); with 0, the
    default, the faster of the fused pass and the two sweeps is chosen for
    the odd iterations and returned as the shape's sor_fused."""
    from .case import Case
    from .initial import make_initial
    q = replace(p, ksp_pc="mg") if p.nx % 2 == 0 else p
    times = {}
    even_fused = mode in (2, 3)
    even = 2 if even_fused else 0
    with Case(q, precision=precision, diagnostics=False, sor_fused=-1) as c:
        c.load(*make_initial(q))
        c.step(1)
        # the Jacobi sweep (and the even-parity SOR sweep while it is split,
        # whose staged form shares the Jacobi's tile rows), unstaged then staged
        for r in ROWS:
            c.set_tiling(r, 4, 4, 8, 8, -1, even or -1)
            # (the even iteration_time sweeps run every other outer iteration)
            times[("rows", r)] = _per_launch(c, "jacobi", 2 * steps)
            if not even_fused:
                times[("rows", r)] += _per_launch(c, "sor_even", 2 * steps)
        best_rows = min(ROWS, key=lambda r: times[("rows", r)])
        for r in JST:
            c.set_tiling(best_rows, 4, 4, 8, 8, r, even or -1)
            times[("jst", r)] = _per_launch(c, "jacobi", 2 * steps)
            if not even_fused:
                times[("jst", r)] += _per_launch(c, "sor_even", 2 * steps)
        times[("jst", -1)] = times[("rows", best_rows)]
        # a larger tile only when it is clearly faster (1 %): the sweeps' sum
        # ties within noise between 8 and 16 rows in FP64, where the 8-row
        # Jacobi alone is the faster
        fastest = min(times[("jst", r)] for r in (-1,) + JST)
        best_jst = next(r for r in (-1,) + JST if times[("jst", r)] <= 1.01 * fastest)
        best_rb = 8
        if even_fused:
            for rb in RB:
                c.set_tiling(best_rows, 2, 4, rb, 8, best_jst, 2)
                times[("rb", rb)] = _per_launch(c, "sor_even", 2 * steps)
            best_rb = min(RB, key=lambda r: times[("rb", r)])
        # the odd iterations: the two sweeps, then the fused pass
        odd = {}
        if mode in (0, -1, 2):
            for ro, d in ODD:
                c.set_tiling(best_rows, ro, d, best_rb, 8, best_jst, even or -1)
                times[("odd", ro, d)] = _per_launch(c, "sor_odd", steps)
                odd[(ro, d, 0)] = 2 * times[("odd", ro, d)]  # two sweeps per iteration
        if mode in (0, 1, 3):
            for rb in RB:
                c.set_tiling(best_rows, 2, 4, best_rb, rb, best_jst, even + 1)
                times[("rb_odd", rb)] = odd[(2, 4, rb)] = _per_launch(c, "sor_odd", steps)
        ro, d, rbo = min(odd, key=odd.get)
        code = (even + (1 if rbo else 0)) or -1
        best = (best_rows, ro, d, best_rb, rbo or 8, best_jst, code)
        # the even iterations' Jacobi step fused with their first colour
        # sweep (sor_fused bit 2) replaces three launches by one, so only
        # whole steps show its worth; kept when clearly faster (1 %)
        if best_jst > 0 and not even_fused and p.ny % 8 == 0 and mode in (0, 1):
            c.set_tiling(*best)
            times[("jsor", 0)] = _per_step(c, 2 * steps)
            with_jsor = (code if code > 0 else 0) | 4
            c.set_tiling(*best[:6], with_jsor)
            times[("jsor", 1)] = _per_step(c, 2 * steps)
            if times[("jsor", 1)] < 0.99 * times[("jsor", 0)]:
                best = best[:6] + (with_jsor,)
    if report is not None:
        report.update({"/".join(map(str, k)): round(v * 1e3, 2) for k, v in times.items()})
    return best


def tuned(p, precision: str = "fp64", mode: int = 0):
    if mode > 0 and mode & 4:
        # the fused Jacobi + colour pass asked for: the shapes of the schedule
        # without it, with bit 2 set
        b = (mode & 3) or -1
        base = tuned(p, precision, 0 if b == 1 else b)  # 1 is the default schedule
        return base[:6] + (((base[6] if len(base) > 6 and base[6] > 0 else 0) | 4),)
    k = key(p, precision, mode)
    if k in _cache:
        return _cache[k]
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            table = json.load(f)
        if k in table:
            _cache[k] = tuple(table[k]["tiles"])
            return _cache[k]
    _cache[k] = tune(p, precision, mode=mode)
    return _cache[k]


def main():
    from .config import CHANNEL_512x128x256, CaseParams
    ap = argparse.ArgumentParser()
    ap.add_argument("grids", nargs="+")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--mode", type=int, default=0,
                    help="SOR schedule (tpls_params.sor_fused): 0 default (odd iterations fused), -1 none, 3 all")
    a = ap.parse_args()
    for gs in a.grids:
        nx, ny, nz = (int(v) for v in gs.split("x"))
        p = replace(CHANNEL_512x128x256, nx=nx, ny=ny, nz=nz) if nx >= 512 else CaseParams(nx=nx, ny=ny, nz=nz)
        rep: dict = {}
        best = tune(p, a.precision, report=rep, mode=a.mode)
        print(json.dumps({key(p, a.precision, a.mode): {"tiles": list(best), "us_per_launch": rep}}), flush=True)


if __name__ == "__main__":
    main()
