"""Benchmark: Mcell-steps/s of the TPLS timestep on one MI355X.

Workload: the plain-FP64 single channel case on the GPU, 512x128x256 cells
(maxl, maxm, maxn = 513, 129, 257), Re 500, viscosity ratio 10, dt 5e-5,
seeded synthetic initial conditions.  A step is one full timestep
(advection, reinitialisation, CSF, momentum assembly, 45 Helmholtz outer
iterations, CG pressure solve to rtol 1e-10, projection).  The whole state is
resident in HBM before the timed region starts.

Prints one JSON line with the throughput, p50/p99 step latency, the roofline
of the dominant kernel (the CG update kernel, timed with HIP events on the
case's stream during the timed steps) and a CPU baseline (the oracle's C
restatement on the host cores, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", default="512x128x256")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--pc", default="none", choices=["none", "mg"],
                    help="pressure CG preconditioner: none (plain CG) or mg (multigrid V-cycle)")
    ap.add_argument("--mg-nu", type=int, default=2, help="V-cycle pre/post smoothing sweeps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(p, its_gpu: float, threads: int):
    """Time the oracle on the host: the non-pressure part of one full step of
    the same grid, plus a capped CG solve whose per-iteration time is scaled
    to the GPU's mean iteration count."""
    import numpy as np
    from dataclasses import replace
    from oracle import binding as B
    from tpls_amd.initial import make_initial

    cap = 60
    q = replace(p, ksp_maxits=cap)
    B.set_threads(threads)
    st = B.OracleState(q.maxl, q.maxm, q.maxn, q.solver_dict())
    st.load(*make_initial(q))
    t_nonp, t_p, its = st.step_timed()
    st.close()
    t_step = t_nonp + (t_p / max(its, 1)) * its_gpu
    val = p.cells * 1e-6 / t_step
    return {
        "value": round(val, 4), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
        "sample": (f"oracle C restatement, {threads} threads: one {p.nx}x{p.ny}x{p.nz} step with the "
                   f"CG capped at {its} iterations (non-pressure {t_nonp:.2f} s, "
                   f"{t_p / max(its, 1) * 1e3:.1f} ms per CG iteration), CG time scaled to the "
                   f"GPU's mean {its_gpu:.0f} iterations per step"),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs one case per GPU on a single GPU; "
                                   "--gpus must be 1 (see DESIGN.md for how the path shards)"}))
        sys.exit(2)
    import numpy as np
    from tpls_amd.case import Case
    from tpls_amd.config import CHANNEL_512x128x256, scaled
    from tpls_amd.initial import make_initial

    nx, ny, nz = (int(v) for v in a.grid.split("x"))
    p = scaled(CHANNEL_512x128x256, nx=nx, ny=ny, nz=nz, ksp_pc=a.pc, ksp_mg_nu=a.mg_nu)
    init = make_initial(p)
    case = Case(p, precision=a.precision)
    case.load(*init)
    del init
    for _ in range(a.warmup):
        case.step(1)
    case.sync()
    case.time_kernels(True)
    lat = []
    its = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        case.step(1)
        case.sync()
        lat.append(time.perf_counter() - s0)
        its.append(case.diag()["ksp_its"])
    case.sync()
    elapsed = time.perf_counter() - t0
    ms, n = case.kernel_time()
    case.time_kernels(False)
    d = case.diag()
    case.close()

    cells = p.cells
    value = cells * a.steps / elapsed * 1e-6
    lat_ms = np.array(lat) * 1e3
    # cg_x per launch: reads p, x, r and writes x, r -> 40 B per unknown
    n_unknowns = (p.maxl - 1) * (p.maxm - 1) * (p.maxn - 1)
    kbytes = 40.0 * n_unknowns
    avg_ms = ms / max(n, 1)
    achieved = kbytes / (avg_ms * 1e-3) / 1e9 if n else 0.0
    out = {
        "metric": "Mcell-steps/s batched-FP32 vs plain-FP64 at stated error; p50/p99 step latency",
        "value": round(value, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32 (CG in f64)",
        "data": "synthetic (seeded two-layer channel flow)",
        "config": {"workload": f"plain {a.precision} single channel case {nx}x{ny}x{nz} cells, "
                               f"Re {p.Re:g}, viscosity ratio {p.mu_plus / p.mu_minus:g}, dt {p.dt:g}, "
                               f"CG rtol {p.ksp_rtol:g}",
                   "solver": "cg" if a.pc == "none" else f"cg + multigrid V({a.mg_nu},{a.mg_nu})",
                   "cg_its_mean": round(float(np.mean(its)), 1), "ksp_reason": d["ksp_reason"]},
        "roofline": {"kernel": "cg_x", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_us": round(avg_ms * 1e3, 2), "launches": n,
                     "algorithmic_bytes_per_launch": int(kbytes)},
    }
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(p, float(np.mean(its)), threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
