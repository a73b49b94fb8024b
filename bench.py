"""Benchmark: Mcell-steps/s of the TPLS timestep on one MI355X.

Workload: the plain-FP64 single channel case on the GPU, 512x128x256 cells
(maxl, maxm, maxn = 513, 129, 257), Re 500, viscosity ratio 10, dt 5e-5,
seeded synthetic initial conditions.  A step is one full timestep
(advection, reinitialisation, CSF, momentum assembly, 45 Helmholtz outer
iterations, CG pressure solve to rtol 1e-10, projection).  The whole state is
resident in HBM before the timed region starts.

The pressure solve is the multigrid-preconditioned CG by default (--pc mg);
a shorter run of the same case with the plain CG (PETSc's -pc_type none
configuration) is reported beside it under "plain_cg".

Prints one JSON line with the throughput, p50/p99 step latency, the roofline
of the dominant kernel (timed with HIP events on the case's stream during the
timed steps; its HBM traffic per launch comes from the rocprofv3 FETCH_SIZE /
WRITE_SIZE summary in profiles/ when one exists for this grid) and a CPU
baseline (the oracle's C restatement on the host cores, on a bounded sample
of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The batched and serving runs keep one stream per case, which HIP maps onto
# GPU_MAX_HW_QUEUES hardware queues.  Measured on the 64-case FP32 batch:
# 1 queue 253, 2 423, 4 604, 8 344, 16 349, 32 320 Mcell-steps/s, so HIP's
# default of 4 is kept unless the environment sets another value.  Set before
# the HIP runtime starts.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "4")

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", default="512x128x256")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--pc", default="mg", choices=["none", "mg"],
                    help="pressure CG preconditioner: none (plain CG) or mg (multigrid V-cycle)")
    ap.add_argument("--mg-nu", type=int, default=2, help="V-cycle pre/post smoothing sweeps")
    ap.add_argument("--roofline-kernel", default="jacobi", choices=["jacobi", "sor_odd", "sor_even", "cg_x", "advect"])
    ap.add_argument("--second-roofline", default="advect",
                    help="a further kernel class timed the same way and reported under rooflines ('' to skip)")
    ap.add_argument("--roofline-steps", type=int, default=2, help="steps over which the kernel is timed")
    ap.add_argument("--mg32-compare", type=int, default=1,
                    help="also time the FP64 case with its V-cycle in FP32 (CG, residual and stopping test "
                         "stay FP64) and report its difference from the all-FP64 run")
    ap.add_argument("--fp32-compare", type=int, default=1,
                    help="also run the case with FP32 storage over the same steps and report its "
                         "throughput and error against this run (0 to skip)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay each step's launches from HIP graphs (identical results)")
    ap.add_argument("--sor-fused", type=int, default=0,
                    help="SOR iterations run as one fused out-of-place pass instead of two in-place colour "
                         "sweeps (identical results): 0 default (odd iteration_time), -1 none, 1 odd, 2 even, "
                         "3 all")
    ap.add_argument("--tiles", default="auto",
                    help="launch shapes of the Helmholtz sweeps: auto (the autotuner's table) or default")
    ap.add_argument("--batch-cases", type=int, default=64,
                    help="cases of the batched FP32 run reported beside the single case (0 to skip)")
    ap.add_argument("--batch-steps", type=int, default=4, help="timed batched steps")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched workload")
    ap.add_argument("--batch-grouped", type=int, default=2,
                    help="2: the batch's Helmholtz loops and pressure solves run grouped (one launch per kernel "
                         "across the cases; the fastest mode for the 64-case batch); 1: the pressure solves only; "
                         "0: each case on its own stream")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="arrival window of the serving run reported beside the single case (0 to skip)")
    ap.add_argument("--serve-rate", type=float, default=1.0,
                    help="Poisson arrival rate of that run, requests/s")
    ap.add_argument("--serve-steps", default="20,200",
                    help="step range per request of that run (the full workload, python -m "
                         "tpls_amd.serve, asks 50-2000 over 10 minutes)")
    ap.add_argument("--batch-warmup", type=int, default=8,
                    help="untimed batched steps (the step's buffer roles cycle with period 6, so 7 steps "
                         "capture every graph)")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="timed steps of the plain-CG comparison run (0 to skip)")
    ap.add_argument("--dim-steps", type=int, default=3,
                    help="timed steps of the same grid with the diffuse interface method (0 to skip)")
    ap.add_argument("--dim-only", action="store_true", help="run only the diffuse-interface workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 64 Mi doubles three times (the reference "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
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


def cpu_one_core():
    """The oracle on one host core, on the CPU-runnable workload (128x32x64,
    Re 100, viscosity ratio 2, dt 1e-4, the same multigrid CG): two steps
    after one untimed step."""
    from dataclasses import replace
    from oracle import binding as B
    from tpls_amd.config import CHANNEL_128x32x64
    from tpls_amd.initial import make_initial
    q = replace(CHANNEL_128x32x64, ksp_pc="mg")
    B.set_threads(1)
    st = B.OracleState(q.maxl, q.maxm, q.maxn, q.solver_dict())
    st.load(*make_initial(q))
    st.step(1)
    t0 = time.perf_counter()
    st.step(2)
    el = time.perf_counter() - t0
    st.close()
    return {"value": round(q.cells * 2 / el * 1e-6, 4), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement, 1 thread: 2 steps of 128x32x64 (multigrid CG), {el:.2f} s"}


KERNEL_NAMES = {"sor_odd": "k_sor2d", "jacobi": "k_jacobi", "sor_even": "k_sor3d", "cg_x": "k_cg_x",
                "advect": "k_advect_phi"}
FUSED_NAMES = {"sor_odd": "k_rb2d", "sor_even": "k_rb3d"}
STAGED_NAMES = {"jacobi": "k_jacobi_st", "sor_even": "k_sor3d_st"}


def kernel_name(kernel: str, split: bool, staged: bool = True) -> str:
    """The kernel a timed class launches: the fused SOR pass unless the two
    in-place sweeps run; the staged Jacobi / even-parity sweep unless the
    unstaged ones were asked for."""
    if not split and kernel in FUSED_NAMES:
        return FUSED_NAMES[kernel]
    if staged and kernel in STAGED_NAMES:
        return STAGED_NAMES[kernel]
    return KERNEL_NAMES[kernel]


def algorithmic_bytes(p, kernel: str, precision: str = "fp64", split: bool = False) -> float:
    """Algorithmic HBM bytes per launch: each distinct array read or written
    once over the interior.  Jacobi: x_old, RHS, viscosity read, x written,
    32 B per cell; fused SOR pass (both colours): x, RHS, viscosity read, x
    written, 32 B per cell; one in-place SOR colour sweep: x, RHS and
    viscosity read over the colour's half plus the other colour's x, half of
    x written, 24 B per cell; CG update: p, x, r read, x, r written, 40 B per
    unknown.  The Helmholtz launches cycle over u, v, w, so their cell count
    is the mean of the three component grids."""
    L, P, N = p.maxl, p.maxm - 1, p.maxn
    if kernel == "cg_x":
        return 40.0 * (L - 1) * P * (N - 1)          # FP64 in both precisions
    if kernel == "advect":
        # phi2, u2, v2, w2 read, conv2_phi written, per p cell
        return 40.0 * (L - 1) * P * (N - 1) * (0.5 if precision == "fp32" else 1.0)
    cells = ((L - 2) * P * (N - 1) + (L - 1) * P * (N - 1) + (L - 1) * P * (N - 2)) / 3.0
    per = 32.0 if kernel == "jacobi" or not split else 24.0
    return per * cells * (0.5 if precision == "fp32" else 1.0)


def measured_traffic(name: str, nx: int, ny: int, nz: int, precision: str):
    """HBM bytes per launch of the named kernel from the committed rocprofv3
    PMC summary of this grid, or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{nx}x{ny}x{nz}_{precision}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == name:
            try:
                return int(float(cells[-1]))
            except ValueError:
                return None
    return None


def plain_cg_run(p, a):
    """The same case with the plain CG: a shorter timed run."""
    import numpy as np
    from dataclasses import replace
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    q = replace(p, ksp_pc="none")
    case = Case(q, precision=a.precision, **case_opts(a))
    case.load(*make_initial(q))
    case.step(1)
    case.sync()
    its = []
    t0 = time.perf_counter()
    for _ in range(a.plain_steps):
        case.step(1)
        case.sync()
        its.append(case.diag()["ksp_its"])
    el = time.perf_counter() - t0
    case.close()
    return {"ms_per_step": round(el / a.plain_steps * 1e3, 3),
            "value": round(q.cells * a.plain_steps / el * 1e-6, 4), "unit": "Mcell-steps/s",
            "steps": a.plain_steps, "warmup": 1, "cg_its_mean": round(float(np.mean(its)), 1)}


def dim_run(p, a):
    """The same grid and solver with the diffuse interface method (idm dim,
    max_iteration_dim relaxation sweeps per step in place of the level-set
    update and reinitialisation); the phase field starts from the
    generator's equilibrium profile."""
    import numpy as np
    from dataclasses import replace
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    q = replace(p, idm="dim")
    case = Case(q, precision=a.precision, **case_opts(a))
    case.load(*make_initial(q))
    for _ in range(max(a.warmup, 1)):
        case.step(1)
    case.sync()
    its = []
    t0 = time.perf_counter()
    for _ in range(a.dim_steps):
        case.step(1)
        case.sync()
        its.append(case.diag()["ksp_its"])
    el = time.perf_counter() - t0
    case.close()
    return {"ms_per_step": round(el / a.dim_steps * 1e3, 3),
            "value": round(q.cells * a.dim_steps / el * 1e-6, 4), "unit": "Mcell-steps/s",
            "steps": a.dim_steps, "warmup": max(a.warmup, 1), "max_iteration_dim": q.dim_its,
            "cg_its_mean": round(float(np.mean(its)), 1)}


def copy_probe():
    import torch
    n = 64 << 20
    src = torch.zeros(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe": {"launches": 3, "bytes_read_per_launch": n * 8,
                                     "bytes_written_per_launch": n * 8}}), flush=True)


COMPARED = ("phi2", "u2", "v2", "w2", "pres")


def tiles_of(a):
    return "auto" if a.tiles == "auto" else None


def case_opts(a) -> dict:
    """Launch options of every case the bench creates (none changes a value)."""
    # sor_fused 0: the tuning table's schedule for the grid
    return {"graphs": bool(a.graphs), "tiles": tiles_of(a), "sor_fused": a.sor_fused or None}


def sor_split(a, kernel: str) -> bool:
    """Whether the timed SOR class runs as two in-place sweeps (else the fused pass)."""
    code = 1 if a.sor_fused == 0 else a.sor_fused
    bit = {"sor_odd": 1, "sor_even": 2}.get(kernel, 0)
    return code < 0 or not code & bit


def timed_variant(p, a, precision: str, **opts):
    """The case with other options over the headline run's warm-up and timed
    steps; returns the open case, the seconds of the timed steps and the CG
    iterations of each."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    case = Case(p, precision=precision, **case_opts(a), **opts)
    case.load(*make_initial(p))
    for _ in range(a.warmup):
        case.step(1)
    case.sync()
    its = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        case.step(1)
        case.sync()
        its.append(case.diag()["ksp_its"])
    return case, time.perf_counter() - t0, its


def field_error(p, case, ref) -> dict:
    """Max and L2 relative difference of the case's fields from ref over the
    interior rows (pressure modulo its mean)."""
    import numpy as np
    P = p.maxm - 1
    err = {}
    for f in COMPARED:
        g, r = case.field(f)[:, 1:P + 1], ref[f][:, 1:P + 1]
        if f == "pres":
            g = g[1:p.maxl, :, 1:p.maxn] - g[1:p.maxl, :, 1:p.maxn].mean()
            r = r[1:p.maxl, :, 1:p.maxn] - r[1:p.maxl, :, 1:p.maxn].mean()
        d = g - r
        err[f] = {"max_rel": float(np.abs(d).max() / np.abs(r).max()),
                  "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(r))}
    return err


def fp32_run(p, a, ref, ref_volume):
    """The same case with FP32 storage over the same warm-up and timed steps:
    throughput and the error of the final state against the FP64 run (max and
    L2 relative; pressure modulo its mean; relative liquid-volume difference)."""
    case, el, _ = timed_variant(p, a, "fp32")
    err = field_error(p, case, ref)
    vol = case.volume()
    case.close()
    return {"value": round(p.cells * a.steps / el * 1e-6, 4), "unit": "Mcell-steps/s",
            "ms_per_step": round(el / a.steps * 1e3, 3), "steps": a.steps, "warmup": a.warmup,
            "error_vs_fp64_after_steps": a.warmup + a.steps, "error": err,
            "volume_rel_diff": abs(vol - ref_volume) / ref_volume}


def mg32_run(p, a, ref):
    """The FP64 case with its multigrid V-cycle in FP32 (tpls_params.mg_fp32):
    storage, the step's arithmetic, the CG vectors, inner products and stopping
    test stay FP64, so the solve meets the same tolerance.  Throughput and the
    difference of the final state from the all-FP64 run."""
    import numpy as np
    case, el, its = timed_variant(p, a, "fp64", mg_fp32=True)
    err = field_error(p, case, ref)
    case.close()
    return {"value": round(p.cells * a.steps / el * 1e-6, 4), "unit": "Mcell-steps/s",
            "ms_per_step": round(el / a.steps * 1e3, 3), "steps": a.steps, "warmup": a.warmup,
            "cg_its_mean": round(float(np.mean(its)), 1),
            "max_rel_diff_vs_fp64_mg": max(e["max_rel"] for e in err.values()),
            "diff_after_steps": a.warmup + a.steps}


def batched_run(a):
    """The batched workload: n cases drawn by seed from the batched
    configuration (grids 64x16x32 .. 256x64x128, uneven sizes), FP32 storage
    with the FP64 pressure solve, advanced together with tpls_batch_step.
    Reports the aggregate Mcell-steps/s over all cases and the batch-step
    latency (every case advances one step per batch step)."""
    import numpy as np
    from tpls_amd.batch import batch_configs
    from tpls_amd.case import Case, batch_step, set_batch_grouped
    from tpls_amd.initial import make_initial
    set_batch_grouped(a.batch_grouped)
    cfgs = batch_configs(a.batch_cases, seed=0)
    cases = []
    for p in cfgs:
        # the finest multigrid level's smoothing pairs as one pass each on
        # every grid of the draw (ny a multiple of 8)
        c = Case(p, precision="fp32", diagnostics=False, mg_pair=1, **case_opts(a))
        c.load(*make_initial(p))
        cases.append(c)
    batch_step(cases, a.batch_warmup)
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.batch_steps):
        s0 = time.perf_counter()
        batch_step(cases, 1)
        lat.append(time.perf_counter() - s0)
    el = time.perf_counter() - t0
    its = [c.diag()["ksp_its"] for c in cases]
    for c in cases:
        c.close()
    cells = sum(p.cells for p in cfgs)
    lat_ms = np.array(lat) * 1e3
    return {"value": round(cells * a.batch_steps / el * 1e-6, 4), "unit": "Mcell-steps/s",
            "cases": len(cfgs), "cells": int(cells), "precision": "fp32 (pressure CG in fp64)",
            "steps": a.batch_steps, "warmup": a.batch_warmup, "ms_per_batch_step": round(el / a.batch_steps * 1e3, 3),
            "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "cg_its_mean": round(float(np.mean(its)), 1),
            "grouping": {0: "each case on its own stream", 1: "pressure solves grouped across the cases",
                         2: "Helmholtz loops and pressure solves grouped across the cases"}[a.batch_grouped],
            "grids": "seeded draw from 64x16x32, 128x32x64, 192x48x96, 256x64x128"}


def serving_run(a):
    """Continuous-batching serving: Poisson arrivals of the batched cases plus
    5% at 512x128x256 over a short window, 200 GB HBM budget, FP32 cases;
    stopped at three times the window.  The default rate keeps the offered
    load near the GPU's batched capacity, so latencies are those of a loaded
    but stable server (the full workload, 20 requests/s of 50-2000 steps,
    offers about 100 times that capacity)."""
    from tpls_amd.serve import Server, poisson_requests
    lo, hi = (int(v) for v in a.serve_steps.split(","))
    reqs = poisson_requests(a.serve_rate, a.serve_seconds, seed=1, steps=(lo, hi))
    offered = sum(r.params.cells * r.steps for r in reqs) / a.serve_seconds * 1e-6
    out = Server(200e9, graphs=bool(a.graphs)).run(reqs, horizon=a.serve_seconds, drain=True,
                                                   max_wall=3 * a.serve_seconds)
    out.update({"arrival_rate_per_s": a.serve_rate, "window_s": a.serve_seconds, "steps_per_request": [lo, hi],
                "offered_mcell_steps_per_s": round(offered, 2), "precision": "fp32 (pressure CG in fp64)"})
    return out


def main():
    a = parse()
    if a.copy_probe:
        copy_probe()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "--gpus must be 1: the benchmark measures one GPU (independent cases "
                                   "shard across GPUs by HBM budget, one large case by slabs along x)"}))
        sys.exit(2)
    import numpy as np
    from tpls_amd.case import Case
    from tpls_amd.config import CHANNEL_512x128x256, scaled
    from tpls_amd.initial import make_initial
    if a.batch_only:
        print(json.dumps({"batched_fp32": batched_run(a)}), flush=True)
        return

    nx, ny, nz = (int(v) for v in a.grid.split("x"))
    p = scaled(CHANNEL_512x128x256, nx=nx, ny=ny, nz=nz, ksp_pc=a.pc, ksp_mg_nu=a.mg_nu)
    if a.dim_only:
        print(json.dumps({f"dim_{a.precision}": dim_run(p, a)}), flush=True)
        return
    init = make_initial(p)
    case = Case(p, precision=a.precision, **case_opts(a))
    # the Jacobi and even-parity sweeps run staged unless the shape says -1
    staged = not (case.tiles and len(case.tiles) > 5 and int(case.tiles[5]) < 0)
    case.load(*init)
    del init
    for _ in range(a.warmup):
        case.step(1)
    case.sync()
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
    d = case.diag()
    mg32 = bool(a.mg32_compare) and a.precision == "fp64" and a.pc == "mg"
    keep = (a.fp32_compare or mg32) and a.precision == "fp64"
    ref = {f: case.field(f) for f in COMPARED} if keep else None
    ref_volume = case.volume() if keep else None
    # the dominant kernel (and a second class), timed with HIP events on the
    # case's stream over further steps (launched eagerly: events bracket each
    # launch)
    timed = {}
    for kern in [a.roofline_kernel] + ([a.second_roofline] if a.second_roofline else []):
        case.time_kernels(kern)
        case.step(a.roofline_steps)
        timed[kern] = case.kernel_time()
        case.time_kernels(None)
    case.close()

    plain = None
    if a.plain_steps > 0 and a.pc != "none":
        plain = plain_cg_run(p, a)

    cells = p.cells
    value = cells * a.steps / elapsed * 1e-6
    lat_ms = np.array(lat) * 1e3
    def roofline_of(kern):
        ms, n = timed[kern]
        split = sor_split(a, kern)
        name = kernel_name(kern, split, staged)
        kbytes = algorithmic_bytes(p, kern, a.precision, split)
        avg_ms = ms / max(n, 1)
        achieved = kbytes / (avg_ms * 1e-3) / 1e9 if n else 0.0
        return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(name, nx, ny, nz, a.precision), "avg_launch_us": round(avg_ms * 1e3, 2),
                "launches": n, "timed_over_steps": a.roofline_steps, "algorithmic_bytes_per_launch": int(kbytes)}

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
                               f"Re {p.Re:g}, viscosity ratio {p.mu_plus / p.mu_minus:g}, density ratio 1 "
                               f"(the reference has no density), dt {p.dt:g}, CG rtol {p.ksp_rtol:g}",
                   "solver": "cg" if a.pc == "none" else f"cg + multigrid V({a.mg_nu},{a.mg_nu})",
                   "launch": "hip graphs" if a.graphs else "eager",
                   "tiles": "autotuned" if a.tiles == "auto" else "default",
                   "cg_its_mean": round(float(np.mean(its)), 1), "ksp_reason": d["ksp_reason"]},
        "roofline": roofline_of(a.roofline_kernel),
    }
    if a.second_roofline:
        out["rooflines"] = {k: roofline_of(k) for k in timed}
    if plain:
        out["plain_cg"] = plain
    if a.dim_steps > 0:
        out[f"dim_{a.precision}"] = dim_run(p, a)
    if ref is not None and a.fp32_compare:
        out["fp32"] = fp32_run(p, a, ref, ref_volume)
    if ref is not None and mg32:
        out["fp64_mg_fp32"] = mg32_run(p, a, ref)
    if a.batch_cases > 0:
        out["batched_fp32"] = batched_run(a)
    if a.serve_seconds > 0:
        out["serving_fp32"] = serving_run(a)
    # throughput ratios of the batched FP32 run against the FP64 single case
    # (multigrid CG) and against the plain FP64 path (plain CG, the
    # reference's -pc_type none configuration)
    # the same-solver ratio (multigrid CG on both sides) is the headline one
    ratios = {}
    if "batched_fp32" in out:
        bv = out["batched_fp32"]["value"]
        ratios["batched_fp32_vs_fp64_mg"] = round(bv / out["value"], 3)
        if plain:
            ratios["batched_fp32_vs_plain_fp64_cg"] = round(bv / plain["value"], 2)
    if "fp32" in out:
        ratios["fp32_vs_fp64_mg"] = round(out["fp32"]["value"] / out["value"], 3)
    if "fp64_mg_fp32" in out:
        ratios["fp64_mg_fp32_vs_fp64_mg"] = round(out["fp64_mg_fp32"]["value"] / out["value"], 3)
    if ratios:
        out["ratios"] = ratios
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(p, float(np.mean(its)), threads)
        out["cpu_baseline"]["one_core"] = cpu_one_core()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
