"""main_ns_hybrid.F90's level-set timestep (main:583-918 and the rotation
at :2480-2509) walked call for call through the Fortran host modules
(tests/fortran/drive_main_step.f90: the drop-in procedures, tpls_mpi's
exchanges, the KSP block of pressure_solver, main's own statements between
the calls), on the 128x32x64 channel, against the oracle's step driver.

The driver runs twice: with every array attached to the device mirror (the
arrays stay resident in HBM and cross the host link only where main's own
statements read what a kernel wrote, or the reverse) and without (every call
copies its arrays both ways).  The two runs must agree bit for bit, and both
must match the oracle within 1e-11 after two steps (normwise relative; the
pressure modulo its constant), with the same CG iteration counts to within
one and the same residual prints.  The per-step times
of both, beside tpls_step's on the same case, are printed.
"""
import os
import subprocess
import time

import numpy as np
import pytest

from oracle import binding as B
from tests.compare import demean, relerr, rows
from tpls_amd.config import CHANNEL_128x32x64
from tpls_amd.initial import make_initial

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_main_step")
NSTEPS = 2


def write_input(path, p, init, nsteps, mirror):
    d = p.solver_dict()
    with open(path, "wb") as f:
        np.array([p.maxl, p.maxm, p.maxn, nsteps, mirror, d["mom_u"], d["mom_v"], d["mom_w"],
                  d["levelset_its"], d["ksp_maxits"]], np.int32).tofile(f)
        np.array([d[k] for k in ("dx", "dy", "dz", "dt", "Re", "mu_plus", "mu_minus", "height",
                                 "smooth_width", "scap", "ksp_rtol", "ksp_abstol", "ksp_dtol")],
                 np.float64).tofile(f)
        for a in init:
            np.asfortranarray(a, dtype=np.float64).ravel(order="F").tofile(f)


def read_output(path, p, nsteps):
    rec = np.dtype([("err", "<f8", 4), ("its", "<i4"), ("reason", "<i4"), ("rnorm", "<f8"), ("sec", "<f8")])
    with open(path, "rb") as f:
        steps = np.fromfile(f, rec, nsteps)
        out = {}
        for name in ("u2", "v2", "w2", "pres", "phi2"):
            shp = B.field_shape(name, p.maxl, p.maxm, p.maxn)
            out[name] = np.fromfile(f, np.float64, int(np.prod(shp))).reshape(shp, order="F")
        ctr = np.fromfile(f, np.int64, 6)
    return steps, out, ctr


def run_driver(tmp_path, p, init, nsteps, mirror):
    fin, fout = tmp_path / f"in{mirror}.bin", tmp_path / f"out{mirror}.bin"
    write_input(fin, p, init, nsteps, mirror)
    print(f"[main-step] driver mirror={mirror} {p.nx}x{p.ny}x{p.nz} {nsteps} steps", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return read_output(fout, p, nsteps), r.stdout


def test_main_step_through_the_modules(tmp_path):
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/fortran/drive_main_step is not built")
    p = CHANNEL_128x32x64
    assert p.pc_code() == 0, "the KSP block runs PETSc's CG without a preconditioner"
    init = make_initial(p)

    (s1, f1, c1), log1 = run_driver(tmp_path, p, init, NSTEPS, 1)
    (s0, f0, c0), log0 = run_driver(tmp_path, p, init, NSTEPS, 0)
    assert "KSP_CONVERGED" in log1 and "iterations =" in log1

    # resident and copying runs: the same kernels on the same data
    for name in f0:
        assert np.array_equal(f0[name], f1[name]), name
    assert np.array_equal(s0["its"], s1["its"]) and np.array_equal(s0["err"], s1["err"])
    assert c1[5] == 48 and c1[0] > 0, c1          # every array attached, data moved
    assert c0[0] == 0 and c0[5] == 0, c0          # nothing attached without the mirror

    B.set_threads(8)
    orc = B.OracleState(p.maxl, p.maxm, p.maxn, p.solver_dict())
    orc.load(*init)
    P = p.maxm - 1
    for s in range(NSTEPS):
        d = orc.step(1)
        assert abs(int(s1["its"][s]) - d["ksp_its"]) <= 1, (s, s1["its"][s], d["ksp_its"])
        assert s1["reason"][s] in (2, 3)
        for q, k in enumerate(("err_phi", "err_u", "err_v", "err_w")):
            assert s1["err"][s][q] == pytest.approx(d[k], rel=1e-6), (s, k)
    errs = {f: relerr(rows(f1[f], P), rows(orc.field(f), P)) for f in ("u2", "v2", "w2", "phi2")}
    errs["pres"] = relerr(rows(demean(f1["pres"], p.maxl, P, p.maxn), P),
                          rows(demean(orc.field("pres"), p.maxl, P, p.maxn), P))
    orc.close()
    print("[main-step] after", NSTEPS, "steps:", " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < 1e-11, (k, v)

    # the device-resident case on the same state, for the step time
    from tpls_amd.case import Case
    gpu = Case(p)
    gpu.load(*init)
    gpu.step(1)
    t0 = time.perf_counter()
    gpu.step(NSTEPS)
    t_case = (time.perf_counter() - t0) / NSTEPS
    gpu.close()
    mb = (c1[0] + c1[1]) / NSTEPS / 1e6
    print(f"[main-step] s/step: modules+mirror {s1['sec'].tolist()} ({mb:.0f} MB over the host link per step, "
          f"{c1[2] / NSTEPS:.0f} faults, {(c1[3] + c1[4]) / NSTEPS:.0f} copies per step), "
          f"modules copying {s0['sec'].tolist()}, "
          f"tpls_step {t_case:.4f}", flush=True)
