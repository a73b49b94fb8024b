"""The multigrid-preconditioned pressure solve on the GPU against the oracle.

1. One V-cycle z = M r (HIP kernels through the library's test hook) equals
   the oracle's orc_mg_apply: both evaluate the same operations in the same
   order without FMA contraction, and every colour sweep is order-free, so
   the bound is a few ulps (1e-13 normwise relative).
2. Whole timesteps with ksp_pc = mg against the oracle's step driver with the
   same preconditioner: 1e-12 after one step, 1e-9 after many, iteration
   counts within one.
3. The multigrid solution agrees with the plain-CG solution to the solver
   tolerance.
"""
import dataclasses

import numpy as np
import pytest

from oracle import binding as B
from tests.compare import demean, relerr, rows
from tests.test_gpu_step import run_pair
from tpls_amd.config import CaseParams, CHANNEL_128x32x64

pytestmark = pytest.mark.gpu

GRIDS = [(16, 8, 12, (1.0, 1.0, 1.0)), (36, 16, 28, (2.0, 1.0, 3.0)),
         (64, 32, 64, (1.0, 4.0, 0.25)), (128, 32, 64, (4096.0, 4096.0, 4096.0)),
         (9, 4, 7, (1.0, 1.0, 1.0))]


@pytest.mark.parametrize("nu", [1, 2])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
def test_vcycle_matches_oracle(grid, nu):
    from tpls_amd import _lib
    nx, ny, nz, (cx, cy, cz) = grid
    diag = 2 * (cx + cy + cz)
    r = np.random.default_rng(nx * 7 + nu).standard_normal(nx * ny * nz)
    print(f"[mg-vcycle] grid {nx}x{ny}x{nz} nu {nu}", flush=True)
    z_o, nl_o = B.mg_apply(r, nx, ny, nz, cx, cy, cz, diag, nu)
    z_g, nl_g = _lib.test_mg_apply(r, nx, ny, nz, cx, cy, cz, diag, nu)
    e = relerr(z_g, z_o)
    print(f"[mg-vcycle] levels {nl_o}/{nl_g} relerr {e:.2e} identical {np.array_equal(z_g, z_o)}",
          flush=True)
    assert nl_o == nl_g
    assert e < 1e-13


MG_SMALL = CaseParams(nx=32, ny=16, nz=24, Re=200.0, mu_plus=5.0, dt=1e-4, noise=1e-2, seed=6,
                      ksp_pc="mg")


def test_ten_steps_small_mg():
    run_pair(MG_SMALL, 10, 1, lambda s: 1e-12 if s == 1 else 1e-9, "mg-small-10")


def test_channel_128x32x64_mg_100_steps():
    p = dataclasses.replace(CHANNEL_128x32x64, ksp_pc="mg")
    run_pair(p, 100, 25, lambda s: 1e-12 if s == 1 else 1e-9, "mg-128x32x64")


def test_mg_and_plain_cg_agree():
    """Same case, both preconditioners on the GPU: the fields after 5 steps
    differ by no more than the solver tolerance allows."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    base = CaseParams(nx=64, ny=16, nz=32, Re=100.0, mu_plus=2.0, dt=1e-4, seed=9)
    init = make_initial(base)
    out = {}
    for pc in ("none", "mg"):
        p = dataclasses.replace(base, ksp_pc=pc)
        c = Case(p)
        c.load(*init)
        c.step(5)
        its = c.diag()["ksp_its"]
        out[pc] = ({f: c.field(f) for f in ("u2", "v2", "w2", "phi2")},
                   demean(c.field("pres"), p.maxl, p.maxm - 1, p.maxn), its)
        c.close()
    P = base.maxm - 1
    errs = {f: relerr(rows(out["mg"][0][f], P), rows(out["none"][0][f], P)) for f in out["mg"][0]}
    errs["pres"] = relerr(rows(out["mg"][1], P), rows(out["none"][1], P))
    print("[mg-vs-cg] its", out["none"][2], out["mg"][2],
          " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert out["mg"][2] < out["none"][2] / 5
    for k, v in errs.items():
        assert v < 1e-8, (k, v)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fp32_vcycle_agrees_with_fp64_vcycle(precision):
    """The FP32 V-cycle (the default in FP32 cases) against the FP64 one on
    the same case: the CG around it stays FP64 with the same stopping test,
    so the fields after 5 steps differ by no more than the solver tolerance
    allows (FP64 case: 1e-8, as plain CG against multigrid; FP32 case: the
    FP32 storage's 1e-6), and the iteration counts stay close."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=64, ny=16, nz=32, Re=100.0, mu_plus=2.0, dt=1e-4, seed=9, ksp_pc="mg")
    init = make_initial(p)
    out = {}
    for m32 in (False, True):
        print(f"[mg-fp32] 64x16x32 {precision} fp32 V-cycle {m32}", flush=True)
        with Case(p, precision=precision, mg_fp32=m32) as c:
            c.load(*init)
            its = []
            for _ in range(5):
                c.step(1)
                its.append(c.diag()["ksp_its"])
            out[m32] = ({f: c.field(f) for f in ("u2", "v2", "w2", "phi2")},
                        demean(c.field("pres"), p.maxl, p.maxm - 1, p.maxn), its)
    P = p.maxm - 1
    errs = {f: relerr(rows(out[True][0][f], P), rows(out[False][0][f], P)) for f in out[True][0]}
    errs["pres"] = relerr(rows(out[True][1], P), rows(out[False][1], P))
    print("[mg-fp32] its", out[False][2], out[True][2], " ".join(f"{k}={v:.1e}" for k, v in errs.items()),
          flush=True)
    tol = 1e-8 if precision == "fp64" else 1e-6
    for k, v in errs.items():
        assert v < tol, (k, v)
    for a, b in zip(out[False][2], out[True][2]):
        assert abs(a - b) <= 2
