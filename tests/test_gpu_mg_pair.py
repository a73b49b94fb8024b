"""The multigrid's finest-level smoothing pairs as one out-of-place pass
(k_mg_pair, tpls_params.mg_pair = 1) give the fields of the two in-place
colour sweeps per pair (mg_pair = -1) bit for bit: a first-colour cell reads
only old second-colour values and a second-colour cell only new first-colour
ones, in both forms.  FP64 and FP32 V-cycles, V(1,1) to V(3,3), and grids
whose march chunks are 1, 2 and 4 levels long (the pass forms the first
colour one level beyond each end of its chunk), three steps each; and level
1's pairs, which run the same way where that level has 2 M unknowns."""
from dataclasses import replace

import numpy as np
import pytest

from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres")
BASE = CaseParams(nx=64, ny=16, nz=32, Re=100.0, mu_plus=2.0, dt=1e-4, seed=9, noise=1e-2, ksp_pc="mg")
GRIDS = {"64x16x32": BASE, "128x32x64": replace(BASE, nx=128, ny=32, nz=64),
         "256x64x32": replace(BASE, nx=256, ny=64, nz=32)}


def fields(p, init, precision, mg_fp32, pair):
    from tpls_amd.case import Case
    with Case(p, precision=precision, mg_fp32=mg_fp32, mg_pair=pair) as c:
        c.load(*init)
        c.step(3)
        return {f: c.field(f) for f in FIELDS}, c.diag()["ksp_its"]


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("precision,mg_fp32", [("fp64", False), ("fp64", True), ("fp32", None)])
def test_pair_pass_equals_the_two_sweeps(grid, precision, mg_fp32):
    from tpls_amd.initial import make_initial
    for nu in (1, 2, 3):
        p = replace(GRIDS[grid], ksp_mg_nu=nu)
        init = make_initial(p)
        print(f"[mg-pair] {grid} {precision} mg_fp32={mg_fp32} V({nu},{nu})", flush=True)
        ref, its_ref = fields(p, init, precision, mg_fp32, -1)
        got, its = fields(p, init, precision, mg_fp32, 1)
        assert its == its_ref
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (grid, precision, mg_fp32, nu, f)


@pytest.mark.parametrize("precision,mg_fp32", [("fp64", False), ("fp32", None)])
def test_level_one_pairs_at_full_size(precision, mg_fp32):
    """At 512x128x256 level 1 (256x64x128, 2 M unknowns) has a second vector
    as well, and its smoothing pairs run as one pass each too: one step,
    V(1,1) and V(2,2), bit for bit the in-place sweeps of every level."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    for nu in (1, 2):
        p = replace(BASE, nx=512, ny=128, nz=256, ksp_mg_nu=nu)
        init = make_initial(p)
        print(f"[mg-pair] 512x128x256 {precision} V({nu},{nu})", flush=True)
        out = []
        for pair in (-1, 1):
            with Case(p, precision=precision, mg_fp32=mg_fp32, mg_pair=pair) as c:
                c.load(*init)
                c.step(1)
                out.append(({f: c.field(f) for f in FIELDS}, c.diag()["ksp_its"]))
        (ref, its_ref), (got, its) = out
        assert its == its_ref
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (precision, nu, f)
