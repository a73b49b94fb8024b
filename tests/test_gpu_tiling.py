"""The Helmholtz sweeps' launch shapes change no computed value.  Every shape
the autotuner may pick -- tile rows, the column sweep's prefetch depth, the
fused SOR pass (both colour sweeps of an iteration in one out-of-place
kernel) with its tile rows, the Jacobi sweep with staged planes -- gives the
fields of the reference's schedule
(two in-place colour sweeps, default shapes) bit for bit.  The grids cover
partial x tiles, y periods that the fused tile rows do not divide, and more
than one k chunk of the even-parity pass."""
import numpy as np
import pytest

from tpls_amd.autotune import ODD, ROWS
from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres")
GRIDS = {
    "72x12x20": CaseParams(nx=72, ny=12, nz=20, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=21, mom_u=6,
                           mom_v=6, mom_w=6),
    "130x8x38": CaseParams(nx=130, ny=8, nz=38, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=22, mom_u=8,
                           mom_v=8, mom_w=8),
    "64x16x24": CaseParams(nx=64, ny=16, nz=24, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=23),
}


def run(p, init, precision, tiles, split):
    """split: True (two sweeps everywhere), False (fused everywhere), None
    (the default schedule) or a tpls_params.sor_fused code."""
    from tpls_amd.case import Case
    fused = split if isinstance(split, int) and not isinstance(split, bool) else (None if split is None else not split)
    with Case(p, precision=precision, tiles=tiles, sor_fused=fused) as c:
        c.load(*init)
        c.step(3)
        return {f: c.field(f) for f in FIELDS}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_launch_shape_gives_the_same_fields(precision):
    from tpls_amd.initial import make_initial
    p = GRIDS["72x12x20"]
    init = make_initial(p)
    ref = run(p, init, precision, None, True)
    shapes = [((r, 4, 4, 4, 8, -1), True) for r in ROWS] + [((4, ro, d), True) for ro, d in ODD]
    shapes += [((r, 4, 4, rb, rbo), False) for r in (1, 4) for rb in (2, 4, 8) for rbo in (2, 4, 8)]
    shapes += [((4, 2, 8, 4, 8, jst), split) for jst in (2, 4, 8, 16) for split in (True, False, None)]
    for t, split in shapes:
        got = run(p, init, precision, t, split)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (t, split, f)
    print(f"[tiling] {precision}: {len(shapes)} launch shapes identical", flush=True)


@pytest.mark.parametrize("grid", ["130x8x38", "64x16x24"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_sor_pass_matches_the_two_sweeps(grid, precision):
    from tpls_amd.initial import make_initial
    p = GRIDS[grid]
    init = make_initial(p)
    ref = run(p, init, precision, None, True)
    for rb in (2, 4, 8):
        got = run(p, init, precision, (4, 2, 8, rb, rb), False)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (grid, rb, f)
    for jst in (2, 4, 8, 16):
        got = run(p, init, precision, (4, 2, 8, 4, 8, jst), True)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (grid, "jst", jst, f)
    print(f"[sor pass] {grid} {precision}: fused == split for tile rows 2/4/8", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_components_with_different_iteration_counts(precision):
    """In FP32 the three components' Helmholtz loops run in lock-step; an
    odd iteration that all three reach runs their fused passes as one launch
    (launch_sor_pair3), the iterations past the shortest loop one component
    at a time (FP64 runs each loop whole).  Outer-loop counts 8, 6 and 4
    (u, v, w) give the reference's schedule bit for bit."""
    import dataclasses
    from tpls_amd.initial import make_initial
    p = dataclasses.replace(GRIDS["64x16x24"], mom_u=8, mom_v=6, mom_w=4)
    init = make_initial(p)
    ref = run(p, init, precision, None, True)
    for t in ((4, 2, 8, 8, 8), (4, 2, 8, 8, 8, 8), (4, 2, 8, 4, 4)):
        got = run(p, init, precision, t, False if len(t) == 5 else 5)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (t, f)


# a column of 40 levels (u, v: 40, w: 39 updated levels per column)
TALL = CaseParams(nx=64, ny=16, nz=41, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=24, mom_u=8, mom_v=8,
                  mom_w=8)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_march_chunk_length_changes_no_value(precision):
    """The staged Jacobi and even-parity SOR sweeps march each block up a
    chunk of levels, prefetching 8 levels ahead into registers and one plane
    ahead into LDS.  Chunks shorter than, equal to and longer than the
    prefetch depth (1, 8, 9, 17 levels; the 16-row Jacobi tiles prefetch 4
    levels: 4 and 5), and one chunk of the whole column, give the fields of
    the unstaged sweeps bit for bit."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = TALL
    init = make_initial(p)
    ref = run(p, init, precision, (4, 4, 4, 8, 8, -1), True)
    for ks in (1, 4, 5, 8, 9, 17, p.nz):
        for jst in (4, 8, 16):
            with Case(p, precision=precision, tiles=(4, 4, 4, 8, 8, jst), sor_fused=False, march_levels=ks) as c:
                c.load(*init)
                c.step(3)
                for f in FIELDS:
                    assert np.array_equal(c.field(f), ref[f]), (ks, jst, f)
    print(f"[march chunk] {precision}: chunks of 1/4/5/8/9/17/{p.nz} levels identical", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_jacobi_fused_with_the_first_colour_sweep(precision):
    """sor_fused bit 2: an even iteration's Jacobi step, its boundary
    conditions and the first colour sweep in one out-of-place pass
    (launch_jsor0).  Every grid, alone and with the odd iterations fused as
    well, gives the fields of the reference's schedule bit for bit: u's outlet
    extrapolation, the inlet copies and w's walls are the boundary kernel's.
    A period that 8 does not divide (72x12x20) takes the unfused path."""
    from tpls_amd.initial import make_initial
    for grid in ("64x16x24", "130x8x38", "72x12x20"):
        p = GRIDS[grid]
        init = make_initial(p)
        ref = run(p, init, precision, None, True)
        for code in (4, 5):
            for jst in (4, 8):
                got = run(p, init, precision, (4, 2, 8, 4, 8, jst), code)
                for f in FIELDS:
                    assert np.array_equal(got[f], ref[f]), (grid, code, jst, f)
    print(f"[jsor] {precision}: fused Jacobi + first colour sweep identical", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_jacobi_colour_pass_chunks(precision):
    """The fused pass forms J one level below and above its chunk of levels:
    chunks of 1, 2, 3, 4, 5, 8, 9 and 17 levels and the whole column give the
    fields of the unfused sweeps bit for bit (u, v: 40 levels, w: 39)."""
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = TALL
    init = make_initial(p)
    ref = run(p, init, precision, (4, 4, 4, 8, 8, -1), True)
    for ks in (1, 2, 3, 4, 5, 8, 9, 17, p.nz):
        with Case(p, precision=precision, tiles=(4, 4, 4, 8, 8, 8), sor_fused=5, march_levels=ks) as c:
            c.load(*init)
            c.step(3)
            for f in FIELDS:
                assert np.array_equal(c.field(f), ref[f]), (ks, f)
    print(f"[jsor chunk] {precision}: chunks identical", flush=True)
