"""The Helmholtz sweeps' launch shapes (tile rows, column-sweep prefetch
depth) change no computed value: every shape the autotuner may pick gives the
default shape's fields bit for bit."""
import numpy as np
import pytest

from tpls_amd.autotune import ODD, ROWS
from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_launch_shape_gives_the_same_fields(precision):
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=72, ny=12, nz=20, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=21, mom_u=6, mom_v=6,
                   mom_w=6)
    init = make_initial(p)
    ref = None
    shapes = [(r, 4, 4) for r in ROWS] + [(4, ro, d) for ro, d in ODD]
    for t in shapes:
        with Case(p, precision=precision, tiles=t) as c:
            c.load(*init)
            c.step(3)
            got = {f: c.field(f) for f in FIELDS}
        if ref is None:
            ref = got
            continue
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (t, f)
    print(f"[tiling] {precision}: {len(shapes)} launch shapes identical", flush=True)
