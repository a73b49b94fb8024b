"""HIP-graph replay of the step (tpls_params.graphs) against eager launches.

A graph is captured per buffer-role pattern of the step and per CG iteration
parity, then replayed; the roles cycle, so 14 steps replay every captured
graph several times, with the program's own eager work (the CG read-backs,
the projection) and synchronisations between replays.  Every step of every
replay must equal the eager run bit for bit."""
import dataclasses

import numpy as np
import pytest

from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres", "conv1_u", "csf_w1", "conv1_phi")


@pytest.mark.parametrize("precision,pc", [("fp64", "mg"), ("fp32", "mg"), ("fp64", "none")])
def test_graph_replay_equals_eager(precision, pc):
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=64, ny=16, nz=32, Re=200.0, mu_plus=4.0, dt=2e-4, noise=1e-2, seed=12, ksp_pc=pc)
    init = make_initial(p)
    eager = Case(p, precision=precision)
    graph = Case(p, precision=precision, graphs=True)
    eager.load(*init)
    graph.load(*init)
    nsteps = 14 if pc == "mg" else 4
    for s in range(1, nsteps + 1):
        eager.step(1)
        graph.step(1)
        de, dg = eager.diag(), graph.diag()
        assert de["ksp_its"] == dg["ksp_its"], s
        assert de["err_u"] == dg["err_u"], s
        for f in FIELDS:
            assert np.array_equal(eager.field(f), graph.field(f)), (s, f)
    print(f"[graphs] {precision} {pc}: {nsteps} steps identical, last its {dg['ksp_its']}", flush=True)
    eager.close()
    graph.close()
