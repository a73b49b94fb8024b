"""Batched stepping (tpls_batch_step): cases of different sizes and precisions
advanced together give, case by case, exactly what each gives alone (same
kernels, same launch shapes, fixed-order reductions: bit-identical)."""
import dataclasses

import numpy as np
import pytest

from tpls_amd.batch import batch_configs
from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres")


def _mix():
    cs = [CaseParams(nx=32, ny=16, nz=24, seed=1),
          CaseParams(nx=64, ny=16, nz=32, Re=300.0, mu_plus=5.0, seed=2, ksp_pc="mg"),
          CaseParams(nx=48, ny=8, nz=16, Re=80.0, mu_plus=0.5, noise=0.05, seed=3)]
    cs += batch_configs(3, seed=7, grids=((64, 16, 32), (128, 32, 64)))
    # diffuse-interface cases beside the level-set ones
    cs += [CaseParams(nx=32, ny=8, nz=16, Re=90.0, seed=4, idm="dim"),
           CaseParams(nx=48, ny=16, nz=24, Re=120.0, mu_plus=3.0, seed=5, idm="dim", ksp_pc="mg")]
    precs = ["fp64", "fp32", "fp64", "fp32", "fp64", "fp32", "fp64", "fp32"]
    return cs, precs


def _run(cases, precs, steps, together, graphs=False, mg_pair=0):
    from tpls_amd.case import Case, batch_step
    from tpls_amd.initial import make_initial
    hs = []
    for p, pr in zip(cases, precs):
        c = Case(p, precision=pr, graphs=graphs, mg_pair=mg_pair)
        c.load(*make_initial(p))
        hs.append(c)
    if together:
        batch_step(hs, steps)
    else:
        for c in hs:
            c.step(steps)
    out = [({f: c.field(f) for f in FIELDS}, c.diag()) for c in hs]
    for c in hs:
        c.close()
    return out


@pytest.mark.parametrize("graphs,grouped", [(False, 1), (True, 1), (False, 2), (True, 2), (False, 0), (True, 0)])
def test_batch_equals_alone(graphs, grouped):
    """Grouping mode 1: the cases' pressure solves run in lock-step, one
    launch per kernel across them; mode 2: the Helmholtz loops and the
    solve's set-up as well; mode 0: each case on its own stream."""
    from tpls_amd.case import set_batch_grouped
    cases, precs = _mix()
    alone = _run(cases, precs, 4, together=False)
    set_batch_grouped(grouped)
    try:
        mixed = _run(cases, precs, 4, together=True, graphs=graphs)
    finally:
        set_batch_grouped(1)
    for q, ((fa, da), (fb, db)) in enumerate(zip(alone, mixed)):
        print(f"[batch] case {q} {cases[q].nx}x{cases[q].ny}x{cases[q].nz} {precs[q]} "
              f"its {da['ksp_its']}/{db['ksp_its']}", flush=True)
        assert da["ksp_its"] == db["ksp_its"]
        assert da["err_phi"] == db["err_phi"]
        for f in FIELDS:
            assert np.array_equal(fa[f], fb[f]), (q, f)


def test_batch_rejects_a_case_twice():
    from tpls_amd import _lib
    from tpls_amd.case import Case, batch_step
    p = CaseParams(nx=32, ny=16, nz=24)
    with Case(p) as c:
        with pytest.raises(_lib.TplsError, match="twice"):
            batch_step([c, c], 1)


def test_a_failing_case_leaves_the_others_unperturbed():
    """One case's step fails (forced on the host before it launches anything):
    tpls_batch_step_status reports it, the case refuses further steps, and
    every other case completes the same steps bit for bit as alone."""
    from tpls_amd import _lib
    from tpls_amd.case import Case, batch_step_status
    from tpls_amd.initial import make_initial
    cs = [CaseParams(nx=32, ny=16, nz=24, seed=1), CaseParams(nx=48, ny=8, nz=16, Re=80.0, seed=2),
          CaseParams(nx=64, ny=16, nz=32, Re=300.0, mu_plus=5.0, seed=3, ksp_pc="mg")]
    want = []
    for p in cs:
        with Case(p, graphs=True) as c:
            c.load(*make_initial(p))
            c.step(3)
            want.append({f: c.field(f) for f in FIELDS})
    hs = []
    for p in cs:
        c = Case(p, graphs=True)
        c.load(*make_initial(p))
        hs.append(c)
    try:
        assert [rc for rc, _ in batch_step_status(hs, 1)] == [0, 0, 0]
        lib = _lib.lib()
        assert lib.tpls_test_fail_next_step(hs[1]._h) == 0
        st = batch_step_status(hs, 2)
        print(f"[batch failure] statuses {st}", flush=True)
        assert st[0][0] == 0 and st[2][0] == 0 and st[1][0] != 0
        assert "injected" in st[1][1]
        with pytest.raises(_lib.TplsError, match="earlier step"):
            hs[1].step(1)
        for q in (0, 2):
            for f in FIELDS:
                assert np.array_equal(hs[q].field(f), want[q][f]), (q, f)
    finally:
        for c in hs:
            c.close()


@pytest.mark.parametrize("grouped", [1, 0])
def test_batch_with_pair_passes_equals_alone(grouped):
    """With tpls_params.mg_pair = 1 every multigrid case whose ny is a
    multiple of 8 runs its finest level's smoothing pairs as one pass each;
    the grouped V-cycle runs them across the cases (k_group of k_mg_pair).
    Each case batched equals itself alone, bit for bit."""
    from tpls_amd.case import set_batch_grouped
    cases, precs = _mix()
    alone = _run(cases, precs, 4, together=False, mg_pair=1)
    set_batch_grouped(grouped)
    try:
        mixed = _run(cases, precs, 4, together=True, mg_pair=1)
    finally:
        set_batch_grouped(1)
    for q, ((fa, da), (fm, dm)) in enumerate(zip(alone, mixed)):
        assert da["ksp_its"] == dm["ksp_its"], q
        for f in FIELDS:
            assert np.array_equal(fa[f], fm[f]), (q, f)
