"""Replay the golden fixtures of tests/golden/ (written from the reference's own
Fortran by tests/golden/make_golden.py) through the oracle, bit for bit, and
through the GPU drop-in entry points and the device-resident case at 1e-12.

These pins hold without the reference checkout: the fixtures carry the seeded
inputs and the reference's outputs.
"""
import glob
import os

import numpy as np
import pytest

from oracle import binding as B
from tests.compare import demean, relerr, rows

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROUTINE_FILES = sorted(f for f in glob.glob(os.path.join(HERE, "*.npz"))
                       if not os.path.basename(f).startswith("step"))
GPU_NAME = {"weno_appox_convection": "wenoappoxconvection"}


def load(path):
    z = np.load(path, allow_pickle=False)
    name = str(z["name"])
    args = []
    for q in range(int(z["argc"])):
        a = z[f"a{q}"]
        args.append(np.asfortranarray(a) if a.ndim else a.item())
    outs = [int(i) for i in z["outs"]]
    want = {i: z[f"o{i}"] for i in outs}
    jr = tuple(int(v) for v in z["jrange"]) if "jrange" in z.files else None
    return name, args, outs, want, [float(s) for s in z["scal"]], jr


def region(a, jr):
    return a[:, jr[0]:jr[1] + 1, :] if jr else a


@pytest.mark.parametrize("path", ROUTINE_FILES, ids=os.path.basename)
def test_oracle_replays_reference_fixture(path):
    name, args, outs, want, scal, jr = load(path)
    got_scal = B.oracle(name, *args)
    for i in outs:
        assert np.array_equal(region(args[i], jr), region(want[i], jr)), (name, i)
    for g, w in zip(got_scal or [], scal):
        assert g == pytest.approx(w, rel=1e-13)


def _step_params(z):
    from tpls_amd.config import CaseParams
    m = int(z["mom"])
    return CaseParams(nx=int(z["nx"]), ny=int(z["ny"]), nz=int(z["nz"]), Re=float(z["Re"]),
                      mu_plus=float(z["mu_plus"]), dt=float(z["dt"]), mom_u=m, mom_v=m, mom_w=m,
                      levelset_its=int(z["levelset_its"]), noise=float(z["noise"]), seed=int(z["seed"]))


def test_oracle_replays_two_reference_steps():
    z = np.load(os.path.join(HERE, "step2.npz"), allow_pickle=False)
    p = _step_params(z)
    st = B.OracleState(p.maxl, p.maxm, p.maxn, p.solver_dict())
    st.load(z["u"], z["v"], z["w"], z["pres"], z["phi"])
    its = [st.step(1)["ksp_its"] for _ in range(2)]
    assert its == [int(v) for v in z["its"]]
    for f, key in (("u2", "u2"), ("v2", "v2"), ("w2", "w2"), ("phi2", "phi2"), ("pres", "pres2")):
        assert np.array_equal(st.field(f), z[key]), f
    st.close()


@pytest.mark.gpu
@pytest.mark.parametrize("path", ROUTINE_FILES, ids=os.path.basename)
def test_gpu_replays_reference_fixture(path):
    from tpls_amd import _lib
    name, args, outs, want, scal, jr = load(path)
    print(f"[golden-gpu] {os.path.basename(path)}", flush=True)
    got_scal = _lib.call(GPU_NAME.get(name, name), *args)
    for i in outs:
        e = relerr(region(args[i], jr), region(want[i], jr))
        assert e < 1e-12, (name, i, e)
    for g, w in zip(got_scal or [], scal):
        assert g == pytest.approx(w, rel=1e-12)


@pytest.mark.gpu
def test_gpu_replays_two_reference_steps():
    from tpls_amd.case import Case
    z = np.load(os.path.join(HERE, "step2.npz"), allow_pickle=False)
    p = _step_params(z)
    P = p.maxm - 1
    with Case(p) as c:
        c.load(z["u"], z["v"], z["w"], z["pres"], z["phi"])
        its = []
        for _ in range(2):
            c.step(1)
            its.append(c.diag()["ksp_its"])
        for f in ("u2", "v2", "w2", "phi2"):
            assert relerr(rows(c.field(f), P), rows(z[f], P)) < 1e-12, f
        e = relerr(rows(demean(c.field("pres"), p.maxl, P, p.maxn), P),
                   rows(demean(z["pres2"], p.maxl, P, p.maxn), P))
        assert e < 1e-12
    assert all(abs(a - int(b)) <= 1 for a, b in zip(its, z["its"]))
