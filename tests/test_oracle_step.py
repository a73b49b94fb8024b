"""The oracle's whole-step driver against a step composed from the reference's
own routines (tests/ref_step.py), bit for bit, plus step-level invariants."""
import numpy as np
import pytest

from oracle import binding as B
from tests.ref_step import RefStep
from tpls_amd.config import CaseParams
from tpls_amd.initial import make_initial

SMALL = CaseParams(nx=16, ny=8, nz=12, Re=100.0, mu_plus=3.0, dt=2e-4, mom_u=6, mom_v=6,
                   mom_w=6, levelset_its=4, noise=1e-2, seed=3)


@pytest.mark.skipif(not B.ref_available(), reason="reference build absent")
def test_step_matches_composed_reference():
    p = SMALL
    init = make_initial(p)
    prm = p.solver_dict()
    st = B.OracleState(p.maxl, p.maxm, p.maxn, prm)
    st.load(*init)
    rs = RefStep(p.maxl, p.maxm, p.maxn, prm, *init)
    for step in range(3):
        d = st.step(1)
        rs.step()
        assert d["ksp_its"] == rs.ksp_its
        assert d["err_phi"] == pytest.approx(rs.err_phi, rel=1e-12)
        for name, arr in (("u2", rs.u2), ("v2", rs.v2), ("w2", rs.w2), ("phi2", rs.phi2),
                          ("pres", rs.pres), ("conv1_u", rs.conv["u"][1]),
                          ("csf_w1", rs.csf["w"][1]), ("conv1_phi", rs.cphi[1]),
                          ("visc2_aug", rs.visc_aug)):
            got = st.field(name)
            if name == "visc2_aug":     # columns -1 and maxl+1 are never written
                got, arr = got[1:-1], arr[1:-1]
            assert np.array_equal(got, arr), (step, name, np.abs(got - arr).max())


def test_oracle_is_thread_count_invariant():
    p = SMALL
    init = make_initial(p)
    out = []
    for nt in (1, 4):
        B.set_threads(nt)
        st = B.OracleState(p.maxl, p.maxm, p.maxn, p.solver_dict())
        st.load(*init)
        st.step(2)
        out.append([st.field(n).copy() for n in ("u2", "v2", "w2", "phi2", "pres")])
    B.set_threads(1)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_quiescent_flat_interface_stays_at_rest():
    """u = v = w = 0, phi = z - h, equal viscosities, no surface tension:
    the step must leave the state at rest (up to the inlet profile, here 0)."""
    p = CaseParams(nx=16, ny=8, nz=12, mu_plus=1.0, scap=0.0, mom_u=4, mom_v=4, mom_w=4,
                   levelset_its=3)
    L, M, N = p.maxl, p.maxm, p.maxn
    u = np.zeros((L + 1, M + 1, N - 1), order="F")
    v = np.zeros_like(u)
    w = np.zeros((L + 1, M + 1, N), order="F")
    pres = np.zeros((L + 1, M + 1, N + 1), order="F")
    z = (np.arange(N + 1) - 0.5) * p.dx
    phi = np.asfortranarray(np.broadcast_to(z - p.height, pres.shape).copy())
    st = B.OracleState(L, M, N, p.solver_dict())
    st.load(u, v, w, pres, phi)
    st.step(2)
    for n in ("u2", "v2", "w2"):
        assert np.abs(st.field(n)).max() < 1e-13, n
    # a flat signed-distance interface is a fixed point of the reinitialisation
    ph = st.field("phi2")
    assert np.abs(ph[1:L, 1:M, 1:N] - phi[1:L, 1:M, 1:N]).max() < 1e-13
