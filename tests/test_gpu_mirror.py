"""The per-routine entry points on arrays attached to the device mirror
against the same calls on plain host arrays (every call copying its arrays
both ways): the same kernels run on the same values, so every output must
be bit-identical.  The sequences follow main_ns_hybrid.F90's use of the
arrays between calls (host statements that read and write them, halo
exchanges, a routine that binds one array twice), so each state change of
the mirror is crossed; the byte counters show what stayed resident.
"""
import numpy as np
import pytest

from tests.fields import level_set, rand, shapes

pytestmark = pytest.mark.gpu

MAXL, MAXM, MAXN = 37, 17, 29
SX, EX, SY, EY, EXMAX = 1, MAXL - 1, 1, MAXM - 1, MAXL - 1
DX = 1.0 / (MAXN - 1)
DT = 2e-4
S = shapes(MAXL, MAXM, MAXN)


@pytest.fixture(scope="module")
def L():
    from tpls_amd import _lib
    return _lib


def run_twice(L, make, seq):
    """seq(arrays) on fresh copies, plain and then attached; returns both
    results (a dict of arrays and scalars) and the attached run's counters."""
    plain = make()
    out0 = seq(plain)
    att = make()
    c0 = L.mirror_counters()
    for a in att.values():
        if isinstance(a, np.ndarray) and a.dtype == np.float64:
            L.mirror_attach(a)
    try:
        out1 = seq(att)
    finally:
        for a in att.values():
            if isinstance(a, np.ndarray) and a.dtype == np.float64:
                L.mirror_detach(a)
    c1 = L.mirror_counters()
    assert c1["attached"] == c0["attached"]
    for k in out0:
        assert np.array_equal(np.asarray(out0[k]), np.asarray(out1[k])), k
    for k in plain:   # every array, after detach, holds what the plain run left
        assert np.array_equal(plain[k], att[k]), k
    return out1, {k: c1[k] - c0[k] for k in c1}


def test_helmholtz_loop_keeps_rhs_and_viscosity_resident(L):
    """main_ns_hybrid.F90:758-808: u3_old = u3, exchanges, Jacobi, the inlet/
    outlet statements, two SOR half sweeps, get_difference."""
    def make():
        r = np.random.default_rng(41)
        return {"x": np.zeros(S["uv"], order="F"), "xo": np.zeros(S["uv"], order="F"),
                "R": rand(r, S["uv"]), "mu": rand(r, S["aug"], scale=0.1, offset=1.0, hw=2),
                "p": rand(r, S["p"])}

    def seq(a):
        x, xo, R, mu, p = a["x"], a["xo"], a["R"], a["mu"], a["p"]
        diffs = []
        for it in range(1, 5):
            xo[...] = x
            L.exchange2d(xo, EX, EY, MAXN - 2, SX, SY, 0)
            L.call("do_jacobi_u", x, xo, R, mu, DX, DX, DX, DT, EX, EY, SX, SY, MAXN, EXMAX)
            x[0, SY:EY + 1, :] = x[1, SY:EY + 1, :]
            x[EXMAX, SY:EY + 1, :] = x[EXMAX - 1, SY:EY + 1, :] + (DT / DX) * (
                p[EXMAX + 1, SY:EY + 1, 1:MAXN] + p[EXMAX - 1, SY:EY + 1, 1:MAXN] - 2.0 * p[EXMAX, SY:EY + 1, 1:MAXN])
            x[EXMAX + 1, SY:EY + 1, :] = 0.0
            L.exchange2d(x, EX, EY, MAXN - 2, SX, SY, 0)
            L.call("do_sor_u", x, R, mu, DX, DX, DX, DT, EX, EY, SX, SY, MAXN, EXMAX, 0, it)
            L.exchange2d(x, EX, EY, MAXN - 2, SX, SY, 0)      # the device holds x: runs there
            L.call("do_sor_u", x, R, mu, DX, DX, DX, DT, EX, EY, SX, SY, MAXN, EXMAX, 1, it)
            diffs.append(L.call("get_difference", x, xo, EX, EY, MAXN - 2, SX, SY, 0)[0])
        return {"x": x.copy(), "diff": np.array(diffs)}

    _, c = run_twice(L, make, seq)
    nb = np.zeros(S["uv"]).nbytes
    # R and mu (and p) go up once, when attached; per iteration x goes up
    # twice and xo once, x comes down twice: nowhere near the 10 uploads and
    # 3 downloads per iteration of the copying calls
    attach = nb * 3 + np.zeros(S["aug"]).nbytes + np.zeros(S["p"]).nbytes
    print(f"[mirror helmholtz] up {c['up']} B, down {c['down']} B, faults {c['faults']}; "
          f"array {nb} B, attach {attach} B", flush=True)
    assert c["up"] <= attach + 4 * 3.2 * nb
    assert c["down"] <= 4 * 2.2 * nb + 3 * nb


def test_levelset_chain_binds_arrays_twice(L):
    """do_tpls (phi3 refreshed on the device), get_levelset_status and
    do_levelset_iteration (phi_reinit bound as input and output), get_csf and
    the exchanges of its outputs, get_viscosity."""
    def make():
        r = np.random.default_rng(42)
        phi2 = level_set(r, MAXL, MAXM, MAXN, DX, noise=0.005)
        d = {"phi2": phi2, "phi3": np.zeros(S["p"], order="F")}
        for q in range(3):
            d[f"c{q}"] = rand(r, S["p"], scale=0.1)
        for n in ("D", "sg", "fx", "fy", "fz", "mu"):
            d[n] = np.zeros(S["p"], order="F")
        d["pr"] = np.zeros(S["p"], order="F")
        d["ist"] = np.zeros(S["p"], np.int32, order="F")
        return d

    def seq(a):
        e1 = L.call("do_tpls", a["phi2"], a["phi3"], a["c0"], a["c1"], a["c2"], SX, EX, SY, EY, MAXN, EXMAX, 6,
                    0.5, DX, DX, DT)[0]
        L.exchange2d(a["phi3"], EX, EY, MAXN, SX, SY, 0)
        L.call("get_levelset_status", a["phi3"], a["D"], a["ist"], a["sg"], SX, EX, SY, EY, MAXN, DX)
        a["pr"][...] = a["phi3"]
        e2 = L.call("do_levelset_iteration", a["phi3"], a["pr"], a["D"], a["ist"], a["sg"], SX, EX, SY, EY, MAXN,
                    EXMAX, DX, 0.5)[0]
        L.call("get_csf", a["fx"], a["fy"], a["fz"], a["pr"], SX, EX, SY, EY, MAXN, EXMAX, DX, DX, DX, 1.0, 1.5 * DX)
        for n in ("fx", "fy", "fz"):
            L.exchange2d(a[n], EX, EY, MAXN, SX, SY, 0)
        L.call("get_viscosity", a["mu"], a["pr"], EX, EY, SX, SY, MAXN, 1.5 * DX, 500.0, 1.0, 10.0)
        return {"e1": e1, "e2": e2, "ist": a["ist"].copy()}

    run_twice(L, make, seq)


def test_dim_binds_phi_new_twice(L):
    def make():
        r = np.random.default_rng(43)
        k = np.arange(MAXN + 1)[None, None, :]
        base = 0.5 + 0.5 * np.tanh(((k - 0.5) * DX - 0.5) / (2 * np.sqrt(2) * 0.5 * DX))
        d = {"phi": np.asfortranarray(base + 0.02 * r.standard_normal(S["p"]) + 0.5)}
        d["phiold"] = np.asfortranarray(d["phi"] + 1e-3 * r.standard_normal(S["p"]))
        d["conv2"] = rand(r, S["p"])
        d["cprev"] = rand(r, S["p"])
        d["phinew"] = np.asfortranarray(r.uniform(-0.2, 1.2, S["p"]))
        return d

    def seq(a):
        res = []
        for its in (5, 4):   # odd and even: the result lands in either bound copy
            res.append(L.call("dim", MAXL, MAXN, DX, DX, DX, 2e-4, 0.5 * DX, 1.0 / (0.5 * DX) ** 2, a["phi"],
                              a["phiold"], a["conv2"], a["cprev"], a["phinew"], SX, EX, SY, EY, 0.45, its)[0])
        return {"res": np.array(res)}

    run_twice(L, make, seq)


def test_momentum_projection_and_pressure_solve(L):
    def make():
        r = np.random.default_rng(44)
        d = {"u": rand(r, S["uv"], scale=0.01), "v": rand(r, S["uv"], scale=0.01), "w": rand(r, S["w"], scale=0.01),
             "mu": rand(r, S["aug"], scale=0.1, offset=1.0, hw=2)}
        d["w"][:, :, 0] = 0.0
        d["w"][:, :, -1] = 0.0
        for n in ("cu", "cv", "du", "dv"):
            d[n] = np.zeros(S["uv"], order="F")
        d["cw"], d["dw"] = np.zeros(S["w"], order="F"), np.zeros(S["w"], order="F")
        d["rhs"], d["p"] = np.zeros(S["p"], order="F"), np.zeros(S["p"], order="F")
        d["inl"] = np.ascontiguousarray(0.5 + 0.01 * r.standard_normal(MAXN - 1))
        return d

    def seq(a):
        args = (SX, SY, EX, EY, MAXN, EXMAX, DX, DX, DX)
        L.call("get_conv_all", a["cu"], a["cv"], a["cw"], a["mu"], a["u"], a["v"], a["w"], *args)
        L.call("get_diffusion_all", a["du"], a["dv"], a["dw"], a["mu"], a["u"], a["v"], a["w"], *args)
        L.call("get_source_pres", a["rhs"], a["u"], a["v"], a["w"], EX, EY, SX, SY, MAXN, DX, DX, DX, DT)
        L.exchange2d(a["rhs"], EX, EY, MAXN, SX, SY, 0)
        its, rnorm, reason = L.pressure_solve(a["p"], a["rhs"], a["u"], a["inl"], MAXL, MAXM, MAXN, DX, DX, DX, DT,
                                              1e-10, 1e-50, 1e5, 1000)
        L.call("get_uvw_pres", a["u"], a["v"], a["w"], a["p"], EX, EY, SX, SY, MAXN, EXMAX, DX, DX, DX, DT, a["inl"])
        return {"its": its, "rnorm": rnorm, "reason": reason}

    out, _ = run_twice(L, make, seq)
    assert out["reason"] in (2, 3)


def test_advection_and_weno(L):
    def make():
        r = np.random.default_rng(45)
        d = {"phi": level_set(r, MAXL, MAXM, MAXN, DX, noise=0.01), "u": rand(r, S["uv"]),
             "v": rand(r, S["uv"]), "w": rand(r, S["w"]), "c": np.full(S["p"], 3.0, order="F")}
        return d

    def seq(a):
        L.call("do_advect_phi", a["phi"], a["u"], a["v"], a["w"], a["c"], SX, EX, SY, EY, MAXL, MAXN, EXMAX,
               DX, DX, DX)
        return {"c": a["c"].copy()}

    run_twice(L, make, seq)
