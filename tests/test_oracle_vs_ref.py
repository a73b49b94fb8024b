"""The C oracle against the reference's own Fortran, routine by routine.

oracle/_ref/libtpls_ref.so is the reference's weno.f90, levelset.f90,
momentum_allflux.f90, jacobi_iteration_allflux.f90, sor_iteration_allflux.f90
and pressure.f90 compiled unmodified by oracle/Makefile (built only where the
reference checkout exists, so these tests skip elsewhere).  Both sides get
identical seeded inputs; every output array must match BIT FOR BIT over the
region the reference defines.
"""
import numpy as np
import pytest

from oracle import binding as B
from tests.fields import level_set, rand, shapes, wrap_y

pytestmark = pytest.mark.skipif(not B.ref_available(), reason="reference build absent")

MAXL, MAXM, MAXN = 17, 9, 13
SX, EX, SY, EY, EXMAX = 1, MAXL - 1, 1, MAXM - 1, MAXL - 1
DX = 1.0 / (MAXN - 1)


def _same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    if not np.array_equal(a, b):
        d = np.max(np.abs(a - b))
        raise AssertionError(f"{what}: not bit-identical, max |diff| = {d:.3e}")


def both(name, make_args, outs):
    """Run oracle and reference on copies of the same inputs; compare outs."""
    a1 = make_args()
    a2 = [x.copy(order="F") if isinstance(x, np.ndarray) else x for x in a1]
    s1 = B.oracle(name, *a1)
    s2 = B.ref(name, *a2)
    for idx in outs:
        _same(a1[idx], a2[idx], f"{name} arg {idx}")
    return a1, a2, s1, s2


def test_weno5_scalar():
    rng = np.random.default_rng(1)
    L = B.oracle_lib().lib
    for _ in range(200):
        v = rng.standard_normal(5) * rng.choice([1e-4, 1.0, 1e3])
        # weno.f90's WENO5 is a module function; call it through the reference
        r = B.ref_lib().lib._QMwenoPweno5
        import ctypes as C
        r.restype = C.c_double
        args = [C.byref(C.c_double(float(x))) for x in v]
        assert L.orc_weno5(*[float(x) for x in v]) == r(*args)


@pytest.mark.parametrize("seed", [0, 1])
def test_weno_appox_convection(seed):
    rng = np.random.default_rng(seed)
    S = shapes(MAXL, MAXM, MAXN)

    def mk():
        r = np.random.default_rng(seed)
        phi = level_set(r, MAXL, MAXM, MAXN, DX, noise=0.01, hw=3)
        return [phi, rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"]),
                SX, EX, SY, EY, MAXL, MAXN,
                np.zeros(S["uv"], order="F"), np.zeros(S["uv"], order="F"), np.zeros(S["w"], order="F")]
    both("weno_appox_convection", mk, [10, 11, 12])


def test_levelset_status_and_iteration():
    S = shapes(MAXL, MAXM, MAXN)

    def mk():
        r = np.random.default_rng(3)
        phi = level_set(r, MAXL, MAXM, MAXN, DX, noise=0.02)
        return [phi, np.zeros(S["p"], order="F"), np.zeros(S["p"], order="F", dtype=np.int32),
                np.zeros(S["p"], order="F"), SX, EX, SY, EY, MAXN, DX]
    a1, a2, _, _ = both("get_levelset_status", mk, [1, 2, 3])
    assert a1[2].sum() > 0, "test data must cross the interface"

    phi, D, ist, sg = a1[0], a1[1], a1[2], a1[3]
    r = np.random.default_rng(4)
    pr0 = wrap_y(np.asfortranarray(phi + 0.01 * r.standard_normal(phi.shape)), 1)
    pr1, pr2 = pr0.copy(order="F"), pr0.copy(order="F")
    e1 = B.oracle("do_levelset_iteration", phi, pr1, D, ist, sg, SX, EX, SY, EY, MAXN, EXMAX, DX, 0.5)
    e2 = B.ref("do_levelset_iteration", phi, pr2, D, ist, sg, SX, EX, SY, EY, MAXN, EXMAX, DX, 0.5)
    # the reference's phi_reinit_updated is an uninitialised automatic array:
    # only i in 0..maxl, j in sy..ey (all k) is defined
    _same(pr1[:, SY:EY + 1, :], pr2[:, SY:EY + 1, :], "do_levelset_iteration phi_reinit")
    # the reference's reduction(MAX:) clause is inert without OpenMP: the sum
    assert e1[0] == pytest.approx(e2[0], rel=1e-13)


def test_get_csf():
    S = shapes(MAXL, MAXM, MAXN)

    def mk():
        r = np.random.default_rng(5)
        phi = level_set(r, MAXL, MAXM, MAXN, DX, amp=0.1, noise=0.002)
        f0 = rand(r, S["p"])
        return [f0.copy(order="F"), f0.copy(order="F"), f0.copy(order="F"), phi,
                SX, EX, SY, EY, MAXN, EXMAX, DX, DX, DX, 1.0, 1.5 * DX]
    a1, _, _, _ = both("get_csf", mk, [0, 1, 2])
    assert np.abs(a1[0]).max() > 0


def test_get_viscosity():
    S = shapes(MAXL, MAXM, MAXN)

    def mk():
        r = np.random.default_rng(6)
        phi = level_set(r, MAXL, MAXM, MAXN, DX, noise=0.01)
        return [rand(r, S["p"]), phi, EX, EY, SX, SY, MAXN, 1.5 * DX, 100.0, 1.0, 2.0]
    both("get_viscosity", mk, [0])


@pytest.mark.parametrize("name", ["get_conv_all", "get_diffusion_all"])
def test_momentum_flux(name):
    S = shapes(MAXL, MAXM, MAXN)

    def mk():
        r = np.random.default_rng(7)
        mu = np.asfortranarray(0.01 + 0.005 * r.random(S["aug"]))
        wrap_y(mu, 2)
        return [rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"]), mu,
                rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"]),
                SX, SY, EX, EY, MAXN, EXMAX, DX, DX, DX]
    both(name, mk, [0, 1, 2])


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_jacobi(comp):
    S = shapes(MAXL, MAXM, MAXN)
    shp = S["w"] if comp == "w" else S["uv"]

    def mk():
        r = np.random.default_rng(8)
        mu = np.asfortranarray(0.01 + 0.005 * r.random(S["aug"]))
        wrap_y(mu, 2)
        x0 = rand(r, shp)
        return [x0.copy(order="F"), rand(r, shp), rand(r, shp), mu, DX, DX, DX, 1e-3,
                EX, EY, SX, SY, MAXN, EXMAX]
    both("do_jacobi_" + comp, mk, [0])


@pytest.mark.parametrize("comp", ["u", "v", "w"])
@pytest.mark.parametrize("it", [1, 2])
@pytest.mark.parametrize("flg", [0, 1])
def test_sor(comp, it, flg):
    S = shapes(MAXL, MAXM, MAXN)
    shp = S["w"] if comp == "w" else S["uv"]

    def mk():
        r = np.random.default_rng(9)
        mu = np.asfortranarray(0.01 + 0.005 * r.random(S["aug"]))
        wrap_y(mu, 2)
        base = [rand(r, shp), rand(r, shp), mu, DX, DX, DX, 1e-3, EX, EY, SX, SY, MAXN]
        if comp == "u":
            base.append(EXMAX)
        return base + [flg, it]
    both("do_sor_" + comp, mk, [0])


def test_get_difference():
    S = shapes(MAXL, MAXM, MAXN)
    r = np.random.default_rng(10)
    a, b = rand(r, S["uv"]), rand(r, S["uv"])
    d1 = B.oracle("get_difference", a, b, EX, EY, MAXN - 2, SX, SY, 0)
    d2 = B.ref("get_difference", a, b, EX, EY, MAXN - 2, SX, SY, 0)[0]
    assert d1 == pytest.approx(d2, rel=1e-13)


def test_pressure_source_and_projection():
    S = shapes(MAXL, MAXM, MAXN)

    def mk_src():
        r = np.random.default_rng(11)
        return [rand(r, S["p"]), rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"]),
                EX, EY, SX, SY, MAXN, DX, DX, DX, 1e-3]
    both("get_source_pres", mk_src, [0])

    def mk_prj():
        r = np.random.default_rng(12)
        return [rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"]), rand(r, S["p"]),
                EX, EY, SX, SY, MAXN, EXMAX, DX, DX, DX, 1e-3,
                np.ascontiguousarray(r.standard_normal(MAXN - 1))]
    both("get_uvw_pres", mk_prj, [0, 1, 2])
