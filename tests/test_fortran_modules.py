"""The Fortran host modules (tpls_amd/fortran/tpls_modules.f90), built with
amdflang, called as main_ns_hybrid.F90 calls the reference modules.

CPU: the modules exist under the reference's module names, and every wrapped
procedure has the reference's dummy argument list.  The argument lists are
compared against the reference source where the checkout exists.
GPU: tests/fortran/drive_modules calls each procedure through the modules
(Fortran -> C ABI -> gfx950 kernels) on the inputs of the golden fixtures, and
the results must match the reference's outputs stored in them to 1e-12.
do_advect_phi and do_tpls take MPI arguments and have no fixture; they are
checked against the oracle.
"""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import binding as B
from tests.compare import demean, relerr
from tests.fields import level_set, rand, shapes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODSRC = os.path.join(ROOT, "tpls_amd", "fortran", "tpls_modules.f90")
MODDIR = os.path.join(ROOT, "tpls_amd", "fortran", "mod")
DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_modules")
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_SRC = "/root/reference/src"

MODULES = {"weno": "weno.f90", "advect_phi": "advect_phi.f90", "levelset": "levelset.f90",
           "two_phase_levelset": "two_phase_levelset.f90", "momentum_allflux": "momentum_allflux.f90",
           "jacobi_iteration_allflux": "jacobi_iteration_allflux.f90",
           "sor_iteration_allflux": "sor_iteration_allflux.f90", "pressure": "pressure.f90",
           "cahn_hilliard_solver": "cahn_hilliard_solver.f90"}
RID = {"weno_appox_convection": 1, "get_levelset_status": 2, "do_levelset_iteration": 3, "get_csf": 4,
       "get_viscosity": 5, "get_conv_all": 6, "get_diffusion_all": 7, "do_jacobi_u": 8, "do_jacobi_v": 9,
       "do_jacobi_w": 10, "do_sor_u": 11, "do_sor_v": 12, "do_sor_w": 13, "get_source_pres": 14,
       "get_uvw_pres": 15, "do_advect_phi": 16, "do_tpls": 17, "dim": 18}


def subroutine_args(text):
    """{name: [dummy argument names]} of every subroutine in Fortran source."""
    text = re.sub(r"!.*", "", text)
    text = re.sub(r"&\s*\n\s*&?", " ", text)
    out = {}
    for m in re.finditer(r"^\s*subroutine\s+(\w+)\s*\(([^)]*)\)", text, re.I | re.M):
        out[m.group(1).lower()] = [a.strip().lower() for a in m.group(2).split(",") if a.strip()]
    return out


def test_modules_built_with_reference_names():
    if not os.path.isdir(MODDIR):
        pytest.skip("Fortran modules not built (amdflang absent)")
    for mod in list(MODULES) + ["tpls_core_c", "tpls_pressure", "tpls_io_native", "tpls_mpi", "tpls_mirror"]:
        assert os.path.exists(os.path.join(MODDIR, mod + ".mod")), mod
    assert os.access(DRIVER, os.X_OK)


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference checkout absent")
def test_argument_lists_match_reference():
    ours = subroutine_args(open(MODSRC).read())
    checked = 0
    for fn in MODULES.values():
        ref = subroutine_args(open(os.path.join(REF_SRC, fn)).read())
        for name, args in ref.items():
            assert name in ours, f"{fn}: {name} not wrapped"
            assert ours[name] == args, (name, ours[name], args)
            checked += 1
    assert checked == 19


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference checkout absent")
def test_pressure_solver_argument_lists_match_reference():
    """The two procedures of pressure_solver.F90 that main calls keep their
    argument lists (pressure_solver.F90:88, 444)."""
    ours = subroutine_args(open(MODSRC).read())
    ref = subroutine_args(open(os.path.join(REF_SRC, "pressure_solver.F90")).read())
    for name in ("output_converged_reason", "copy_pressure_out_of_petsc"):
        assert ours[name] == ref[name], (name, ours[name], ref[name])


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference checkout absent")
def test_mpi_argument_lists_match_reference():
    """Module tpls_mpi keeps mpi.f90's argument lists for the procedures
    main_ns_hybrid.F90 and its set-up call (the exchanges, the decomposition,
    the neighbours, the stride datatypes); exchange2d_augaug* are called only
    inside advect_phi.f90, which the GPU modules replace."""
    ours = subroutine_args(open(MODSRC).read())
    ref = subroutine_args(open(os.path.join(REF_SRC, "mpi.f90")).read())
    names = [n for n in ref if not n.startswith("exchange2d_augaug")]
    assert len(names) == 11
    for name in names:
        assert name in ours, name
        assert ours[name] == ref[name], (name, ours[name], ref[name])


def _write_stream(path, rid, maxl, maxm, maxn, args):
    with open(path, "wb") as f:
        f.write(np.array([rid, maxl, maxm, maxn], dtype=np.int32).tobytes())
        for a in args:
            if isinstance(a, np.ndarray):
                f.write(np.asfortranarray(a).ravel(order="F").tobytes())
            elif isinstance(a, (int, np.integer)):
                f.write(np.int32(a).tobytes())
            else:
                f.write(np.float64(a).tobytes())


def _run_driver(tmp_path, rid, maxl, maxm, maxn, args, outs):
    """Run the driver; outs lists (shape, dtype) of the outputs in order
    (shape () for a scalar).  Returns the output arrays."""
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_stream(fin, rid, maxl, maxm, maxn, args)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    res, off = [], 0
    for shape, dt in outs:
        n = int(np.prod(shape)) if shape else 1
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += n * np.dtype(dt).itemsize
        res.append(a.reshape(shape, order="F") if shape else float(a[0]))
    assert off == len(raw)
    return res


# maxl, maxm, maxn of every fixture (tests/golden/make_golden.py)
FIXTURE_GRID = (9, 5, 9)
FIXTURES = sorted(f for f in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(f).startswith("step"))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_modules_reproduce_reference_fixture(path, tmp_path):
    from tests.test_golden import load, region
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/fortran/drive_modules is not built")
    name, args, outs, want, scal, jr = load(path)
    maxl, maxm, maxn = FIXTURE_GRID
    spec = [(want[i].shape, want[i].dtype) for i in outs] + [((), np.float64) for _ in scal]
    print(f"[fortran] {os.path.basename(path)} via the Fortran modules", flush=True)
    got = _run_driver(tmp_path, RID[name], maxl, maxm, maxn, args, spec)
    for q, i in enumerate(outs):
        if want[i].dtype == np.int32:
            assert np.array_equal(region(got[q], jr), region(want[i], jr)), name
        else:
            e = relerr(region(got[q], jr), region(want[i], jr))
            assert e < 1e-12, (name, i, e)
    for q, w in enumerate(scal):
        assert got[len(outs) + q] == pytest.approx(w, rel=1e-12)


MAXL, MAXM, MAXN = 17, 9, 13
DX = 1.0 / (MAXN - 1)


@pytest.mark.gpu
def test_modules_advect_phi_and_tpls(tmp_path):
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/fortran/drive_modules is not built")
    S = shapes(MAXL, MAXM, MAXN)
    r = np.random.default_rng(31)
    phi = level_set(r, MAXL, MAXM, MAXN, DX, noise=0.01)
    u, v, w = rand(r, S["uv"]), rand(r, S["uv"]), rand(r, S["w"])
    co = np.zeros(S["p"], order="F")
    B.oracle("do_advect_phi", phi.copy(order="F"), u.copy(order="F"), v.copy(order="F"),
             w.copy(order="F"), co, 1, MAXL - 1, 1, MAXM - 1, MAXL, MAXN, MAXL - 1, DX, DX, DX)
    (cg,) = _run_driver(tmp_path, 16, MAXL, MAXM, MAXN, [phi, u, v, w, np.zeros(S["p"]), DX, DX, DX],
                        [(S["p"], np.float64)])
    assert relerr(cg, co) < 1e-12

    cs = [rand(r, S["p"], scale=0.1) for _ in range(3)]
    p3o = np.zeros(S["p"], order="F")
    eo = B.oracle("do_tpls", phi.copy(order="F"), p3o, *[c.copy(order="F") for c in cs], 1, MAXL - 1, 1,
                  MAXM - 1, MAXN, MAXL - 1, 10, 0.5, DX, DX, 2e-4)[0]
    p3g, eg = _run_driver(tmp_path, 17, MAXL, MAXM, MAXN, [phi, np.zeros(S["p"]), *cs, 10, 0.5, DX, DX, 2e-4],
                          [(S["p"], np.float64), ((), np.float64)])
    assert relerr(p3g[:, 1:MAXM], p3o[:, 1:MAXM]) < 1e-12
    assert eg == pytest.approx(eo, rel=1e-9)


@pytest.mark.gpu
def test_module_dim(tmp_path):
    """cahn_hilliard_solver::dim through the Fortran module (reference
    argument list, MPI arguments ignored) against the oracle, bit for bit."""
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/fortran/drive_modules is not built")
    S = shapes(MAXL, MAXM, MAXN)
    r = np.random.default_rng(37)
    k = np.arange(MAXN + 1)[None, None, :]
    base = 0.5 + 0.5 * np.tanh(((k - 0.5) * DX - 0.5) / (2 * np.sqrt(2) * 0.5 * DX))
    phi = np.asfortranarray(base + 0.02 * r.standard_normal(S["p"]) + 0.5)
    phi[:, 0], phi[:, MAXM] = phi[:, MAXM - 1], phi[:, 1]
    conv2, cprev = rand(r, S["p"], scale=0.1), rand(r, S["p"], scale=0.1)
    pnew = np.asfortranarray(np.clip(phi - 0.5, 0, 1))
    dt, epn, Pe, height, its = 2e-4, 0.5 * DX, 1.0 / (0.5 * DX) ** 2, 0.45, 6
    co, no = cprev.copy(order="F"), pnew.copy(order="F")
    B.oracle("dim", MAXL, MAXN, DX, DX, DX, dt, epn, Pe, phi, phi, conv2, co, no, 1, MAXL - 1, 1, MAXM - 1,
             height, its)
    print(f"[fortran] dim via module cahn_hilliard_solver, {its} sweeps", flush=True)
    cg, ng = _run_driver(tmp_path, 18, MAXL, MAXM, MAXN, [phi, phi, conv2, cprev, pnew, its, DX, dt, epn, Pe, height],
                         [(S["p"], np.float64), (S["p"], np.float64)])
    assert np.array_equal(cg, co)
    assert np.array_equal(ng[:, 1:MAXM], no[:, 1:MAXM])


@pytest.mark.gpu
def test_main_pressure_block_through_pressure_solver(tmp_path):
    """main_ns_hybrid.F90:896-913 as main writes it (pres = pres_old;
    KSPSolve(ksp, PETSC_NULL_OBJECT, PETSC_NULL_OBJECT, ierr);
    output_converged_reason; KSPGetResidualNorm; KSPGetIterationNumber;
    copy_pressure_out_of_petsc) through modules tpls_state and
    pressure_solver, against the oracle's solve of the same system."""
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/fortran/drive_modules is not built")
    S = shapes(MAXL, MAXM, MAXN)
    r = np.random.default_rng(41)
    rhs, u3 = rand(r, S["p"]), rand(r, S["uv"])
    inlet = np.ascontiguousarray(r.standard_normal(MAXN - 1))
    p0 = rand(r, S["p"], scale=0.01)
    dt, rtol = 2e-4, 1e-10
    po = p0.copy(order="F")
    reason_o, its_o, _ = B.pressure_solve(po, rhs, u3, inlet, MAXL, MAXM, MAXN, DX, DX, DX, dt, rtol=rtol)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_stream(fin, 19, MAXL, MAXM, MAXN, [p0, rhs, u3, inlet, DX, DX, DX, dt, rtol])
    res = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print("[fortran] main's KSP block:", res.stdout.strip().replace("\n", " | "), flush=True)
    assert "KSP_CONVERGED_RTOL" in res.stdout
    raw = open(fout, "rb").read()
    n = int(np.prod(S["p"]))
    pg = np.frombuffer(raw, np.float64, n).reshape(S["p"], order="F")
    its_g, reason_g = np.frombuffer(raw, np.int32, 2, offset=8 * n)
    assert reason_g == reason_o == 2
    assert abs(int(its_g) - its_o) <= 1
    P = MAXM - 1
    assert relerr(demean(pg, MAXL, P, MAXN)[:, 1:P + 1], demean(po, MAXL, P, MAXN)[:, 1:P + 1]) < 1e-8


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference checkout absent")
def test_io_argument_lists_match_reference():
    """Module tpls_io_native's procedures keep io.F90's dummy argument lists."""
    ours = subroutine_args(open(MODSRC).read())
    ref = subroutine_args(open(os.path.join(REF_SRC, "io.F90")).read())
    for name in ("save_data", "load_data", "load_data_into_array"):
        assert ours[name] == ref[name], (name, ours[name], ref[name])


def test_io_native_module_known_answers(tmp_path):
    """io_test.f90's known answers through the Fortran module (host code, no
    GPU): tests/fortran/drive_io_native.f90."""
    exe = os.path.join(ROOT, "tests", "fortran", "drive_io_native")
    if not os.access(exe, os.X_OK):
        pytest.skip("tests/fortran/drive_io_native not built")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ALL PASSED")
