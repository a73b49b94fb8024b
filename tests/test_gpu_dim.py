"""The diffuse interface method on the GPU against the CPU oracle.

* tpls_dim (cahn_hilliard_solver.f90:41, through the C ABI) against orc_dim
  on identical seeded inputs: FP64 bit for bit (both evaluate every
  expression as written, without FMA contraction), ConvPrevPhi, the iterate
  with its boundary planes, and the last sweep's largest change.
* Whole timesteps of an idm dim case (main_ns_hybrid.F90:583-918 with the
  branch of :606-611) against the oracle's step driver, with the level-set
  step's tolerances: 1e-12 after one step, 1e-9 after ten (normwise
  relative; pressure modulo its constant); FP32 storage: phase field and
  viscosity within 1e-5 after three steps, the flow within 1e-2.
* HIP-graph replay and checkpoint/restore of a dim case, bit for bit.

The oracle is pinned by tests/test_oracle_dim.py.
"""
import numpy as np
import pytest

from oracle import binding as B
from tests.compare import demean, relerr, rows
from tpls_amd.config import CaseParams
from tpls_amd.initial import make_initial

pytestmark = pytest.mark.gpu

MAXL, MAXM, MAXN = 41, 17, 29
SX, EX, SY, EY = 1, MAXL - 1, 1, MAXM - 1
H = 1.0 / (MAXN - 1)
SHAPE = (MAXL + 1, MAXM + 1, MAXN + 1)


def wrapped(a):
    a = np.asfortranarray(a)
    a[:, 0] = a[:, MAXM - 1]
    a[:, MAXM] = a[:, 1]
    return a


@pytest.mark.parametrize("seed,its", [(0, 0), (1, 1), (2, 18), (3, 5)])
def test_dim_routine_is_bit_identical_to_the_oracle(seed, its):
    from tpls_amd import _lib
    rng = np.random.default_rng(seed)
    k = np.arange(MAXN + 1)[None, None, :]
    base = 0.5 + 0.5 * np.tanh(((k - 0.5) * H - 0.5) / (2 * np.sqrt(2) * 0.5 * H))
    phi = wrapped(base + 0.02 * rng.standard_normal(SHAPE) + 0.5)   # phi2 + 0.5, as main passes it
    phiold = wrapped(phi + 1e-3 * rng.standard_normal(SHAPE))
    conv2 = wrapped(rng.standard_normal(SHAPE))
    cprev = wrapped(rng.standard_normal(SHAPE))
    phinew = wrapped(rng.uniform(-0.2, 1.2, SHAPE))
    dt, epn, Pe, height = 2e-4, 0.5 * H, 1.0 / (0.5 * H) ** 2, 0.45
    args = (MAXL, MAXN, H, H, H, dt, epn, Pe)
    co, no = cprev.copy(order="F"), phinew.copy(order="F")
    cg, ng = cprev.copy(order="F"), phinew.copy(order="F")
    print(f"[dim routine] seed {seed} its {its} grid {MAXL}x{MAXM}x{MAXN}", flush=True)
    ro = B.oracle("dim", *args, phi, phiold, conv2, co, no, SX, EX, SY, EY, height, its)[0]
    rg = _lib.call("dim", *args, phi, phiold, conv2, cg, ng, SX, EX, SY, EY, height, its)[0]
    assert np.array_equal(cg, co)
    assert np.array_equal(rows(ng, MAXM - 1), rows(no, MAXM - 1))
    # the y halo rows of the interior columns are the periodic wrap on both sides
    assert np.array_equal(ng[1:MAXL, 0], no[1:MAXL, 0]) and np.array_equal(ng[1:MAXL, MAXM], no[1:MAXL, MAXM])
    assert rg == ro
    if its:
        assert ro > 0


def step_pair(p: CaseParams, nsteps: int, precision: str, tol, shift_phi: float = 0.0):
    from tpls_amd.case import Case
    B.set_threads(8)
    init = list(make_initial(p))
    init[4] = np.asfortranarray(init[4] - shift_phi)
    orc = B.OracleState(p.maxl, p.maxm, p.maxn, p.solver_dict())
    orc.load(*init)
    gpu = Case(p, precision=precision)
    gpu.load(*init)
    P = p.maxm - 1
    for s in range(1, nsteps + 1):
        do = orc.step(1)
        gpu.step(1)
        dg = gpu.diag()
        if precision == "fp64":
            assert abs(do["ksp_its"] - dg["ksp_its"]) <= 1, (s, do["ksp_its"], dg["ksp_its"])
        errs = {f: relerr(rows(gpu.field(f), P), rows(orc.field(f), P))
                for f in ("u2", "v2", "w2", "phi2", "convprev_phi", "visc2")}
        errs["pres"] = relerr(rows(demean(gpu.field("pres"), p.maxl, P, p.maxn), P),
                              rows(demean(orc.field("pres"), p.maxl, P, p.maxn), P))
        res = abs(dg["err_phi"] - do["err_phi"]) / max(do["err_phi"], 1e-300)
        print(f"[dim step {precision}] step {s} its {do['ksp_its']}/{dg['ksp_its']} res={res:.1e} "
              + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
        for k, v in errs.items():
            assert v < tol(s, k), (s, k, v)
        # the last sweep's largest change: a small difference of iterates
        if precision == "fp64":
            assert res < 1e-6, (s, res)
    gpu.close()
    orc.close()


def test_dim_one_step_matches_the_oracle():
    p = CaseParams(nx=32, ny=16, nz=24, Re=100.0, mu_plus=2.0, dt=1e-4, seed=5, idm="dim")
    step_pair(p, 1, "fp64", lambda s, f: 1e-12)


def test_dim_ten_steps_match_the_oracle():
    p = CaseParams(nx=48, ny=16, nz=24, Re=200.0, mu_plus=5.0, dt=1e-4, noise=1e-2, seed=6, idm="dim",
                   dim_its=12)
    step_pair(p, 10, "fp64", lambda s, f: 1e-12 if s == 1 else 1e-9)


def test_dim_fp32_storage_tracks_the_oracle():
    """FP32 storage: the phase field and viscosity to FP32 rounding, the flow
    to the looser bound its sensitivity allows.  The phase field starts in
    the range the step keeps it in (phi3 = phiNew - 0.5 with phiNew in
    [0, 1]): from the generator's 0..1 profile the first step moves the
    interface by half a phase unit and the kick amplifies rounding in the
    velocities (measured 3e-4 after two steps)."""
    p = CaseParams(nx=32, ny=16, nz=24, Re=100.0, mu_plus=2.0, dt=1e-4, seed=7, idm="dim")
    step_pair(p, 3, "fp32", lambda s, f: 1e-5 if f in ("phi2", "visc2") else 1e-2, shift_phi=0.5)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dim_graph_replay_equals_eager(precision):
    from tpls_amd.case import Case
    p = CaseParams(nx=64, ny=16, nz=32, Re=200.0, mu_plus=4.0, dt=2e-4, noise=1e-2, seed=12, ksp_pc="mg",
                   idm="dim")
    init = make_initial(p)
    with Case(p, precision=precision) as eager, Case(p, precision=precision, graphs=True) as graph:
        eager.load(*init)
        graph.load(*init)
        for s in range(1, 9):
            eager.step(1)
            graph.step(1)
            de, dg = eager.diag(), graph.diag()
            assert de["ksp_its"] == dg["ksp_its"] and de["err_phi"] == dg["err_phi"], s
            for f in ("u2", "v2", "w2", "phi2", "pres", "convprev_phi"):
                assert np.array_equal(eager.field(f), graph.field(f)), (s, f)
    print(f"[dim graphs] {precision}: 8 steps identical", flush=True)


@pytest.mark.parametrize("precision,shape", [("fp64", (64, 16, 32)), ("fp32", (72, 12, 40)),
                                             ("fp64", (40, 6, 23))])
def test_dim_fused_sweep_equals_two_launch_sweep(precision, shape):
    """One out-of-place pass per sweep against the Laplacian + in-place update
    (partial tiles in x, y and k in the last two grids; an odd sweep count
    ends in the second buffer)."""
    from tpls_amd.case import Case
    nx, ny, nz = shape
    p = CaseParams(nx=nx, ny=ny, nz=nz, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=13, idm="dim", dim_its=7)
    init = make_initial(p)
    with Case(p, precision=precision) as f, Case(p, precision=precision, dim_split=True) as sp:
        f.load(*init)
        sp.load(*init)
        for s in range(1, 4):
            f.step(1)
            sp.step(1)
            assert f.diag()["err_phi"] == sp.diag()["err_phi"], s
            for fld in ("phi2", "convprev_phi", "u2", "pres"):
                assert np.array_equal(f.field(fld), sp.field(fld)), (s, fld)
    print(f"[dim fused] {precision} {nx}x{ny}x{nz}: 3 steps identical", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dim_sweep_chunk_length_changes_no_value(precision):
    """The fused sweep marches each block up a chunk of levels with a
    one-level register prefetch and an LDS ring; chunks of 1, 2, 8, 17 levels
    and the whole column give the two-launch sweeps' fields bit for bit."""
    from tpls_amd.case import Case
    p = CaseParams(nx=64, ny=16, nz=40, Re=150.0, mu_plus=3.0, dt=2e-4, noise=1e-2, seed=14, idm="dim", dim_its=5)
    init = make_initial(p)
    with Case(p, precision=precision, dim_split=True) as sp:
        sp.load(*init)
        sp.step(2)
        want = {f: sp.field(f) for f in ("phi2", "convprev_phi", "u2", "pres")}
        werr = sp.diag()["err_phi"]
    for ks in (1, 2, 8, 17, p.nz):
        with Case(p, precision=precision, march_levels=ks) as f:
            f.load(*init)
            f.step(2)
            assert f.diag()["err_phi"] == werr, ks
            for fld, w in want.items():
                assert np.array_equal(f.field(fld), w), (ks, fld)
    print(f"[dim chunk] {precision}: chunks of 1/2/8/17/{p.nz} levels identical", flush=True)


def test_dim_volume_uses_the_shifted_phase():
    from tpls_amd.case import Case
    p = CaseParams(nx=32, ny=16, nz=24, seed=4, idm="dim")
    with Case(p) as c:
        c.load(*make_initial(p))
        c.step(2)
        phi = c.field("phi2")[1:p.maxl, 1:p.maxm, 1:p.maxn] - 0.5
        v_gpu = c.volume()
    sw = p.smooth_width
    pi = float(np.float32(3.1415926535))
    H = np.where(phi < -sw, 0.0, np.where(phi > sw, 1.0, 0.5 + phi / (2 * sw) + np.sin(pi * phi / sw) / (2 * pi)))
    assert v_gpu == pytest.approx(H.sum() * p.dx ** 3, rel=1e-12)


def test_dim_restored_case_steps_bit_identically(tmp_path):
    from tpls_amd.case import Case
    p = CaseParams(nx=40, ny=12, nz=20, Re=60.0, mu_plus=3.0, dt=1e-4, seed=9, idm="dim", dim_its=10)
    ck = tmp_path / "dim.ckpt"
    with Case(p) as a:
        a.load(*make_initial(p))
        a.step(3)
        a.checkpoint(ck)
        a.step(3)
        want = {f: a.field(f) for f in ("u2", "v2", "w2", "phi2", "pres", "convprev_phi")}
    with Case(p) as b:
        b.restore(ck)
        b.step(3)
        for f, w in want.items():
            assert np.array_equal(b.field(f).view(np.int64), w.view(np.int64)), f
    print("[dim restore] 3+3 steps identical", flush=True)
