"""A case read from a TPLS configuration directory on the GPU
(tpls_case_load_config, main_ns_hybrid.F90:242-410) and fields saved from
the device in the data-file format (tpls_case_save_field).

A case configured from files must step bit-identically to the same case
built from the arrays and parameters directly (grid spacings that F20.10
represents exactly, so both hold the same parameters); a field saved from the
device must read back bit-identically to the field copied from it.
"""
import os

import numpy as np
import pytest

from tpls_amd import io
from tpls_amd.config import CaseParams
from tpls_amd.initial import make_initial

pytestmark = pytest.mark.gpu

FIELDS = ("u2", "v2", "w2", "phi2", "pres")


def write_case(d, p):
    init = make_initial(p)
    io.write_initial_conditions(d, p, init)
    io.write_options_file(os.path.join(d, "tpls_config.opt"), p, n_timesteps=3)
    return init


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_case_from_config_steps_like_case_from_arrays(tmp_path, precision):
    from tpls_amd.case import Case
    # nz = 32: dx = 1/32 and smooth_width = 1.5/32 are exact in F20.10
    p = CaseParams(nx=48, ny=16, nz=32, Re=80.0, mu_plus=2.0, dt=1e-4, seed=6)
    init = write_case(str(tmp_path), p)
    print(f"case from config, 48x16x32 {precision}", flush=True)
    with Case.from_config(tmp_path, tmp_path / "tpls_config.opt", precision=precision) as a, \
            Case(p, precision=precision) as b:
        assert a.run.n_timesteps == 3
        b.load(*init)
        for f in FIELDS:
            assert np.array_equal(a.field(f), b.field(f)), f
        a.step(a.run.n_timesteps)
        b.step(3)
        for f in FIELDS:
            assert np.array_equal(a.field(f), b.field(f)), f


def test_dim_case_from_config_steps_like_case_from_arrays(tmp_path):
    from tpls_amd.case import Case
    p = CaseParams(nx=32, ny=8, nz=16, Re=80.0, mu_plus=2.0, dt=1e-4, seed=7, idm="dim", dim_its=9)
    init = write_case(str(tmp_path), p)
    print("dim case from config, 32x8x16", flush=True)
    with Case.from_config(tmp_path, tmp_path / "tpls_config.opt") as a, Case(p) as b:
        assert a.params.idm == "dim" and a.params.dim_its == 9
        b.load(*init)
        a.step(3)
        b.step(3)
        for f in FIELDS + ("convprev_phi",):
            assert np.array_equal(a.field(f), b.field(f)), f


def test_save_field_reads_back(tmp_path):
    from tpls_amd.case import Case
    p = CaseParams(nx=32, ny=8, nz=16, seed=3)
    print("save_field, 32x8x16", flush=True)
    with Case(p) as c:
        c.load(*make_initial(p))
        c.step(1)
        for name, desc in (("u2", "velocity (u)"), ("w2", "velocity (w)"), ("pres", "pressure"),
                           ("phi2", "level-set (phi)")):
            f = tmp_path / f"{name}.dat"
            c.save_field(name, f, desc)
            data, lx, ly, mx, my, nx, ny, ierr = io.load_data(f)
            want = c.field(name)
            assert ierr == 0 and (lx, mx, nx) == (0, 0, 0)
            assert data.shape == want.shape
            assert np.array_equal(data.view(np.int64), want.view(np.int64)), name


def test_load_config_errors(tmp_path):
    from tpls_amd.case import Case
    p = CaseParams(nx=16, ny=8, nz=16, seed=1)
    write_case(str(tmp_path), p)
    print("load_config errors, 16x8x16", flush=True)
    with Case(p) as c:
        c.load_config(tmp_path)
        q = CaseParams(nx=16, ny=8, nz=12, seed=1)
        other = tmp_path / "other"
        write_case(str(other), q)
        os.replace(other / "initial_w.dat", tmp_path / "initial_w.dat")
        with pytest.raises(io.ConfigError) as e:
            c.load_config(tmp_path)
        assert e.value.ierr == 3 and "initial_w.dat data dimensions do not match" in str(e.value)
        write_case(str(tmp_path), p)
        c.load_config(tmp_path)
        os.remove(tmp_path / "initial_viscosity.dat")
        with pytest.raises(io.ConfigError) as e:
            c.load_config(tmp_path)
        assert e.value.ierr == 200


@pytest.mark.parametrize("precision,pc,graphs", [("fp64", "none", False), ("fp32", "mg", True)])
def test_restored_case_steps_bit_identically(tmp_path, precision, pc, graphs):
    """Run 3 steps, snapshot, run 4 more; a fresh case restored from the
    snapshot and run 4 steps ends with the same bits in every field."""
    from tpls_amd.case import Case
    p = CaseParams(nx=40, ny=12, nz=20, Re=60.0, mu_plus=3.0, dt=1e-4, seed=9, ksp_pc=pc)
    print(f"checkpoint/restore 40x12x20 {precision} {pc} graphs={graphs}", flush=True)
    ck = tmp_path / "case.ckpt"
    with Case(p, precision=precision, graphs=graphs) as a:
        a.load(*make_initial(p))
        a.step(3)
        a.checkpoint(ck)
        a.step(4)
        want = {f: a.field(f) for f in FIELDS + ("conv0_u", "csf_w1", "conv1_phi", "pres_old")}
        dw = a.diag()
        assert a.steps == 7
    with Case(p, precision=precision, graphs=graphs) as b:
        b.restore(ck)
        assert b.steps == 3
        b.step(4)
        assert b.steps == 7
        for f, w in want.items():
            assert np.array_equal(b.field(f).view(np.int64), w.view(np.int64)), f
        db = b.diag()
        assert db["ksp_its"] == dw["ksp_its"] and db["err_u"] == dw["err_u"]


def test_restore_rejects_other_cases_and_corruption(tmp_path):
    from tpls_amd import _lib
    from tpls_amd.case import Case
    p = CaseParams(nx=16, ny=8, nz=16, seed=2)
    print("restore errors, 16x8x16", flush=True)
    ck = tmp_path / "a.ckpt"
    with Case(p) as a:
        a.load(*make_initial(p))
        a.step(1)
        a.checkpoint(ck)
        with Case(CaseParams(nx=16, ny=8, nz=12, seed=2)) as b:
            with pytest.raises(_lib.TplsError, match="checkpoint of a 17x9x17 FP64 case"):
                b.restore(ck)
        with Case(p, precision="fp32") as b:
            with pytest.raises(_lib.TplsError, match="FP64 case"):
                b.restore(ck)
        raw = bytearray(ck.read_bytes())
        raw[len(raw) // 2] ^= 0x40
        bad = tmp_path / "bad.ckpt"
        bad.write_bytes(bytes(raw))
        with pytest.raises(_lib.TplsError, match="checksum mismatch"):
            a.restore(bad)
        bad.write_bytes(bytes(raw[: len(raw) // 3]))
        with pytest.raises(_lib.TplsError, match="truncated"):
            a.restore(bad)
        with pytest.raises(_lib.TplsError, match="not a case checkpoint"):
            a.restore(__file__)


def test_run_driver_outputs_and_resume(tmp_path):
    """python -m tpls_amd.run on a configuration directory: the reference's
    output files at its cadence, their content taken at the reference's point
    of the iteration, and a resumed run ending in the same state."""
    from tpls_amd import run
    from tpls_amd.case import Case
    p = CaseParams(nx=24, ny=8, nz=16, seed=5)
    cfg, out = tmp_path / "cfg", tmp_path / "out"
    io.write_initial_conditions(cfg, p, make_initial(p))
    opts = cfg / "tpls_config.opt"
    io.write_options_file(opts, p, n_timesteps=4, phi_dat_frequency=2, uvw_dat_frequency=4)
    print("run driver, 24x8x16, 4 iterations", flush=True)
    assert run.main(["-d", str(cfg), "-f", str(opts), "--out", str(out), "--checkpoint", str(tmp_path / "a.ckpt")]) == 0
    assert sorted(os.listdir(out)) == sorted(["AA_initial_phi.dat", "phi_channel_        2.dat",
                                              "phi_channel_        4.dat", "uvw_channel_        4.dat",
                                              "fieldbackup1.dat", "fieldbackup0.dat"])
    with Case.from_config(cfg, opts) as c:
        l, m, n = c.shape
        prm = c._prm
        c.step(3)
        pre = {f: c.field(f) for f in ("phi2", "u2", "v2", "w2")}
        inlet = c.field("u_inlet")
        c.step(1)
        G = lambda name, a: run.gathered(name, a, l, m, n, prm.dz, prm.height, inlet)  # noqa: E731
        io.channel_output_phi(G("phi", pre["phi2"]), l, m, n, prm.dx, prm.dy, prm.dz, tmp_path / "phi.dat")
        io.channel_output_uvw(G("u", pre["u2"]), G("v", pre["v2"]), G("w", pre["w2"]), G("pres", c.field("pres")),
                              l, m, n, prm.dx, prm.dy, prm.dz, tmp_path / "uvw.dat")
        assert (tmp_path / "phi.dat").read_bytes() == (out / "phi_channel_        4.dat").read_bytes()
        assert (tmp_path / "uvw.dat").read_bytes() == (out / "uvw_channel_        4.dat").read_bytes()
        c.checkpoint(tmp_path / "direct.ckpt")
    assert (tmp_path / "direct.ckpt").read_bytes() == (tmp_path / "a.ckpt").read_bytes()
    # the same four iterations as two runs of two
    io.write_options_file(opts, p, n_timesteps=2, phi_dat_frequency=2, uvw_dat_frequency=4)
    assert run.main(["-d", str(cfg), "-f", str(opts), "--out", str(tmp_path / "o1"),
                     "--checkpoint", str(tmp_path / "b2.ckpt")]) == 0
    io.write_options_file(opts, p, n_timesteps=4, phi_dat_frequency=2, uvw_dat_frequency=4)
    assert run.main(["-d", str(cfg), "-f", str(opts), "--out", str(tmp_path / "o2"), "--restore",
                     str(tmp_path / "b2.ckpt"), "--checkpoint", str(tmp_path / "b4.ckpt")]) == 0
    assert (tmp_path / "b4.ckpt").read_bytes() == (tmp_path / "a.ckpt").read_bytes()
    assert (tmp_path / "o2" / "phi_channel_        4.dat").read_bytes() == (out / "phi_channel_        4.dat").read_bytes()
