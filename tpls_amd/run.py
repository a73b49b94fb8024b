"""A TPLS run on one GPU: the program flow of main_ns_hybrid.F90 around the
device-resident timestep.

    python -m tpls_amd.run [-d CONFIG_DIR] [-f OPTIONS_FILE] [--precision fp64|fp32]
                           [--ksp-pc none|mg] [--graphs] [--tiles auto]
                           [--out DIR] [--checkpoint FILE --checkpoint-every N] [--restore FILE]

-d and -f are the reference's command-line options (option_names.F90:30-32;
defaults '.' and 'tpls_config.opt', main_ns_hybrid.F90:58-66).  The run reads
the configuration and the initial_*.dat files (main_ns_hybrid.F90:242-410),
writes AA_initial_phi.dat (main_ns_hybrid.F90:395-398), then advances
num_timesteps steps.  With backup_text_format, every phi_dat_frequency-th
iteration writes phi_channel_<iteration>.dat, every uvw_dat_frequency-th of
those uvw_channel_<iteration>.dat, and a text backup alternately to
fieldbackup1.dat and fieldbackup0.dat (main_ns_hybrid.F90:975-2478), with the
reference's file names (the iteration in an I9 field).  Their content is what
the reference writes at that point of an iteration: the velocities, phi and
convection histories of the iteration's start, the pressure of its end, each
array as the reference's gather leaves it (interior copied, then the ghost
rules of main_ns_hybrid.F90:1070-2430).  The NetCDF outputs (backup_hdf5_format)
are not written.

--checkpoint writes a full-state snapshot (tpls_case_checkpoint) every N
iterations and at the end; --restore resumes from one, continuing the
iteration count (no initial files are read then).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import io
from .case import Case


def gathered(name: str, a: np.ndarray, maxl: int, maxm: int, maxn: int, dz: float = 0.0, height: float = 0.0,
             u_inlet: np.ndarray | None = None) -> np.ndarray:
    """An array as main_ns_hybrid.F90's gather to the master process leaves
    it: the interior i = 1..maxl-1, j = 1..maxm-1 copied into a zeroed array,
    then that array's ghost-cell rules."""
    g = np.zeros_like(a, order="F")
    g[1:maxl, 1:maxm, :] = a[1:maxl, 1:maxm, :]

    def wrap_y():
        g[:, 0, :] = g[:, maxm - 1, :]
        g[:, maxm, :] = g[:, 1, :]

    if name in ("pres", "pres_old", "conv0_phi", "conv1_phi", "visc"):  # main:1070-1080 and alike
        g[0] = g[1]
        g[maxl] = g[maxl - 1]
        wrap_y()
        g[:, :, maxn] = g[:, :, maxn - 1]
        g[:, :, 0] = g[:, :, 1]
    elif name == "phi":  # main:1264-1281
        k = np.arange(1, maxn)
        g[0, 1:maxm, 1:maxn] = (k * dz - 0.5 * dz - height)[None, :]
        g[maxl] = g[maxl - 1]
        wrap_y()
        g[:, :, maxn] = g[:, :, maxn - 1]
        g[:, :, 0] = g[:, :, 1]
    elif name in ("conv0_u", "conv1_u", "conv0_v", "conv1_v"):  # main:1757-1765 and alike
        g[0] = g[1]
        g[maxl] = g[maxl - 1]
        wrap_y()
    elif name in ("conv0_w", "conv1_w"):  # main:2324-2334, 2420-2430
        g[0] = g[1]
        g[maxl] = g[maxl - 1]
        wrap_y()
        g[:, :, 0] = 0.0
        g[:, :, maxn - 1] = 0.0
    elif name == "u":  # main:1659-1670
        g[0, :, :] = u_inlet[None, :]
        g[maxl - 1] = g[maxl - 2]
        g[maxl] = g[maxl - 2]
        wrap_y()
    elif name == "v":  # main:1944-1951
        g[0] = 0.0
        g[maxl] = g[maxl - 1]
        wrap_y()
    elif name == "w":  # main:2227-2238
        g[:, :, 0] = 0.0
        g[:, :, maxn - 1] = 0.0
        g[0] = 0.0
        g[maxl] = 0.0
        wrap_y()
    else:
        raise ValueError(name)
    return g


def iteration_file(prefix: str, iteration: int) -> str:
    """write(filename,'(A,I9,A)') prefix, iteration, '.dat'"""
    return f"{prefix}{iteration:9d}.dat"


class Outputs:
    """The text outputs of one iteration, captured around its step."""

    PRE = ("u2", "v2", "w2", "phi2", "conv0_u", "conv1_u", "conv0_v", "conv1_v", "conv0_w", "conv1_w",
           "conv0_phi", "conv1_phi", "pres_old")

    def __init__(self, case: Case, prm, out_dir: str):
        self.c = case
        self.prm = prm
        self.out = out_dir
        self.maxl, self.maxm, self.maxn = case.shape
        self.backups = 0
        self.u_inlet = case.field("u_inlet")

    def before(self) -> dict:
        return {f: self.c.field(f) for f in self.PRE}

    def after(self, iteration: int, pre: dict, uvw: bool) -> None:
        l, m, n = self.maxl, self.maxm, self.maxn
        p = self.prm
        G = lambda name, a: gathered(name, a, l, m, n, p.dz, p.height, self.u_inlet)  # noqa: E731
        phi = G("phi", pre["phi2"])
        pres = G("pres", self.c.field("pres"))
        path = lambda f: os.path.join(self.out, f)  # noqa: E731
        io.channel_output_phi(phi, l, m, n, p.dx, p.dy, p.dz, path(iteration_file("phi_channel_", iteration)))
        u, v, w = G("u", pre["u2"]), G("v", pre["v2"]), G("w", pre["w2"])
        if uvw:
            io.channel_output_uvw(u, v, w, pres, l, m, n, p.dx, p.dy, p.dz,
                                  path(iteration_file("uvw_channel_", iteration)))
        self.backups += 1
        arrays = {"u2": u, "v2": v, "w2": w, "phi": phi, "pres": pres, "pres_old": G("pres_old", pre["pres_old"])}
        for f in ("conv0_u", "conv1_u", "conv0_v", "conv1_v", "conv0_w", "conv1_w", "conv0_phi", "conv1_phi"):
            arrays[f] = G(f, pre[f])
        io.backup_channel(arrays, l, m, n, path(f"fieldbackup{self.backups % 2}.dat"))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-d", dest="config_dir", default=".")
    ap.add_argument("-f", dest="options_file", default="tpls_config.opt")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--ksp-pc", default="none", choices=("none", "mg"))
    ap.add_argument("--ksp-rtol", type=float, default=1e-10)
    ap.add_argument("--graphs", action="store_true")
    ap.add_argument("--tiles", default=None, help="'auto' for the tuned launch shapes")
    ap.add_argument("--out", default=".", help="directory of the text outputs")
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--checkpoint-every", type=int, default=0)
    ap.add_argument("--restore", default=None)
    a = ap.parse_args(argv)

    t0 = time.perf_counter()
    print(f"Loading options from options file: {a.options_file}", flush=True)
    case = Case.from_config(a.config_dir, a.options_file, precision=a.precision, graphs=a.graphs, tiles=a.tiles,
                            ksp_pc=a.ksp_pc, ksp_rtol=a.ksp_rtol)
    run, prm = case.run, case._prm
    os.makedirs(a.out, exist_ok=True)
    first = 1
    if a.restore:
        case.restore(a.restore)
        first = case.steps + 1
        print(f"Resumed from {a.restore} after iteration {case.steps}", flush=True)
    else:
        # the initial level set as its file holds it (main_ns_hybrid.F90:390-398)
        print(" Outputing initial phi", flush=True)
        l, m, n = case.shape
        phi0 = np.zeros((l + 1, m + 1, n + 1), order="F")
        io.load_data_into_array(os.path.join(a.config_dir, io.INITIAL_FILES["phi"]), phi0, 0, l, 0, m, 0, n)
        io.channel_output_phi(phi0, l, m, n, prm.dx, prm.dy, prm.dz, os.path.join(a.out, "AA_initial_phi.dat"))
    print(f" dt= {prm.dt!r}\n cfl= {run.maxu * prm.dt / prm.dx!r}\n n_timesteps= {run.n_timesteps}\n"
          f" (dx,dy,dz)= {prm.dx!r} {prm.dy!r} {prm.dz!r}", flush=True)
    outs = Outputs(case, prm, a.out) if run.backup_text_format else None
    t1 = time.perf_counter()
    for it in range(first, run.n_timesteps + 1):
        due = outs is not None and it % run.phi_dat_frequency == 0
        pre = outs.before() if due else None
        case.step(1)
        d = case.diag()
        print(f"iteration {it}: err_phi {d['err_phi']:.6e} err_u {d['err_u']:.6e} err_v {d['err_v']:.6e} "
              f"err_w {d['err_w']:.6e} ksp_its {d['ksp_its']} ksp_reason {d['ksp_reason']}", flush=True)
        if due:
            outs.after(it, pre, it % run.uvw_dat_frequency == 0)
        if a.checkpoint and a.checkpoint_every and it % a.checkpoint_every == 0:
            case.checkpoint(a.checkpoint)
    case.sync()
    t2 = time.perf_counter()
    if a.checkpoint:
        case.checkpoint(a.checkpoint)
    steps = max(0, run.n_timesteps + 1 - first)
    print(f" Timing: setup: {t1 - t0:.3f}\n Timing: iteration_loop: {t2 - t1:.3f}\n"
          f" Timing: per_iteration: {(t2 - t1) / max(steps, 1):.6f}", flush=True)
    case.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
