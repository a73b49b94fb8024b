"""A device-resident TPLS case: all state arrays of main_ns_hybrid.F90 live in
HBM and ``step()`` advances whole timesteps on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import CaseParams


def field_shape(name: str, maxl: int, maxm: int, maxn: int):
    """Fortran extents of a case field (main_ns_hybrid.F90:416-549)."""
    if name == "u_inlet":
        return (maxn - 1,)
    if name == "visc2_aug":
        return (maxl + 3, maxm + 3, maxn + 1)
    if name in ("w2", "w3") or name.endswith("_w") or name.startswith("csf_w"):
        return (maxl + 1, maxm + 1, maxn)
    if name in ("phi2", "phi3", "pres", "pres_old", "conv0_phi", "conv1_phi", "conv2_phi",
                "fx_csf", "fy_csf", "fz_csf", "visc2", "rhs_p", "convprev_phi"):
        return (maxl + 1, maxm + 1, maxn + 1)
    return (maxl + 1, maxm + 1, maxn - 1)


def fused_code(sor_fused) -> int:
    """tpls_params.sor_fused from None (the default), a bool (all or no
    SOR iterations fused) or the code itself (-1 .. 7; bit 2 fuses an even
    iteration's Jacobi step with its first colour sweep)."""
    if sor_fused is None:
        return 0
    if isinstance(sor_fused, bool):
        return 3 if sor_fused else -1
    return int(sor_fused)


def make_params(p: CaseParams, precision: str = "fp64", diagnostics: bool = True,
                graphs: bool = False, tiles=None, sor_fused=None, mg_fp32=None,
                dim_split: bool = False, march_levels: int = 0, mg_pair: int = 0) -> _lib.Params:
    prm = _lib.Params()
    prm.march_levels = int(march_levels)
    prm.mg_pair = int(mg_pair)
    prm.mg_fp32 = {None: 0, True: 1, False: -1}[mg_fp32]
    prm.dim_split = 1 if dim_split else 0
    for k, v in p.solver_dict().items():
        setattr(prm, k, v)
    prm.precision = {"fp64": _lib.FP64, "fp32": _lib.FP32}[precision]
    prm.diagnostics = 1 if diagnostics else 0
    prm.graphs = 1 if graphs else 0
    prm.sor_fused = fused_code(sor_fused)
    if tiles:
        prm.tile_rows, prm.tile_rows_odd, prm.prefetch_odd = (int(v) for v in tiles[:3])
        if len(tiles) > 3:
            prm.tile_rows_rb, prm.tile_rows_rb_odd = (int(v) for v in tiles[3:5])
        if len(tiles) > 5:
            prm.tile_rows_jst = int(tiles[5])
        if len(tiles) > 6 and sor_fused is None:
            prm.sor_fused = int(tiles[6])
    return prm


# The tpls_params fields a TPLS configuration sets (include/tpls_io.h).
FILE_FIELDS = ("dx", "dy", "dz", "dt", "Re", "mu_plus", "mu_minus", "height", "smooth_width", "scap",
               "mom_u", "mom_v", "mom_w", "levelset_its", "idm_dim", "epn", "Pe", "dim_its")


class Case:
    def __init__(self, params: CaseParams, precision: str = "fp64", diagnostics: bool = True,
                 graphs: bool = False, tiles=None, sor_fused=None, file_params: _lib.Params | None = None,
                 mg_fp32=None, dim_split: bool = False, march_levels: int = 0, mg_pair: int = 0):
        """graphs: replay each step's launches from HIP graphs (same results).
        tiles: (rows, rows_odd, prefetch_odd[, rows_rb, rows_rb_odd[, rows_jst[, sor_fused]]]) launch
        shapes of the Helmholtz sweeps, "auto" for the tuned shapes of this
        grid (tpls_amd.autotune), or None for the defaults.  sor_fused: which
        SOR iterations run as one out-of-place pass instead of two in-place
        colour sweeps (None: the default, odd iteration_time; True / False:
        all / none; or the library's code).  Every choice computes the same
        values.  mg_fp32: precision of the multigrid V-cycle (None: the
        case's precision; True: FP32; False: FP64); the CG around it stays
        FP64.  dim_split: with idm dim, run each relaxation sweep as two
        launches instead of one fused pass (same values).  march_levels: levels
        per block of the k-marching sweeps (0: sized to the device; same
        values).  mg_pair: the multigrid's finest-level smoothing pairs as one
        pass each (0: on 2 M cells or more, 1: always, -1: never; same
        values).  file_params: the tpls_params of a configuration
        (tpls_amd.io.read_configuration), whose file-backed fields replace
        those params implies."""
        if tiles == "auto":
            from .autotune import tuned
            tiles = tuned(params, precision, fused_code(sor_fused))
        params.check()
        self.params = params
        self.precision = precision
        self.shape = (params.maxl, params.maxm, params.maxn)
        self.tiles = tiles
        self._prm = make_params(params, precision, diagnostics, graphs, tiles, sor_fused, mg_fp32, dim_split,
                                march_levels, mg_pair)
        if file_params is not None:
            for k in FILE_FIELDS:
                setattr(self._prm, k, getattr(file_params, k))
        h = C.c_void_p()
        L = _lib.lib()
        _lib.check(L.tpls_case_create(C.byref(self._prm), *self.shape, C.byref(h)), "tpls_case_create")
        self._h = h

    @classmethod
    def from_config(cls, config_dir, options_file, precision: str = "fp64", diagnostics: bool = True,
                    graphs: bool = False, tiles=None, sor_fused=None, **solver):
        """A case configured and initialised as main_ns_hybrid.F90:242-410
        does it: the options file and config_dir/initial_config.dat give the
        parameters (exactly as written there), the initial_*.dat files of
        config_dir the state.  solver: the CaseParams fields the files do not
        hold (ksp_pc, ksp_rtol, ...).  The run configuration is kept as
        ``case.run``."""
        from . import io
        prm, run = io.read_configuration(config_dir, options_file)
        dim = dict(idm="dim", epn=run.epn, Pe=run.Pe, dim_its=run.max_iteration_dim) if run.idm_dim else {}
        p = CaseParams(nx=run.maxl - 1, ny=run.maxm - 1, nz=run.maxn - 1, Re=prm.Re, mu_minus=prm.mu_minus,
                       mu_plus=prm.mu_plus, height=prm.height, scap=prm.scap, dt=prm.dt,
                       smooth_width_scale=prm.smooth_width * (run.maxn - 1), mom_u=prm.mom_u,
                       mom_v=prm.mom_v, mom_w=prm.mom_w, levelset_its=prm.levelset_its, **dim, **solver)
        case = cls(p, precision, diagnostics, graphs, tiles, sor_fused, file_params=prm)
        case.run = run
        case.load_config(config_dir)
        return case

    def load_config(self, config_dir) -> None:
        """Initial state from config_dir/initial_*.dat (main_ns_hybrid.F90:380-393)."""
        from . import io
        ierr = C.c_int(0)
        rc = io._io().tpls_case_load_config(self._h, io._b(config_dir), C.byref(ierr))
        if rc == -8:
            raise io.ConfigError(ierr.value, _lib.lib().tpls_last_error().decode(errors="replace"))
        _lib.check(rc, "tpls_case_load_config")

    def save_field(self, name: str, filename, description: str) -> None:
        """save_data of a field (io.F90:47-73) straight from the device."""
        from . import io
        _lib.check(io._io().tpls_case_save_field(self._h, _lib.FIELD_ID[name], io._b(filename),
                                                 description.encode()), f"save {name}")

    def checkpoint(self, filename) -> None:
        """Full-state snapshot (tpls_case_checkpoint): a case restored from
        it steps bit-identically to this one."""
        from . import io
        _lib.check(io._io().tpls_case_checkpoint(self._h, io._b(filename)), "checkpoint")

    def restore(self, filename) -> None:
        from . import io
        _lib.check(io._io().tpls_case_restore(self._h, io._b(filename)), "restore")

    @property
    def steps(self) -> int:
        from . import io
        n = C.c_long(0)
        _lib.check(io._io().tpls_case_steps(self._h, C.byref(n)), "steps")
        return n.value

    @staticmethod
    def device_bytes(params: CaseParams, precision: str = "fp64") -> int:
        return int(_lib.lib().tpls_case_bytes(params.maxl, params.maxm, params.maxn,
                                              {"fp64": 0, "fp32": 1}[precision]))

    def load(self, u, v, w, pres, phi) -> None:
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (u, v, w, pres, phi)]
        _lib.check(_lib.lib().tpls_case_load(self._h, *[a.ctypes.data_as(_lib._D) for a in arrs]),
                   "tpls_case_load")

    def field(self, name: str) -> np.ndarray:
        out = np.zeros(field_shape(name, *self.shape), order="F")
        _lib.check(_lib.lib().tpls_case_get_field(self._h, _lib.FIELD_ID[name],
                                                  out.ctypes.data_as(_lib._D)), f"get {name}")
        return out

    def set_field(self, name: str, arr) -> None:
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != field_shape(name, *self.shape):
            raise ValueError(f"{name}: shape {a.shape} != {field_shape(name, *self.shape)}")
        _lib.check(_lib.lib().tpls_case_set_field(self._h, _lib.FIELD_ID[name],
                                                  a.ctypes.data_as(_lib._D)), f"set {name}")

    def step(self, n: int = 1) -> None:
        _lib.check(_lib.lib().tpls_step(self._h, int(n)), "tpls_step")

    def diag(self) -> dict:
        d = _lib.Diag()
        _lib.check(_lib.lib().tpls_case_diag(self._h, C.byref(d)), "tpls_case_diag")
        return {f: getattr(d, f) for f, _ in _lib.Diag._fields_}

    def sync(self) -> None:
        _lib.check(_lib.lib().tpls_case_sync(self._h), "tpls_case_sync")

    TIMED = {None: 0, False: 0, "cg_x": 1, True: 1, "jacobi": 2, "sor_odd": 3, "sor_even": 4, "advect": 5}

    def time_kernels(self, kernel=True) -> None:
        """Bracket every launch of one kernel class with HIP events on the
        case's stream: "cg_x", "jacobi", "sor_odd", "sor_even" (True = "cg_x";
        None or False stops timing)."""
        _lib.check(_lib.lib().tpls_case_time_kernels(self._h, self.TIMED[kernel]), "time_kernels")

    def kernel_time(self):
        ms, n = C.c_double(0), C.c_long(0)
        _lib.check(_lib.lib().tpls_case_kernel_time(self._h, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value

    def set_tiling(self, rows: int, rows_odd: int, prefetch_odd: int, rows_rb: int = 8, rows_rb_odd: int = 8,
                   rows_jst: int = 4, sor_fused: int = 0) -> None:
        _lib.check(_lib.lib().tpls_case_set_tiling(self._h, rows, rows_odd, prefetch_odd, rows_rb, rows_rb_odd,
                                                   rows_jst, sor_fused), "set_tiling")
        self.tiles = (rows, rows_odd, prefetch_odd, rows_rb, rows_rb_odd, rows_jst, sor_fused)

    def volume(self) -> float:
        """Liquid volume: sum of H_eps(phi2) dx dy dz over the interior, FP64."""
        v = C.c_double(0)
        _lib.check(_lib.lib().tpls_case_volume(self._h, C.byref(v)), "volume")
        return v.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().tpls_case_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_step(cases, nsteps: int = 1) -> None:
    """Advance several cases together (tpls_batch_step): each case runs on its
    own stream and the GPU overlaps them; every case computes what
    Case.step would compute for it alone."""
    arr = (C.c_void_p * len(cases))(*[c._h for c in cases])
    _lib.check(_lib.lib().tpls_batch_step(arr, len(cases), int(nsteps)), "tpls_batch_step")


def set_batch_grouped(mode) -> None:
    """How batch_step runs a batch (tpls_batch_set_grouped): 0 (or False)
    each case on its own stream; 1 (or True, the default) the pressure
    solves grouped, one launch per kernel across the cases; 2 the Helmholtz
    loops and the solve's set-up grouped as well."""
    _lib.check(_lib.lib().tpls_batch_set_grouped(int(mode)), "tpls_batch_set_grouped")


def batch_step_status(cases, nsteps: int = 1) -> list:
    """batch_step that isolates failures (tpls_batch_step_status): returns one
    status per case, 0 or the error code of that case's step (its message
    then in the second element of the pair); a failed case stops, the others
    complete their steps."""
    n = len(cases)
    arr = (C.c_void_p * n)(*[c._h for c in cases])
    st = (C.c_int * n)()
    rc = _lib.lib().tpls_batch_step_status(arr, n, int(nsteps), st)
    msg = _lib.lib().tpls_last_error().decode(errors="replace") if rc else ""
    return [(int(v), msg if v else "") for v in st]
