"""ctypes bindings for the CPU oracle and for the reference-compiled routines.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s CPU-baseline leg, always as the checker or the CPU
baseline, never as the product path.

Two libraries are wrapped:

* ``oracle/liboracle.so`` -- the C restatement (``tpls_oracle.c``); ints and
  doubles by value, arrays as pointers to Fortran-ordered numpy arrays.
* ``oracle/_ref/libtpls_ref.so`` -- the reference's own Fortran modules
  (weno, levelset, momentum_allflux, jacobi_iteration_allflux,
  sor_iteration_allflux, pressure) compiled by ``oracle/Makefile``.  Every
  Fortran argument is passed by reference.  Present only where the
  reference checkout was available at build time.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TPLS_ORACLE_SO names another build of tpls_oracle.c (tests/test_oracle_glue.py
# loads copies with planted mistakes through it)
ORACLE_SO = os.environ.get("TPLS_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtpls_ref.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

# Argument kinds: a = double array, I = int32 array, i = int, l = long,
# d = double, o = double output scalar.
SPECS = {
    "weno_appox_convection": ("aaaaiiiiiiaaa", "weno", "wenoappoxconvection"),
    "get_levelset_status": ("aaIaiiiiid", "levelset", "get_levelset_status"),
    "do_levelset_iteration": ("aaaIaiiiiiiddo", "levelset", "do_levelset_iteration"),
    "get_csf": ("aaaaiiiiiiddddd", "levelset", "get_csf"),
    "get_diffusion_all": ("aaaaaaaiiiiiiddd", "momentum_allflux", "get_diffusion_all"),
    "get_conv_all": ("aaaaaaaiiiiiiddd", "momentum_allflux", "get_conv_all"),
    "get_viscosity": ("aaiiiiidddd", "momentum_allflux", "get_viscosity"),
    "do_jacobi_u": ("aaaaddddiiiiii", "jacobi_iteration_allflux", "do_jacobi_u"),
    "do_jacobi_v": ("aaaaddddiiiiii", "jacobi_iteration_allflux", "do_jacobi_v"),
    "do_jacobi_w": ("aaaaddddiiiiii", "jacobi_iteration_allflux", "do_jacobi_w"),
    "get_difference": ("aaiiiiiio", "jacobi_iteration_allflux", "get_difference"),
    "do_sor_u": ("aaaddddiiiiiiii", "sor_iteration_allflux", "do_sor_u"),
    "do_sor_v": ("aaaddddiiiiiii", "sor_iteration_allflux", "do_sor_v"),
    "do_sor_w": ("aaaddddiiiiiii", "sor_iteration_allflux", "do_sor_w"),
    "get_source_pres": ("aaaaiiiiidddd", "pressure", "get_source_pres"),
    "get_uvw_pres": ("aaaaiiiiiidddda", "pressure", "get_uvw_pres"),
}
ORACLE_ONLY = {
    "do_advect_phi": "aaaaaiiiiiiiddd",
    "do_tpls": "aaaaaiiiiiiidoddd",
    "dim": "iiddddddaaaaaiiiidio",
    # main_ns_hybrid.F90's own statements of the step
    "csf_faces": "aaaaaaiii",
    "visc_aug": "aaaiiiddddd",
    "momentum_rhs": "aaaaaaaaald",
    "u_inout_bc": "aaiiidd",
    "v_bc": "aiii",
    "w_bc": "aiii",
}


def _ptr(arr, kind):
    if kind == "a":
        assert arr.dtype == np.float64 and (arr.flags.f_contiguous or arr.ndim == 1), \
            "double arrays must be Fortran-contiguous float64"
        return arr.ctypes.data_as(_D)
    assert arr.dtype == np.int32 and (arr.flags.f_contiguous or arr.ndim == 1)
    return arr.ctypes.data_as(_I)


class _Lib:
    def __init__(self, path, by_reference):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run: make -C oracle)")
        self.lib = C.CDLL(path)
        self.by_ref = by_reference

    def call(self, sym, spec, args, restype=None):
        fn = getattr(self.lib, sym)
        fn.restype = restype
        cargs, outs = [], []
        n_in = len(spec) - spec.count("o")
        assert n_in == len(args), (sym, n_in, len(args))
        it = iter(args)
        for kind in spec:
            a = None if kind == "o" else next(it)
            if kind in "aI":
                cargs.append(_ptr(a, kind))
            elif kind == "i":
                cargs.append(C.byref(C.c_int(int(a))) if self.by_ref else C.c_int(int(a)))
            elif kind == "l":
                cargs.append(C.c_long(int(a)))
            elif kind == "d":
                cargs.append(C.byref(C.c_double(float(a))) if self.by_ref else C.c_double(float(a)))
            elif kind == "o":
                box = C.c_double(0.0)
                outs.append(box)
                cargs.append(C.byref(box))
            else:
                raise ValueError(kind)
        ret = fn(*cargs)
        return ret, [b.value for b in outs]


_oracle = None
_ref = None


def oracle_lib() -> _Lib:
    global _oracle
    if _oracle is None:
        _oracle = _Lib(ORACLE_SO, by_reference=False)
        L = _oracle.lib
        L.orc_weno5.restype = C.c_double
        L.orc_weno5.argtypes = [C.c_double] * 5
        L.orc_get_difference.restype = C.c_double
        L.orc_state_create.restype = C.c_void_p
        L.orc_state_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcParams)]
        L.orc_state_destroy.argtypes = [C.c_void_p]
        L.orc_state_field.restype = _D
        L.orc_state_field.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long)]
        L.orc_state_load.argtypes = [C.c_void_p] + [_D] * 5
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(OrcDiag)]
        L.orc_step_timed.argtypes = [C.c_void_p, _D, _D, _I]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_pressure_solve.restype = C.c_int
        L.orc_mg_apply.restype = C.c_int
        L.orc_mg_apply.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int, _D, _D]
        L.orc_dim_diagonal.restype = C.c_double
        L.orc_dim_diagonal.argtypes = [C.c_double] * 6
        L.orc_dim_inlet.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _D]
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> _Lib:
    global _ref
    if _ref is None:
        _ref = _Lib(REF_SO, by_reference=True)
    return _ref


def oracle(name, *args):
    """Call orc_<name>; returns output scalars (list) or the C return value."""
    if name == "get_difference":
        L = oracle_lib()
        ret, _ = L.call("orc_get_difference", "aaiiiiii", args, restype=C.c_double)
        return ret
    spec = SPECS[name][0] if name in SPECS else ORACLE_ONLY[name]
    _, outs = oracle_lib().call("orc_" + name, spec, args)
    return outs


def ref(name, *args):
    """Call the reference routine <module>::<name>; returns output scalars."""
    spec, mod, proc = SPECS[name]
    _, outs = ref_lib().call(f"_QM{mod}P{proc}", spec, args)
    return outs


def pressure_solve(pres, rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt, rtol=1e-10,
                   abstol=1e-50, dtol=1e5, maxits=10000, pc=0):
    """orc_pressure_solve on Fortran-ordered arrays (pres updated in place);
    returns (reason, its, rnorm)."""
    L = oracle_lib().lib
    its, rn = C.c_int(0), C.c_double(0.0)
    ptr = lambda a: a.ctypes.data_as(_D)
    reason = L.orc_pressure_solve(ptr(pres), ptr(rhs_p), ptr(u3), ptr(u_inlet), int(maxl), int(maxm),
                                  int(maxn), C.c_double(dx), C.c_double(dy), C.c_double(dz),
                                  C.c_double(dt), C.c_double(rtol), C.c_double(abstol),
                                  C.c_double(dtol), int(maxits), int(pc), C.byref(its), C.byref(rn))
    return reason, its.value, rn.value


def mg_apply(r, nx, ny, nz, cx, cy, cz, diag, nu=1):
    """z = M r, one symmetric V-cycle of the multigrid preconditioner on an
    nx x ny x nz reduced operator (r, z flat, x fastest); returns (z, levels)."""
    L = oracle_lib().lib
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    nl = L.orc_mg_apply(nx, ny, nz, cx, cy, cz, diag, int(nu), r.ctypes.data_as(_D),
                        z.ctypes.data_as(_D))
    return z, nl


def set_threads(n: int) -> None:
    oracle_lib().lib.orc_set_threads(int(n))


# ---------------------------------------------------------------------------
# whole-step state
# ---------------------------------------------------------------------------
class OrcParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dx", "dy", "dz", "dt", "Re", "mu_plus", "mu_minus", "height",
                 "smooth_width", "scap")] + \
               [(n, C.c_int) for n in ("mom_u", "mom_v", "mom_w", "levelset_its")] + \
               [(n, C.c_double) for n in ("ksp_rtol", "ksp_abstol", "ksp_dtol")] + \
               [("ksp_maxits", C.c_int), ("ksp_pc", C.c_int), ("idm_dim", C.c_int),
                ("epn", C.c_double), ("Pe", C.c_double), ("dim_its", C.c_int)]


# OrcParams fields a parameter dict may leave out (the level-set method's
# defaults).
OPTIONAL_PARAMS = {"ksp_pc": 0, "idm_dim": 0, "epn": 0.0, "Pe": 1.0, "dim_its": 0}


class OrcDiag(C.Structure):
    _fields_ = [("err_phi", C.c_double), ("err_u", C.c_double), ("err_v", C.c_double),
                ("err_w", C.c_double), ("ksp_its", C.c_int), ("ksp_reason", C.c_int),
                ("ksp_rnorm", C.c_double)]


# Field ids, in the order of the ORC_* enum in tpls_oracle.h.
FIELD_NAMES = [
    "u2", "v2", "w2", "u3", "v3", "w3", "phi2", "phi3", "pres", "pres_old",
    "conv0_u", "conv1_u", "conv2_u", "conv0_v", "conv1_v", "conv2_v",
    "conv0_w", "conv1_w", "conv2_w",
    "csf_u0", "csf_u1", "csf_u2", "csf_v0", "csf_v1", "csf_v2", "csf_w0", "csf_w1", "csf_w2",
    "conv0_phi", "conv1_phi", "conv2_phi", "fx_csf", "fy_csf", "fz_csf",
    "visc2", "visc2_aug", "rhs_p", "diff_u", "diff_v", "diff_w", "rhs_u", "rhs_v", "rhs_w",
    "u_inlet", "convprev_phi",
]
FIELD_ID = {n: i for i, n in enumerate(FIELD_NAMES)}


def field_shape(name, maxl, maxm, maxn):
    """Fortran shape (extents) of a state field."""
    if name == "u_inlet":
        return (maxn - 1,)
    if name == "visc2_aug":
        return (maxl + 3, maxm + 3, maxn + 1)
    if name in ("w2", "w3") or name.endswith("_w") or name in ("csf_w0", "csf_w1", "csf_w2"):
        return (maxl + 1, maxm + 1, maxn)
    if name in ("phi2", "phi3", "pres", "pres_old", "conv0_phi", "conv1_phi", "conv2_phi",
                "fx_csf", "fy_csf", "fz_csf", "visc2", "rhs_p", "convprev_phi"):
        return (maxl + 1, maxm + 1, maxn + 1)
    return (maxl + 1, maxm + 1, maxn - 1)


def dim_diagonal(dx, dy, dz, dt, epn, Pe) -> float:
    return oracle_lib().lib.orc_dim_diagonal(dx, dy, dz, dt, epn, Pe)


def dim_inlet(maxn, dz, height, epn) -> np.ndarray:
    prof = np.zeros(maxn + 1)
    oracle_lib().lib.orc_dim_inlet(int(maxn), float(dz), float(height), float(epn), prof.ctypes.data_as(_D))
    return prof


class OracleState:
    """The oracle's copy of the full solver state (main_ns_hybrid.F90 arrays)."""

    def __init__(self, maxl, maxm, maxn, params: dict):
        self.L = oracle_lib().lib
        self.shape = (maxl, maxm, maxn)
        p = OrcParams()
        for f in fields_of(OrcParams):
            setattr(p, f, params.get(f, OPTIONAL_PARAMS[f]) if f in OPTIONAL_PARAMS else params[f])
        self._p = p
        self.h = self.L.orc_state_create(maxl, maxm, maxn, C.byref(p))

    def field(self, name) -> np.ndarray:
        cnt = C.c_long(0)
        ptr = self.L.orc_state_field(self.h, FIELD_ID[name], C.byref(cnt))
        shp = field_shape(name, *self.shape)
        flat = np.ctypeslib.as_array(ptr, shape=(cnt.value,))
        return flat.reshape(shp, order="F")

    def load(self, u, v, w, p, phi):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (u, v, w, p, phi)]
        self.L.orc_state_load(self.h, *[a.ctypes.data_as(_D) for a in arrs])

    def step(self, n=1) -> dict:
        d = OrcDiag()
        self.L.orc_step(self.h, int(n), C.byref(d))
        return {f: getattr(d, f) for f, _ in OrcDiag._fields_}

    def step_timed(self):
        t0, t1, its = C.c_double(), C.c_double(), C.c_int()
        self.L.orc_step_timed(self.h, C.byref(t0), C.byref(t1), C.byref(its))
        return t0.value, t1.value, its.value

    def close(self):
        if self.h:
            self.L.orc_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fields_of(struct):
    return [n for n, _ in struct._fields_]
