"""ctypes binding of the HIP library ``tpls_amd/libtpls_hip.so``.

The library is the only compute path: importing this module fails loudly when
it has not been built (``python -m tpls_amd.build``), and every wrapper
raises ``TplsError`` on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtpls_hip.so")


class TplsError(RuntimeError):
    pass


class Params(C.Structure):
    """tpls_params of include/tpls_core.h"""
    _fields_ = [(n, C.c_double) for n in
                ("dx", "dy", "dz", "dt", "Re", "mu_plus", "mu_minus", "height",
                 "smooth_width", "scap")] + \
               [(n, C.c_int) for n in ("mom_u", "mom_v", "mom_w", "levelset_its")] + \
               [(n, C.c_double) for n in ("ksp_rtol", "ksp_abstol", "ksp_dtol")] + \
               [("ksp_maxits", C.c_int), ("precision", C.c_int), ("diagnostics", C.c_int),
                ("ksp_pc", C.c_int), ("graphs", C.c_int), ("tile_rows", C.c_int),
                ("tile_rows_odd", C.c_int), ("prefetch_odd", C.c_int), ("sor_fused", C.c_int),
                ("tile_rows_rb", C.c_int), ("tile_rows_rb_odd", C.c_int), ("tile_rows_jst", C.c_int),
                ("mg_fp32", C.c_int), ("idm_dim", C.c_int), ("epn", C.c_double), ("Pe", C.c_double),
                ("dim_its", C.c_int), ("dim_split", C.c_int),
                ("march_levels", C.c_int), ("mg_pair", C.c_int)]


class Diag(C.Structure):
    _fields_ = [("err_phi", C.c_double), ("err_u", C.c_double), ("err_v", C.c_double),
                ("err_w", C.c_double), ("ksp_its", C.c_int), ("ksp_reason", C.c_int),
                ("ksp_rnorm", C.c_double)]


FP64, FP32 = 0, 1

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

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_CASE = C.c_void_p

# argument kinds of the per-routine entry points: a = double array,
# I = int32 array, i = int, d = double, o = double* output scalar
ROUTINES = {
    "wenoappoxconvection": "aaaaiiiiiiaaa",
    "do_advect_phi": "aaaaaiiiiiiiddd",
    "do_tpls": "aaaaaiiiiiiidoddd",
    "dim": "iiddddddaaaaaiiiidio",
    "get_levelset_status": "aaIaiiiiid",
    "do_levelset_iteration": "aaaIaiiiiiiddo",
    "get_csf": "aaaaiiiiiiddddd",
    "get_diffusion_all": "aaaaaaaiiiiiiddd",
    "get_conv_all": "aaaaaaaiiiiiiddd",
    "get_viscosity": "aaiiiiidddd",
    "do_jacobi_u": "aaaaddddiiiiii",
    "do_jacobi_v": "aaaaddddiiiiii",
    "do_jacobi_w": "aaaaddddiiiiii",
    "get_difference": "aaiiiiiio",
    "do_sor_u": "aaaddddiiiiiiii",
    "do_sor_v": "aaaddddiiiiiii",
    "do_sor_w": "aaaddddiiiiiii",
    "get_source_pres": "aaaaiiiiidddd",
    "get_uvw_pres": "aaaaiiiiiidddda",
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TplsError(f"{LIB_PATH} is not built; run `python -m tpls_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.tpls_last_error.restype = C.c_char_p
        L.tpls_case_bytes.restype = C.c_size_t
        L.tpls_case_bytes.argtypes = [C.c_int] * 4
        L.tpls_case_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.POINTER(_CASE)]
        L.tpls_case_destroy.argtypes = [_CASE]
        L.tpls_case_load.argtypes = [_CASE] + [_D] * 5
        L.tpls_case_get_field.argtypes = [_CASE, C.c_int, _D]
        L.tpls_case_set_field.argtypes = [_CASE, C.c_int, _D]
        L.tpls_step.argtypes = [_CASE, C.c_int]
        L.tpls_case_diag.argtypes = [_CASE, C.POINTER(Diag)]
        L.tpls_case_sync.argtypes = [_CASE]
        L.tpls_case_time_kernels.argtypes = [_CASE, C.c_int]
        L.tpls_case_kernel_time.argtypes = [_CASE, _D, C.POINTER(C.c_long)]
        L.tpls_pressure_solve.argtypes = [_D, _D, _D, _D, C.c_int, C.c_int, C.c_int] + \
            [C.c_double] * 7 + [C.c_int, _I, _D, _I]
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().tpls_last_error().decode(errors="replace")
        raise TplsError(f"{what} failed with status {rc}: {msg}")


def _ptr(a, kind):
    if kind == "a":
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)):
            raise TypeError("double arrays must be Fortran-contiguous float64 numpy arrays")
        return a.ctypes.data_as(_D)
    if not (isinstance(a, np.ndarray) and a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim == 1)):
        raise TypeError("integer arrays must be Fortran-contiguous int32 numpy arrays")
    return a.ctypes.data_as(_I)


def call(name: str, *args):
    """Call tpls_<name> with the reference's argument order (outputs scalars
    of kind 'o' are returned as a list)."""
    spec = ROUTINES[name]
    fn = getattr(lib(), "tpls_" + name)
    cargs, outs = [], []
    it = iter(args)
    n_in = len(spec) - spec.count("o")
    if len(args) != n_in:
        raise TypeError(f"tpls_{name} takes {n_in} arguments, got {len(args)}")
    for kind in spec:
        if kind in "aI":
            cargs.append(_ptr(next(it), kind))
        elif kind == "i":
            cargs.append(C.c_int(int(next(it))))
        elif kind == "d":
            cargs.append(C.c_double(float(next(it))))
        else:
            box = C.c_double(0.0)
            outs.append(box)
            cargs.append(C.byref(box))
    check(fn(*cargs), "tpls_" + name)
    return [b.value for b in outs]


def pressure_solve(pres, rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt,
                   rtol, abstol, dtol, maxits):
    its, rnorm, reason = C.c_int(0), C.c_double(0), C.c_int(0)
    check(lib().tpls_pressure_solve(_ptr(pres, "a"), _ptr(rhs_p, "a"), _ptr(u3, "a"),
                                    _ptr(u_inlet, "a"), maxl, maxm, maxn, dx, dy, dz, dt,
                                    rtol, abstol, dtol, maxits, C.byref(its), C.byref(rnorm),
                                    C.byref(reason)), "tpls_pressure_solve")
    return its.value, rnorm.value, reason.value


def exchange2d(uu, ex, ey, ez, sx, sy, sz, kind=0):
    """tpls_exchange2d: mpi.f90's exchange2d (kind 0), exchange2d_aug1 (1) or
    exchange2d_aug2 (2) on one rank, in place."""
    f = lib().tpls_exchange2d
    f.argtypes = [_D] + [C.c_int] * 7
    check(f(_ptr(uu, "a"), ex, ey, ez, sx, sy, sz, kind), "tpls_exchange2d")


def mirror_attach(a):
    """Attach a Fortran-contiguous float64 array to the device mirror."""
    f = lib().tpls_mirror_attach
    f.argtypes = [C.c_void_p, C.c_longlong]
    _ptr(a, "a")
    check(f(a.ctypes.data, a.nbytes), "tpls_mirror_attach")


def mirror_detach(a):
    f = lib().tpls_mirror_detach
    f.argtypes = [C.c_void_p]
    check(f(a.ctypes.data), "tpls_mirror_detach")


def mirror_counters() -> dict:
    """Bytes up / down over the host link, faults resolved, copies, arrays attached."""
    c = (C.c_longlong * 6)()
    check(lib().tpls_mirror_counters(c), "tpls_mirror_counters")
    return dict(zip(("up", "down", "faults", "uploads", "downloads", "attached"), list(c)))


def test_mg_apply(r, nx, ny, nz, cx, cy, cz, diag, nu):
    """Test hook: z = M r, one V-cycle of the pressure solve's multigrid
    preconditioner on the GPU (r flat, x fastest); returns (z, levels)."""
    f = lib().tpls_test_mg_apply
    f.restype = C.c_int
    f.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int, C.c_void_p, C.c_void_p]
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    nl = f(nx, ny, nz, cx, cy, cz, diag, nu, r.ctypes.data, z.ctypes.data)
    check(min(nl, 0), "tpls_test_mg_apply")
    return z, nl
