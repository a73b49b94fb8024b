"""One TPLS timestep composed from the reference's own compiled routines.

Every module procedure (WENO, levelset status / iteration, CSF, viscosity,
conv/diffusion, Jacobi, SOR, divergence, projection) is the reference Fortran
in oracle/_ref/libtpls_ref.so.  Only main_ns_hybrid.F90's own statements are
restated here, in numpy, at the reference's exchange points: the AB3 / CN
right-hand sides, boundary conditions, outer Helmholtz loops, the phi
augmentation of do_advect_phi and the level-set driver do_tpls (which need the
tpls_mpi module and so cannot be compiled alone), and the history rotation.
The pressure solve is the oracle's (PETSc is unavailable).

Used by tests/test_oracle_step.py to pin the oracle's step driver.
"""
import numpy as np

from oracle import binding as B


def xchg(a, hw=1):
    """exchange2d* on one y-periodic rank: halo rows <- periodic wrap."""
    ny = a.shape[1] - 2 * hw
    for d in range(1, hw + 1):
        a[:, hw - d, :] = a[:, hw - d + ny, :]
        a[:, hw + ny - 1 + d, :] = a[:, hw + d - 1, :]


def zeros(shape):
    return np.zeros(shape, order="F")


class RefStep:
    def __init__(self, maxl, maxm, maxn, prm, u, v, w, pres, phi):
        self.maxl, self.maxm, self.maxn = maxl, maxm, maxn
        self.P = prm
        L, M, N = maxl, maxm, maxn
        uv, ws, ps = (L + 1, M + 1, N - 1), (L + 1, M + 1, N), (L + 1, M + 1, N + 1)
        self.u2, self.v2, self.w2 = (np.array(a, order="F", dtype=float) for a in (u, v, w))
        self.pres = np.array(pres, order="F", dtype=float)
        self.phi2 = np.array(phi, order="F", dtype=float)
        self.u_inlet = np.ascontiguousarray(self.u2[0, 0, :])
        self.pres_old = zeros(ps)
        self.phi3 = zeros(ps)
        self.conv = {c: [zeros(uv if c != "w" else ws) for _ in range(3)] for c in "uvw"}
        self.csf = {c: [zeros(uv if c != "w" else ws) for _ in range(3)] for c in "uvw"}
        self.cphi = [zeros(ps) for _ in range(3)]
        self.fx, self.fy, self.fz = zeros(ps), zeros(ps), zeros(ps)
        self.visc2 = zeros(ps)
        self.visc_aug = zeros((L + 3, M + 3, N + 1))
        self.u3, self.v3, self.w3 = zeros(uv), zeros(uv), zeros(ws)

    def visc_of_phi(self, pv):
        P = self.P
        sw = P["smooth_width"]
        pi = float(np.float32(3.14159265359))
        H = np.where(pv < -sw, 0.0, np.where(
            pv > sw, 1.0, 0.5 + (pv / (2.0 * sw)) + (1.0 / (2.0 * pi)) * np.sin(pi * pv / sw)))
        return (P["mu_plus"] * H + P["mu_minus"] * (1.0 - H)) / P["Re"]

    def step(self):
        L, M, N = self.maxl, self.maxm, self.maxn
        sx, ex, sy, ey, exm = 1, L - 1, 1, M - 1, L - 1
        P = self.P
        dx, dy, dz, dt = P["dx"], P["dy"], P["dz"], P["dt"]
        # ---- do_advect_phi (advect_phi.f90:58-118) ----
        aug = zeros((L + 5, M + 5, N + 1))
        aug[3:L + 2, 3:M + 2, :] = self.phi2[1:L, 1:M, :]      # (sx:ex, sy:ey)
        aug[2, 3:M + 2, :] = self.phi2[0, 1:M, :]
        aug[L + 2, 3:M + 2, :] = self.phi2[L, 1:M, :]
        xchg(aug, 3)
        dfx, dfy, dfz = zeros(self.u2.shape), zeros(self.u2.shape), zeros(self.w2.shape)
        B.ref("weno_appox_convection", aug, self.u2, self.v2, self.w2, sx, ex, sy, ey, L, N,
              dfx, dfy, dfz)
        for a in (dfx, dfy, dfz, self.u2, self.v2, self.w2):
            xchg(a)
        c2 = self.cphi[2]
        c2[:] = 0.0
        i = slice(1, L)
        jc, jm = slice(1, M), slice(0, M - 1)
        kc, km = slice(1, N), slice(0, N - 1)
        u2, v2, w2 = self.u2, self.v2, self.w2
        c2[i, jc, kc] = ((u2[1:L, jm, km] * dfx[1:L, jm, km] - u2[0:L - 1, jm, km] * dfx[0:L - 1, jm, km]) / dx
                         + (v2[i, jc, km] * dfy[i, jc, km] - v2[i, jm, km] * dfy[i, jm, km]) / dy
                         + (w2[i, jm, kc] * dfz[i, jm, kc] - w2[i, jm, km] * dfz[i, jm, km]) / dz)
        # ---- do_tpls (two_phase_levelset.f90:56-116) ----
        c0, c1 = self.cphi[0], self.cphi[1]
        rhs = (23.0 / 12.0) * c2 - (4.0 / 3.0) * c1 + (5.0 / 12.0) * c0
        xchg(rhs)
        p3 = self.phi3
        p3[1:L, 1:M, 1:N] = self.phi2[1:L, 1:M, 1:N] - dt * rhs[1:L, 1:M, 1:N]
        zv = np.arange(1, N) * dz - 0.5 * dz
        p3[0, 1:M, 1:N] = (zv - P["height"])[None, :]
        p3[L, 1:M, 1:N] = p3[L - 1, 1:M, 1:N]
        p3[0:L + 1, 1:M, 0] = 2.0 * p3[0:L + 1, 1:M, 1] - p3[0:L + 1, 1:M, 2]
        p3[0:L + 1, 1:M, N] = 2.0 * p3[0:L + 1, 1:M, N - 1] - p3[0:L + 1, 1:M, N - 2]
        xchg(p3)
        D, sg = zeros(p3.shape), zeros(p3.shape)
        ist = np.zeros(p3.shape, order="F", dtype=np.int32)
        B.ref("get_levelset_status", p3, D, ist, sg, sx, ex, sy, ey, N, dx)
        pr = p3.copy(order="F")
        err = 0.0
        for _ in range(P["levelset_its"]):
            xchg(pr)
            err = B.ref("do_levelset_iteration", p3, pr, D, ist, sg, sx, ex, sy, ey, N, exm, dx,
                        P["height"])[0]
        p3[:] = pr
        self.err_phi = err
        # ---- CSF (main_ns_hybrid.F90:618-641) ----
        xchg(p3)
        B.ref("get_csf", self.fx, self.fy, self.fz, p3, sx, ex, sy, ey, N, exm, dx, dy, dz,
              P["scap"], P["smooth_width"])
        for a in (self.fx, self.fy, self.fz):
            xchg(a)
        cu, cv, cw = self.csf["u"][2], self.csf["v"][2], self.csf["w"][2]
        fx, fy, fz = self.fx, self.fy, self.fz
        cu[1:L, 1:M, 0:N - 1] = (fx[2:L + 1, 2:M + 1, 1:N] + fx[1:L, 2:M + 1, 1:N]) / 2.0
        cv[1:L, 1:M, 0:N - 1] = (fy[1:L, 2:M + 1, 1:N] + fy[1:L, 1:M, 1:N]) / 2.0
        cw[1:L, 1:M, 0:N] = (fz[1:L, 2:M + 1, 1:N + 1] + fz[1:L, 2:M + 1, 0:N]) / 2.0
        # ---- viscosity (main_ns_hybrid.F90:646-718) ----
        B.ref("get_viscosity", self.visc2, self.phi2, ex, ey, sx, sy, N, P["smooth_width"], P["Re"],
              P["mu_minus"], P["mu_plus"])
        va = self.visc_aug           # index offset: aug i -> i+1, j -> j+1
        va[2:L + 1, 2:M + 1, :] = self.visc2[1:L, 1:M, :]
        m0 = self.visc_of_phi(self.phi2[0, 1:M, :])
        m1 = self.visc_of_phi(self.phi2[L, 1:M, :])
        va[1, 2:M + 1, :] = m0
        self.visc2[0, 1:M, :] = m0
        va[L + 1, 2:M + 1, :] = m1
        self.visc2[L, 1:M, :] = m1
        # exchange2d_aug1 then aug2 on (sx-2:ex+2, sy-2:ey+2)
        ny = M - 1
        va[1:L + 2, 1, :] = va[1:L + 2, 1 + ny, :]
        va[1:L + 2, 2 + ny, :] = va[1:L + 2, 2, :]
        va[0:L + 3, 0, :] = va[0:L + 3, ny, :]
        va[0:L + 3, 3 + ny, :] = va[0:L + 3, 3, :]
        xchg(self.visc2)
        # ---- momentum fluxes and AB3 + CN right-hand sides (main:724-757) ----
        for a in (self.u2, self.v2, self.w2):
            xchg(a)
        cv2 = [self.conv[c][2] for c in "uvw"]
        B.ref("get_conv_all", *cv2, va, self.u2, self.v2, self.w2, sx, sy, ex, ey, N, exm, dx, dy, dz)
        diff = [zeros(self.u2.shape), zeros(self.u2.shape), zeros(self.w2.shape)]
        B.ref("get_diffusion_all", *diff, va, self.u2, self.v2, self.w2, sx, sy, ex, ey, N, exm,
              dx, dy, dz)
        R = {}
        for n, c in enumerate("uvw"):
            x2 = (self.u2, self.v2, self.w2)[n]
            C0, C1, C2 = self.conv[c]
            S0, S1, S2 = self.csf[c]
            R[c] = (x2 - dt * ((23.0 / 12.0) * C2 - (4.0 / 3.0) * C1 + (5.0 / 12.0) * C0)
                    + (dt / 2.0) * diff[n]
                    + dt * ((23.0 / 12.0) * S2 - (4.0 / 3.0) * S1 + (5.0 / 12.0) * S0))
            R[c] = np.asfortranarray(R[c])
        # ---- Helmholtz loops (main:758-887) ----
        pres = self.pres
        u3 = self.u3
        u3[:] = 0.0

        def ubc():
            u3[0, 1:M, :] = u3[1, 1:M, :]
            u3[exm, 1:M, :] = u3[exm - 1, 1:M, :] + (dt / dx) * (
                pres[exm + 1, 1:M, 1:N] + pres[exm - 1, 1:M, 1:N] - 2.0 * pres[exm, 1:M, 1:N])
            u3[exm + 1, 1:M, :] = 0.0
        for it in range(1, P["mom_u"] // 2 + 1):
            uo = u3.copy(order="F")
            xchg(uo)
            B.ref("do_jacobi_u", u3, uo, R["u"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, exm)
            ubc()
            xchg(u3)
            B.ref("do_sor_u", u3, R["u"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, exm, 0, it)
            xchg(u3)
            B.ref("do_sor_u", u3, R["u"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, exm, 1, it)
            ubc()
        xchg(pres)
        v3 = self.v3
        v3[:] = 0.0
        for it in range(1, P["mom_v"] // 2 + 1):
            vo = v3.copy(order="F")
            xchg(vo)
            B.ref("do_jacobi_v", v3, vo, R["v"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, exm)
            xchg(v3)
            B.ref("do_sor_v", v3, R["v"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, 0, it)
            xchg(v3)
            B.ref("do_sor_v", v3, R["v"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, 1, it)
            v3[0, 1:M, :] = v3[1, 1:M, :]
            v3[exm + 1, 1:M, :] = v3[exm, 1:M, :]
        w3 = self.w3
        w3[:] = 0.0
        for it in range(1, P["mom_w"] // 2 + 1):
            wo = w3.copy(order="F")
            xchg(wo)
            B.ref("do_jacobi_w", w3, wo, R["w"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, exm)
            xchg(w3)
            B.ref("do_sor_w", w3, R["w"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, 0, it)
            xchg(w3)
            B.ref("do_sor_w", w3, R["w"], va, dx, dy, dz, dt, ex, ey, sx, sy, N, 1, it)
            w3[1:L, 1:M, 0] = 0.0
            w3[1:L, 1:M, N - 1] = 0.0
            w3[0, 1:M, 1:N - 1] = w3[1, 1:M, 1:N - 1]
            w3[0, 1:M, 0] = 0.0
            w3[0, 1:M, N - 1] = 0.0
            w3[exm + 1, 1:M, :] = w3[exm, 1:M, :]
        # ---- divergence, pressure solve, projection (main:889-918) ----
        for a in (u3, v3, w3):
            xchg(a)
        rhs_p = zeros(pres.shape)
        B.ref("get_source_pres", rhs_p, u3, v3, w3, ex, ey, sx, sy, N, dx, dy, dz, dt)
        xchg(rhs_p)
        pres[:] = self.pres_old
        L_ = B.oracle_lib().lib
        import ctypes as C
        its, rn = C.c_int(0), C.c_double(0)
        D_ = C.POINTER(C.c_double)
        L_.orc_pressure_solve(pres.ctypes.data_as(D_), rhs_p.ctypes.data_as(D_), u3.ctypes.data_as(D_),
                              self.u_inlet.ctypes.data_as(D_), L, M, N, C.c_double(dx), C.c_double(dy),
                              C.c_double(dz), C.c_double(dt), C.c_double(P["ksp_rtol"]),
                              C.c_double(P["ksp_abstol"]), C.c_double(P["ksp_dtol"]),
                              P["ksp_maxits"], P.get("ksp_pc", 0), C.byref(its), C.byref(rn))
        self.ksp_its = its.value
        B.ref("get_uvw_pres", u3, v3, w3, pres, ex, ey, sx, sy, N, exm, dx, dy, dz, dt, self.u_inlet)
        # ---- rotation (main:2480-2509) ----
        for c in "uvw":
            self.conv[c] = [self.conv[c][1], self.conv[c][2], self.conv[c][0]]
            self.csf[c] = [self.csf[c][1], self.csf[c][2], self.csf[c][0]]
        self.cphi = [self.cphi[1], self.cphi[2], self.cphi[0]]
        self.pres_old = pres.copy(order="F")
        self.u2, self.v2, self.w2 = u3.copy(order="F"), v3.copy(order="F"), w3.copy(order="F")
        self.phi2 = p3.copy(order="F")
