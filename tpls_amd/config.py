"""Case parameters of the TPLS stencil core and the standard workloads.

The parameter set is the subset of TPLS's options that the per-timestep core
reads (tpls_config.opt:29-34 for the iteration counts; initial_config.opt for
re, mu_minus, mu_plus, height, scap, dt, smooth_width_scale; the KSP
tolerances replace PETSc's runtime options, configure_mpi_petsc.F90:143-183).
Grids are isotropic with dx = dy = dz = 1/(maxn-1), as the reference's
initial-condition generator makes them (twophase_initialisation_wave.F90:66-68).
A grid of nx x ny x nz cells has maxl, maxm, maxn = nx+1, ny+1, nz+1.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass, field, replace


@dataclass
class CaseParams:
    nx: int
    ny: int
    nz: int
    Re: float = 100.0
    mu_minus: float = 1.0          # viscosity below the interface (phi < 0)
    mu_plus: float = 2.0           # viscosity above the interface (phi > 0)
    height: float = 0.5
    scap: float = 1.0
    smooth_width_scale: float = 1.5
    dt: float = 1e-4
    mom_u: int = 30
    mom_v: int = 30
    mom_w: int = 30
    levelset_its: int = 10
    ksp_rtol: float = 1e-10
    ksp_abstol: float = 1e-50
    ksp_dtol: float = 1e5
    ksp_maxits: int = 10000
    ksp_pc: str = "none"           # "none" (plain CG) or "mg" (multigrid-preconditioned CG)
    ksp_mg_nu: int = 2             # pre- and post-smoothing sweeps of the V-cycle
    noise: float = 1e-3            # amplitude of the seeded uniform perturbation
    seed: int = 0
    # interface method (initial_config.opt "idm"): "lsm" level set or "dim"
    # diffuse interface (Cahn-Hilliard, cahn_hilliard_solver.f90)
    idm: str = "lsm"
    epn: float | None = None       # DIM interface width; None: 0.5 dz (twophase_initialisation_wave.F90:71)
    Pe: float | None = None        # DIM Peclet number; None: 1/epn^2 (twophase_initialisation_wave.F90:72)
    dim_its: int = 18              # max_iteration_dim (tpls_config.opt:55)

    @property
    def maxl(self) -> int:
        return self.nx + 1

    @property
    def maxm(self) -> int:
        return self.ny + 1

    @property
    def maxn(self) -> int:
        return self.nz + 1

    @property
    def dx(self) -> float:
        return 1.0 / (self.maxn - 1)

    @property
    def smooth_width(self) -> float:
        return self.smooth_width_scale * self.dx

    @property
    def dim_epn(self) -> float:
        return 0.5 * self.dx if self.epn is None else self.epn

    @property
    def dim_Pe(self) -> float:
        e = self.dim_epn
        return 1.0 / (e * e) if self.Pe is None else self.Pe

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def check(self) -> None:
        if min(self.nx, self.ny, self.nz) < 4:
            raise ValueError("grid must be at least 4 cells in every direction")
        if self.ny % 2:
            # red-black colouring stays consistent across the periodic y seam
            # only for an even number of y cells (sor_iteration_allflux.f90)
            raise ValueError("ny (= maxm-1) must be even")
        for n in ("mom_u", "mom_v", "mom_w"):
            if getattr(self, n) < 2:
                raise ValueError(f"{n} must be >= 2")
        if self.idm not in ("lsm", "dim"):
            raise ValueError("idm must be lsm or dim")
        if self.idm == "dim" and self.dim_its < 1:
            raise ValueError("dim_its (max_iteration_dim) must be >= 1")

    def solver_dict(self) -> dict:
        d = self.dx
        return dict(dx=d, dy=d, dz=d, dt=self.dt, Re=self.Re, mu_plus=self.mu_plus,
                    mu_minus=self.mu_minus, height=self.height,
                    smooth_width=self.smooth_width, scap=self.scap,
                    mom_u=self.mom_u, mom_v=self.mom_v, mom_w=self.mom_w,
                    levelset_its=self.levelset_its, ksp_rtol=self.ksp_rtol,
                    ksp_abstol=self.ksp_abstol, ksp_dtol=self.ksp_dtol,
                    ksp_maxits=self.ksp_maxits, ksp_pc=self.pc_code(),
                    idm_dim=1 if self.idm == "dim" else 0, epn=self.dim_epn, Pe=self.dim_Pe,
                    dim_its=self.dim_its)

    def pc_code(self) -> int:
        """Preconditioner code of the solver interface: 0 none, n >= 1 the
        multigrid V(n,n) cycle."""
        return {"none": 0, "mg": self.ksp_mg_nu}[self.ksp_pc]


# Plain method, FP64, single case, CPU-runnable: 128x32x64, Re 100, viscosity
# ratio 2, dt 1e-4, 100 steps, noise 1e-3, CG rtol 1e-10.
CHANNEL_128x32x64 = CaseParams(nx=128, ny=32, nz=64, Re=100.0, mu_plus=2.0, dt=1e-4)
# Plain method, FP64, single case on the GPU: 512x128x256 (16.8M cells),
# Re 500, viscosity ratio 10, dt 5e-5, 500 steps.  The density ratio 1.2 of
# this workload has no counterpart in TPLS (constant density); it runs at 1.
CHANNEL_512x128x256 = CaseParams(nx=512, ny=128, nz=256, Re=500.0, mu_plus=10.0, dt=5e-5)
# The batched-workload grid menu.
BATCH_GRIDS = [(64, 16, 32), (128, 32, 64), (192, 48, 96), (256, 64, 128)]


def scaled(base: CaseParams, **kw) -> CaseParams:
    return replace(base, **kw)
