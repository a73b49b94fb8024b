/*
 * CPU oracle for the TPLS per-timestep stencil core.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Python host package, the Fortran wrapper modules) links, loads or calls
 * this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the CPU timing
 * baseline.
 *
 * Each orc_* routine restates one TPLS routine (Q-Leap-Networks/tpls,
 * the src/ Fortran files) in plain C with the SAME argument order and the SAME Fortran
 * array shapes: every array argument points at the element with the lowest
 * Fortran index and is laid out column-major (x fastest).  Integers and
 * doubles are passed by value.  The single-rank domain is assumed
 * (sx = 1, ex = maxl-1, sy = 1, ey = maxm-1); MPI halo exchanges become the
 * periodic copies a one-rank y-periodic Cartesian communicator performs.
 *
 * Floating-point rules: compiled with -ffp-contract=off and without
 * fast-math, so each expression rounds exactly as written, which is how
 * gfortran/flang evaluate the reference on x86-64 (no FMA contraction).
 *
 * Parity pinning (see tests/test_oracle_vs_ref.py, tests/test_oracle_step.py
 * and tests/test_oracle_pressure.py): every routine that exists in a
 * dependency-free reference module is checked bit-for-bit against the
 * reference Fortran compiled by oracle/Makefile into oracle/_ref/.  The step
 * driver's main-program pieces (CSF face averages, inlet/outlet viscosity,
 * AB3 right-hand sides, velocity boundary conditions, history rotation) and
 * the two modules that need MPI (advect_phi.f90, two_phase_levelset.f90)
 * are pinned by what the mathematics fixes (tests/test_oracle_glue.py:
 * exactness on uniform, linear and quadratic-in-time data, the reference's
 * own routines applied to the planes main builds, zero divergence after the
 * step), and that file shows that copies of this oracle with plausible
 * mistakes planted in those statements fail it.  The
 * pressure solve restates PETSc 3.6 KSPCG (pc none, constant null space) on
 * the symmetric reduced system: PETSc is not available here, so its parity
 * with PETSc itself (iterates, iteration counts) is UNPINNED.  Its solution is
 * pinned against a scipy direct solve of the full DMDA system that
 * compute_matrix / compute_rhs assemble, ghost rows included, and the
 * multigrid preconditioner by symmetry, definiteness and contraction checks.
 */
#ifndef TPLS_ORACLE_H
#define TPLS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference routine restatements (argument order as in the Fortran) ---- */

/* weno.f90:175-215 */
double orc_weno5(double v1, double v2, double v3, double v4, double v5);
/* weno.f90:18-172 */
void orc_weno_appox_convection(const double *phi, const double *u, const double *v,
                               const double *w, int sx, int ex, int sy, int ey, int l,
                               int n, double *dfx, double *dfy, double *dfz);
/* advect_phi.f90:19-119 (MPI arguments dropped; single rank) */
void orc_do_advect_phi(double *phi2, double *u2, double *v2, double *w2, double *conv2_phi,
                       int sx, int ex, int sy, int ey, int maxl, int maxn, int ex_max,
                       double dx, double dy, double dz);
/* two_phase_levelset.f90:20-118 (MPI arguments dropped; single rank) */
void orc_do_tpls(double *phi2, double *phi3, double *conv0_phi, double *conv1_phi,
                 double *conv2_phi, int sx, int ex, int sy, int ey, int maxn, int ex_max,
                 int max_iteration_levelset, double height, double *err2, double dx,
                 double dz, double dt);
/* cahn_hilliard_solver.f90:41-384, dim (MPI arguments dropped; single rank).
 * phi and phiOld are the phi2+0.5 and phi_old+0.5 temporaries that
 * main_ns_hybrid.F90:608 passes; phiNew is the initial iterate on entry and
 * the last one, with its boundary planes, on return.  res: the largest
 * |change| of the last sweep (the reference computes it and never uses it).
 * The reference module uses the MPI module for its halo exchanges, so it
 * cannot be compiled here; this restatement is pinned by closed forms, by
 * the fixed point of its sweep against a direct solve of the linear system
 * the sweep relaxes, and by that system's diagonal (tests/test_oracle_dim.py). */
void orc_dim(int maxl, int maxn, double dx, double dy, double dz, double dt, double epn,
             double Pe, const double *phi, const double *phiOld, const double *conv2_phi,
             double *ConvPrevPhi, double *phiNew, int sx, int ex, int sy, int ey, double height,
             int max_iteration_dim, double *res);
/* The scalar factors of dim: diagonal (cahn_hilliard_solver.f90:243-245) and
 * the inlet profile phiNew(0,.,k), k = 0..maxn (k = 0 and maxn unused)
 * (cahn_hilliard_solver.f90:350-352). */
double orc_dim_diagonal(double dx, double dy, double dz, double dt, double epn, double Pe);
void orc_dim_inlet(int maxn, double dz, double height, double epn, double *prof);
/* ---- main-program statements of the step (main_ns_hybrid.F90) ----
 * Single rank, arrays in the shapes of main_ns_hybrid.F90:416-549 (p shape
 * (0:maxl,0:maxm,0:maxn), u/v (0:maxl,0:maxm,0:maxn-2), w (0:maxl,0:maxm,
 * 0:maxn-1), visc2_aug (-1:maxl+1,-1:maxm+1,0:maxn)).  The reference has them
 * inline in the main program, which needs MPI, PETSc and netCDF and cannot be
 * compiled here; each is pinned by what the mathematics fixes
 * (tests/test_oracle_glue.py). */
/* main_ns_hybrid.F90:629-641: CSF forces averaged onto the velocity faces */
void orc_csf_faces(const double *fx, const double *fy, const double *fz, double *csf_u,
                   double *csf_v, double *csf_w, int maxl, int maxm, int maxn);
/* main_ns_hybrid.F90:654-710: visc2_aug's interior and the inlet/outlet
 * planes of visc2 and visc2_aug from phi2 - shift */
void orc_visc_aug(double *visc2, double *visc2_aug, const double *phi2, int maxl, int maxm,
                  int maxn, double smooth_width, double Re, double mu_minus, double mu_plus,
                  double shift);
/* main_ns_hybrid.F90:743-757: one component's AB3 + Crank-Nicolson RHS */
void orc_momentum_rhs(double *R, const double *x2, const double *conv2, const double *conv1,
                      const double *conv0, const double *diff, const double *csf2,
                      const double *csf1, const double *csf0, long n, double dt);
/* main_ns_hybrid.F90:764-776 / 787-799: u3 inlet copy and outlet condition */
void orc_u_inout_bc(double *u3, const double *pres, int maxl, int maxm, int maxn, double dt,
                    double dx);
/* main_ns_hybrid.F90:826-836 and 864-879: v3 and w3 conditions */
void orc_v_bc(double *v3, int maxl, int maxm, int maxn);
void orc_w_bc(double *w3, int maxl, int maxm, int maxn);

/* levelset.f90:17-94 */
void orc_get_levelset_status(const double *phi, double *D_mx, int *istatus, double *sign_mx,
                             int sx, int ex, int sy, int ey, int n, double h);
/* levelset.f90:97-237 */
void orc_do_levelset_iteration(const double *phi, double *phi_reinit, const double *D_mx,
                               const int *istatus, const double *sign_mx, int sx, int ex,
                               int sy, int ey, int n, int ex_max, double h, double height,
                               double *error_val);
/* levelset.f90:241-414 */
void orc_get_csf(double *fx, double *fy, double *fz, const double *fi, int sx, int ex, int sy,
                 int ey, int n, int ex_max, double dx, double dy, double dz, double scap,
                 double smooth_width);
/* momentum_allflux.f90:22-214 */
void orc_get_diffusion_all(double *diff_u, double *diff_v, double *diff_w, const double *visc,
                           const double *u2, const double *v2, const double *w2, int sx,
                           int sy, int ex, int ey, int maxn, int ex_max, double dx, double dy,
                           double dz);
/* momentum_allflux.f90:224-471 */
void orc_get_conv_all(double *conv_u2, double *conv_v2, double *conv_w2, const double *visc,
                      const double *u2, const double *v2, const double *w2, int sx, int sy,
                      int ex, int ey, int maxn, int ex_max, double dx, double dy, double dz);
/* momentum_allflux.f90:474-521 */
void orc_get_viscosity(double *visc, const double *phi, int ex, int ey, int sx, int sy,
                       int maxn, double smooth_width, double Re, double mu_minus,
                       double mu_plus);
/* jacobi_iteration_allflux.f90:21-99 / 102-192 / 195-282 */
void orc_do_jacobi_u(double *u3, const double *u3_old, const double *RHS, const double *visc,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx,
                     int sy, int maxn, int ex_max);
void orc_do_jacobi_v(double *v3, const double *v3_old, const double *RHS, const double *visc,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx,
                     int sy, int maxn, int ex_max);
void orc_do_jacobi_w(double *w3, const double *w3_old, const double *RHS, const double *visc,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx,
                     int sy, int maxn, int ex_max);
/* jacobi_iteration_allflux.f90:285-311 */
double orc_get_difference(const double *ua, const double *ub, int ex, int ey, int ez, int sx,
                          int sy, int sz);
/* sor_iteration_allflux.f90:19-149 / 152-277 / 280-377 */
void orc_do_sor_u(double *u3, const double *RHS, const double *visc, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max,
                  int flg, int iteration_time);
void orc_do_sor_v(double *v3, const double *RHS, const double *visc, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int flg,
                  int iteration_time);
void orc_do_sor_w(double *w3, const double *RHS, const double *visc, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int flg,
                  int iteration_time);
/* pressure.f90:16-56 */
void orc_get_source_pres(double *rhs, const double *u3, const double *v3, const double *w3,
                         int ex, int ey, int sx, int sy, int maxn, double dx, double dy,
                         double dz, double dt);
/* pressure.f90:59-134 */
void orc_get_uvw_pres(double *u3, double *v3, double *w3, const double *p, int ex, int ey,
                      int sx, int sy, int maxn, int ex_max, double dx, double dy, double dz,
                      double dt, const double *u_inlet);

/* ---- pressure Poisson solve: PETSc 3.6 KSPCG restated on the reduced system ----
 * pressure_solver.F90:234-441 (RHS, operator, null space) and
 * main_ns_hybrid.F90:896-913 (initial guess pres_old, solve, copy-out).
 * pres: in = initial guess (interior used), out = solution incl. ghost planes
 * and periodic y halos.  Returns the KSP converged reason (PETSc codes). */
int orc_pressure_solve(double *pres, const double *RHS_p, const double *u3,
                       const double *u_inlet, int maxl, int maxm, int maxn, double dx,
                       double dy, double dz, double dt, double rtol, double abstol,
                       double dtol, int maxits, int pc, int *its, double *rnorm);
/* Preconditioner of the solve: pc = 0 none (PETSc -pc_type none), pc = n >= 1
 * the geometric multigrid V(n,n) cycle defined in tpls_oracle.c.  The stopping
 * test is the null-space-projected residual norm |r - mean(r)| <= rtol |b| for
 * both (PETSc's KSP_NORM_UNPRECONDITIONED; identical to its default norm when
 * pc = 0).  orc_mg_apply: z = M r for an nx x ny x nz reduced operator, flat
 * arrays with x fastest; returns the number of levels. */
int orc_mg_apply(int nx, int ny, int nz, double cx, double cy, double cz, double diag, int nu,
                 const double *r, double *z);

/* ---- whole-step driver: main_ns_hybrid.F90:583-918, 2480-2509 ---- */
typedef struct {
    double dx, dy, dz, dt, Re, mu_plus, mu_minus, height, smooth_width, scap;
    int mom_u, mom_v, mom_w, levelset_its;
    double ksp_rtol, ksp_abstol, ksp_dtol;
    int ksp_maxits;
    int ksp_pc;  /* 0 none, n >= 1 multigrid V(n,n) */
    int idm_dim; /* 0: level-set method (do_tpls), 1: diffuse interface (dim) */
    double epn, Pe;
    int dim_its; /* max_iteration_dim */
} orc_params;

typedef struct {
    double err_phi, err_u, err_v, err_w;
    int ksp_its, ksp_reason;
    double ksp_rnorm;
} orc_diag;

typedef struct orc_state orc_state;

/* Field identifiers; the numbering matches TPLS_F_* in include/tpls_core.h. */
enum {
    ORC_U2 = 0, ORC_V2, ORC_W2, ORC_U3, ORC_V3, ORC_W3,
    ORC_PHI2, ORC_PHI3, ORC_PRES, ORC_PRES_OLD,
    ORC_CONV0_U, ORC_CONV1_U, ORC_CONV2_U, ORC_CONV0_V, ORC_CONV1_V, ORC_CONV2_V,
    ORC_CONV0_W, ORC_CONV1_W, ORC_CONV2_W,
    ORC_CSF_U0, ORC_CSF_U1, ORC_CSF_U2, ORC_CSF_V0, ORC_CSF_V1, ORC_CSF_V2,
    ORC_CSF_W0, ORC_CSF_W1, ORC_CSF_W2,
    ORC_CONV0_PHI, ORC_CONV1_PHI, ORC_CONV2_PHI,
    ORC_FX_CSF, ORC_FY_CSF, ORC_FZ_CSF,
    ORC_VISC2, ORC_VISC2_AUG, ORC_RHS_P,
    ORC_DIFF_U, ORC_DIFF_V, ORC_DIFF_W, ORC_RHS_U, ORC_RHS_V, ORC_RHS_W,
    ORC_U_INLET, ORC_CONVPREV_PHI,
    ORC_NFIELDS
};

orc_state *orc_state_create(int maxl, int maxm, int maxn, const orc_params *prm);
void orc_state_destroy(orc_state *s);
/* Pointer to a field (Fortran shape, lowest index first) and its element count. */
double *orc_state_field(orc_state *s, int id, long *count);
/* Load initial conditions exactly as main_ns_hybrid.F90:412-549 does, from
 * u, v, w, pressure and phi arrays of the initial-file shapes. */
void orc_state_load(orc_state *s, const double *u, const double *v, const double *w,
                    const double *p, const double *phi);
/* Advance nsteps timesteps; diag receives the last step's diagnostics. */
void orc_step(orc_state *s, int nsteps, orc_diag *diag);
/* Time the pieces of one step separately (seconds), for the CPU baseline:
 * t[0] = everything except the pressure solve, t[1] = pressure solve,
 * its = CG iterations in that solve. */
void orc_step_timed(orc_state *s, double *t_nonp, double *t_p, int *its);
/* Threads the oracle's OpenMP loops may use (1 = fully sequential). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
