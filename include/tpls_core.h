/*
 * tpls_core.h -- C ABI of the MI355X-native TPLS stencil core.
 *
 * Two layers:
 *
 *  1. Per-routine drop-in entry points.  One entry per reference procedure on
 *     the per-timestep path, with the reference's argument order and meaning
 *     (Fortran arrays passed as pointers to their first element, column-major,
 *     with the reference's index bounds; integers and doubles by value).
 *     Host arrays go in and out; the computation runs on the GPU.  Each
 *     comment cites the reference interface it replaces.  The MPI-only
 *     arguments of the reference (neighbours, communicator, stride types) are
 *     dropped: the whole domain lives on one GPU, so the entry points require
 *     the single-rank decomposition sx = 1, ex = ex_max = maxl-1, sy = 1,
 *     ey = maxm-1 (and return TPLS_ERR_DECOMP otherwise).
 *
 *  2. A device-resident case: all state arrays of main_ns_hybrid.F90 stay in
 *     HBM and tpls_step() advances whole timesteps (main_ns_hybrid.F90:583-2509
 *     without I/O) with the same kernels.
 *
 * Every function returns TPLS_OK (0) or a negative TPLS_ERR_* code;
 * tpls_last_error() describes the last failure of the calling thread.
 * Calls on one case must not overlap; different cases may be driven from
 * different threads.
 */
#ifndef TPLS_CORE_H
#define TPLS_CORE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TPLS_OK = 0,
    TPLS_ERR_HIP = -1,       /* a HIP runtime call failed */
    TPLS_ERR_ARG = -2,       /* bad argument (null pointer, size, field id) */
    TPLS_ERR_DECOMP = -3,    /* not the single-rank decomposition */
    TPLS_ERR_NOMEM = -4,     /* device allocation failed */
    TPLS_ERR_NOGPU = -5      /* no usable gfx950 device */
};

/* Storage precision of a case. */
enum { TPLS_FP64 = 0, TPLS_FP32 = 1 };

/* Case parameters: the option values the core reads (tpls_config.opt:29-34,
 * initial_config.dat keys written by create_initial_conditions_program.F90:135-169),
 * plus the Krylov tolerances PETSc would take from its runtime options
 * (configure_mpi_petsc.F90:158, 177-182). */
typedef struct {
    double dx, dy, dz, dt, Re, mu_plus, mu_minus, height, smooth_width, scap;
    int mom_u, mom_v, mom_w, levelset_its;
    double ksp_rtol, ksp_abstol, ksp_dtol;
    int ksp_maxits;
    int precision;    /* TPLS_FP64 or TPLS_FP32 */
    int diagnostics;  /* 1: compute the per-step residual prints of
                         main_ns_hybrid.F90:603-605, 806-808, 843-845, 885-887 */
    int ksp_pc;       /* preconditioner of the pressure CG, the choice the
                         reference leaves to PETSc's options database
                         (KSPSetFromOptions, configure_mpi_petsc.F90:158):
                         0 none (-pc_type none); n >= 1 geometric multigrid,
                         one symmetric V(n,n) cycle per iteration (the class
                         of -pc_type mg).  The stopping test is the
                         unpreconditioned residual norm in both cases. */
    int graphs;       /* 1: replay each step's launches from HIP graphs
                         (captured on first use per buffer-role pattern);
                         identical results, far fewer host launches */
    int tile_rows;    /* launch shape of the Helmholtz sweeps (0 = default):
                         rows per 64-wide tile of the Jacobi and even-parity
                         SOR sweeps, 1/2/4/8 */
    int tile_rows_odd;  /* the same for the odd-parity (column) SOR sweep */
    int prefetch_odd;   /* levels the column sweep loads ahead, 4 or 8 */
    int sor_fused;      /* which SOR iterations run both colour sweeps in
                           one out-of-place pass over HBM instead of the
                           reference's two in-place sweeps, with identical
                           results (every cell sees the same inputs):
                           -1 none, 1 odd iteration_time (the column
                           sweeps), 2 even, 3 all; 0: default (1).  Bit 2
                           (4, 5): where the even iterations run split
                           with staged tiles and the y period is a
                           multiple of 8, their Jacobi step, its boundary
                           conditions and the first colour sweep run as
                           one pass (again identical results) */
    int tile_rows_rb;     /* rows per tile of that pass, even parity: 2/4/8 */
    int tile_rows_rb_odd; /* the same, odd parity (0 = default for both) */
    int tile_rows_jst;    /* Jacobi sweep with its planes staged through LDS,
                             tiles of 2/4/8/16 rows (0: default, 4); -1: the
                             unstaged sweep with tile_rows instead (same
                             results either way) */
    int mg_fp32;          /* precision of the multigrid V-cycle (ksp_pc >= 1):
                             0 default, FP32 in TPLS_FP32 cases and FP64 in
                             TPLS_FP64 cases; 1 FP32; -1 FP64.  The CG's inner
                             products, residual and stopping test stay FP64
                             either way; an FP32 cycle changes the iterates
                             within the solve's tolerance, not the tolerance */
    int idm_dim;          /* interface method (initial_config.opt "idm",
                             configure.F90:320-350): 0 the level-set method
                             (do_tpls), 1 the diffuse interface method
                             (dim, cahn_hilliard_solver.f90:41-384, called as
                             main_ns_hybrid.F90:606-611) */
    double epn, Pe;       /* DIM interface width and Peclet number
                             (get_dim_configuration, configure.F90:404-431) */
    int dim_its;          /* max_iteration_dim (configure.F90:437-465), >= 1 */
    int dim_split;        /* 1: each DIM relaxation sweep as two launches (the
                             Laplacian, then the in-place update) instead of
                             one fused out-of-place pass; identical results */
    int march_levels;     /* levels each block of the k-marching sweeps (the
                             staged Jacobi and even-parity SOR sweeps, the
                             fused DIM sweep) covers: 0 sizes the chunks to
                             the device's resident blocks; n >= 1 fixes n
                             (clamped to the column height).  Identical
                             results either way */
    int mg_pair;          /* the multigrid's finest level runs each
                             smoothing pair (both colour sweeps) as one
                             out-of-place pass: 0 where the level has
                             2 M cells or more, 1 wherever its ny is a
                             multiple of 8, -1 never.  Identical results */
} tpls_params;

/* Per-step diagnostics (the values main_ns_hybrid.F90 prints). */
typedef struct {
    double err_phi, err_u, err_v, err_w;
    int ksp_its, ksp_reason; /* PETSc KSPConvergedReason codes */
    double ksp_rnorm;
} tpls_diag;

/* Field identifiers of a case (main_ns_hybrid.F90:416-549 arrays). */
enum {
    TPLS_F_U2 = 0, TPLS_F_V2, TPLS_F_W2, TPLS_F_U3, TPLS_F_V3, TPLS_F_W3,
    TPLS_F_PHI2, TPLS_F_PHI3, TPLS_F_PRES, TPLS_F_PRES_OLD,
    TPLS_F_CONV0_U, TPLS_F_CONV1_U, TPLS_F_CONV2_U, TPLS_F_CONV0_V, TPLS_F_CONV1_V, TPLS_F_CONV2_V,
    TPLS_F_CONV0_W, TPLS_F_CONV1_W, TPLS_F_CONV2_W,
    TPLS_F_CSF_U0, TPLS_F_CSF_U1, TPLS_F_CSF_U2, TPLS_F_CSF_V0, TPLS_F_CSF_V1, TPLS_F_CSF_V2,
    TPLS_F_CSF_W0, TPLS_F_CSF_W1, TPLS_F_CSF_W2,
    TPLS_F_CONV0_PHI, TPLS_F_CONV1_PHI, TPLS_F_CONV2_PHI,
    TPLS_F_FX_CSF, TPLS_F_FY_CSF, TPLS_F_FZ_CSF,
    TPLS_F_VISC2, TPLS_F_VISC2_AUG, TPLS_F_RHS_P,
    TPLS_F_DIFF_U, TPLS_F_DIFF_V, TPLS_F_DIFF_W, TPLS_F_RHS_U, TPLS_F_RHS_V, TPLS_F_RHS_W,
    TPLS_F_U_INLET,
    TPLS_F_CONVPREV_PHI,  /* ConvPrevPhi, the DIM explicit-term history */
    TPLS_NFIELDS
};

typedef struct tpls_case tpls_case;

const char *tpls_last_error(void);

/* ---- device-resident case ---- */

/* Device bytes a case of this size and precision allocates. */
size_t tpls_case_bytes(int maxl, int maxm, int maxn, int precision);
/* Allocate a case on the current device.  Replaces the array allocation of
 * main_ns_hybrid.F90:416-549 and the DMDA/KSP set-up of
 * configure_mpi_petsc.F90:61-212. */
int tpls_case_create(const tpls_params *prm, int maxl, int maxm, int maxn, tpls_case **out);
int tpls_case_destroy(tpls_case *c);
/* Initial state exactly as main_ns_hybrid.F90:412-549 builds it from the five
 * initial_*.dat arrays (u, v, w, pressure, phi in their Fortran shapes):
 * every other array zero, pres_old = 0, u_inlet = u(0,0,:). */
int tpls_case_load(tpls_case *c, const double *u, const double *v, const double *w,
                   const double *pres, const double *phi);
/* Copy a field to / from a host array of its Fortran shape (halo rows of the
 * periodic y direction are returned as the periodic wrap). */
int tpls_case_get_field(tpls_case *c, int field, double *host);
int tpls_case_set_field(tpls_case *c, int field, const double *host);
/* Advance nsteps timesteps (main_ns_hybrid.F90:583-918, 2480-2509). */
int tpls_step(tpls_case *c, int nsteps);
/* Advance several cases (of any sizes and precisions) by nsteps timesteps
 * together: every case runs on its own stream and each phase of the step is
 * enqueued for all cases before the host waits on any, so the GPU overlaps
 * them.  Each case computes exactly what tpls_step would compute for it
 * alone.  A batched counterpart of main_ns_hybrid.F90:583-918 for
 * independent cases; the reference runs one case per job. */
int tpls_batch_step(tpls_case **cases, int ncases, int nsteps);
/* tpls_batch_step with each case's outcome in status[q] (TPLS_OK or the
 * error of that case's step).  A case whose step fails stops there and is
 * marked failed (later steps of it are refused); the other cases complete
 * their steps unperturbed.  Returns TPLS_OK, or the first failure's code
 * with tpls_last_error() naming its case.  A diverged pressure solve is not
 * a failure here: it is reported in the case's tpls_diag.ksp_reason (< 0),
 * as KSPGetConvergedReason reports it (main_ns_hybrid.F90:903-911). */
int tpls_batch_step_status(tpls_case **cases, int ncases, int nsteps, int *status);
/* How tpls_batch_step runs a batch: mode 0 runs each case's step on its own
 * stream; mode 1 (the default) runs the pressure solves grouped: every CG
 * iteration and every multigrid kernel once across the cases, on one stream,
 * in lock-step; mode 2 groups the Helmholtz loops and the solve's set-up as
 * well.  Each case computes exactly what its own step computes in every
 * mode.  Batches with a case under kernel timing run per case. */
int tpls_batch_set_grouped(int mode);
/* Diagnostics of the last step. */
int tpls_case_diag(tpls_case *c, tpls_diag *out);
/* Change the launch shapes of the Helmholtz sweeps (the tile_rows,
 * tile_rows_odd, prefetch_odd, tile_rows_rb, tile_rows_rb_odd, tile_rows_jst
 * and sor_fused parameters) of an existing case; used by the autotuner.
 * Results do not depend on them. */
int tpls_case_set_tiling(tpls_case *c, int rows, int rows_odd, int prefetch_odd, int rows_rb, int rows_rb_odd,
                         int rows_jst, int sor_fused);
/* Liquid volume V = sum over the interior cells of H_eps(phi2) dx dy dz,
 * accumulated in FP64, with the smoothed Heaviside of get_viscosity
 * (momentum_allflux.f90:499-512); of phi2 - 0.5 with idm_dim, as the
 * viscosity is (main_ns_hybrid.F90:650-651).  The reference has no mass diagnostic; this
 * one measures the volume drift of a run against another (e.g. FP32 storage
 * against FP64). */
int tpls_case_volume(tpls_case *c, double *volume);
/* Wait for all work of the case. */
int tpls_case_sync(tpls_case *c);
/* Kernel timing: while a kernel class is selected, every launch of that
 * class is bracketed by HIP events on the case's stream (the stream it runs
 * on); tpls_case_kernel_time waits for the case and returns the summed
 * milliseconds and launch count since the selection.  Launches of the CG
 * update that return at once after convergence are not counted. */
enum {
    TPLS_TIME_OFF = 0,
    TPLS_TIME_CG_X = 1,      /* CG update x += alpha p, r -= alpha A p */
    TPLS_TIME_JACOBI = 2,    /* Jacobi sweep (do_jacobi_u/v/w) */
    TPLS_TIME_SOR_ODD = 3,   /* SOR half sweep, odd iteration_time (k-ascending columns) */
    TPLS_TIME_SOR_EVEN = 4,  /* SOR half sweep, even iteration_time (3-D colouring) */
    TPLS_TIME_ADVECT = 5     /* WENO5 level-set advection (do_advect_phi) */
};
int tpls_case_time_kernels(tpls_case *c, int kernel_class);
int tpls_case_kernel_time(tpls_case *c, double *ms, long *launches);

/* ---- per-routine drop-in entry points (host arrays) ---- */

/* weno.f90:18  WENOAppoxConvection(phi,u,v,w,sx,ex,sy,ey,l,n,dfx,dfy,dfz) */
int tpls_wenoappoxconvection(const double *phi, const double *u, const double *v, const double *w,
                             int sx, int ex, int sy, int ey, int l, int n, double *dfx, double *dfy,
                             double *dfz);
/* advect_phi.f90:19  do_advect_phi (MPI arguments dropped) */
int tpls_do_advect_phi(double *phi2, double *u2, double *v2, double *w2, double *conv2_phi, int sx,
                       int ex, int sy, int ey, int maxl, int maxn, int ex_max, double dx, double dy,
                       double dz);
/* two_phase_levelset.f90:20  do_tpls (MPI arguments dropped) */
int tpls_do_tpls(double *phi2, double *phi3, double *conv0_phi, double *conv1_phi,
                 double *conv2_phi, int sx, int ex, int sy, int ey, int maxn, int ex_max,
                 int max_iteration_levelset, double height, double *err2, double dx, double dz,
                 double dt);
/* cahn_hilliard_solver.f90:41  dim(maxl,maxn,dx,dy,dz,dt,epn,Pe,phi,phiOld,conv2_phi,
 * ConvPrevPhi,phiNew,sx,ex,sy,ey,sz_p,ez_p,height,...,max_iteration_dim) with the
 * MPI arguments (sz_p = 0, ez_p = maxn, strides, neighbours, communicator)
 * dropped.  phi and phiOld are the phi2+0.5 / phi_old+0.5 arrays the caller
 * passes (main_ns_hybrid.F90:608).  phiNew: initial iterate in, last
 * iterate out (boundary planes included); ConvPrevPhi updated in place.  res (may be NULL): the largest
 * |change| of the last sweep, which the reference computes and discards. */
int tpls_dim(int maxl, int maxn, double dx, double dy, double dz, double dt, double epn, double Pe,
             const double *phi, const double *phiOld, const double *conv2_phi, double *ConvPrevPhi,
             double *phiNew, int sx, int ex, int sy, int ey, double height, int max_iteration_dim,
             double *res);
/* levelset.f90:17  get_levelset_status(phi,D_mx,istatus,sign_mx,sx,ex,sy,ey,n,h) */
int tpls_get_levelset_status(const double *phi, double *D_mx, int *istatus, double *sign_mx, int sx,
                             int ex, int sy, int ey, int n, double h);
/* levelset.f90:97  do_levelset_iteration(...,error_val) */
int tpls_do_levelset_iteration(const double *phi, double *phi_reinit, const double *D_mx,
                               const int *istatus, const double *sign_mx, int sx, int ex, int sy,
                               int ey, int n, int ex_max, double h, double height,
                               double *error_val);
/* levelset.f90:241  get_csf(fx,fy,fz,fi,sx,ex,sy,ey,n,ex_max,dx,dy,dz,scap,smooth_width) */
int tpls_get_csf(double *fx, double *fy, double *fz, const double *fi, int sx, int ex, int sy,
                 int ey, int n, int ex_max, double dx, double dy, double dz, double scap,
                 double smooth_width);
/* momentum_allflux.f90:22  get_diffusion_all */
int tpls_get_diffusion_all(double *diff_u, double *diff_v, double *diff_w, const double *viscosity,
                           const double *u2, const double *v2, const double *w2, int sx, int sy,
                           int ex, int ey, int maxn, int ex_max, double dx, double dy, double dz);
/* momentum_allflux.f90:224  get_conv_all */
int tpls_get_conv_all(double *conv_u2, double *conv_v2, double *conv_w2, const double *viscosity,
                      const double *u2, const double *v2, const double *w2, int sx, int sy, int ex,
                      int ey, int maxn, int ex_max, double dx, double dy, double dz);
/* momentum_allflux.f90:474  get_viscosity */
int tpls_get_viscosity(double *viscosity, const double *phi, int ex, int ey, int sx, int sy,
                       int maxn, double smooth_width, double Re, double mu_minus, double mu_plus);
/* jacobi_iteration_allflux.f90:21 / 102 / 195  do_jacobi_u / _v / _w */
int tpls_do_jacobi_u(double *u3, const double *u3_old, const double *RHS, const double *viscosity,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx, int sy,
                     int maxn, int ex_max);
int tpls_do_jacobi_v(double *v3, const double *v3_old, const double *RHS, const double *viscosity,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx, int sy,
                     int maxn, int ex_max);
int tpls_do_jacobi_w(double *w3, const double *w3_old, const double *RHS, const double *viscosity,
                     double dx, double dy, double dz, double dt, int ex, int ey, int sx, int sy,
                     int maxn, int ex_max);
/* jacobi_iteration_allflux.f90:285  get_difference(ua,ub,ex,ey,ez,sx,sy,sz,diff) */
int tpls_get_difference(const double *ua, const double *ub, int ex, int ey, int ez, int sx, int sy,
                        int sz, double *diff);
/* get_difference on arrays the caller already holds in device memory, for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that keep every buffer resident, harnesses that forbid hipMalloc).  No
 * counterpart in the reference: it allocates nothing and copies nothing.
 * ua, ub: device arrays of the bounds (sx-1:ex+1, sy-1:ey+1, sz:ez),
 * unpadded, x fastest; diff: one double of device memory, which receives
 * the sum when `stream` (a hipStream_t; NULL for the null stream) reaches
 * it.  Its only other device memory is a file-scope array of 1024 partial
 * sums (8 KiB) per GPU, held until the process ends; the reduction always
 * runs that many blocks, so no array size needs more.  Calls that overlap on
 * one GPU race on that array.  TPLS_ERR_ARG for host memory or empty bounds,
 * TPLS_ERR_DECOMP unless sx = sy = 1 and sz = 0. */
int tpls_get_difference_noalloc(const double *ua, const double *ub, int ex, int ey, int ez, int sx,
                                int sy, int sz, double *diff, void *stream);
/* sor_iteration_allflux.f90:19 / 152 / 280  do_sor_u / _v / _w (only do_sor_u
 * takes ex_max, as in the reference) */
int tpls_do_sor_u(double *u3, const double *RHS, const double *viscosity, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max,
                  int flg, int iteration_time);
int tpls_do_sor_v(double *v3, const double *RHS, const double *viscosity, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int flg,
                  int iteration_time);
int tpls_do_sor_w(double *w3, const double *RHS, const double *viscosity, double dx, double dy,
                  double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int flg,
                  int iteration_time);
/* pressure.f90:16  get_source_pres(rhs,u3,v3,w3,ex,ey,sx,sy,maxn,dx,dy,dz,dt) */
int tpls_get_source_pres(double *rhs, const double *u3, const double *v3, const double *w3, int ex,
                         int ey, int sx, int sy, int maxn, double dx, double dy, double dz,
                         double dt);
/* pressure.f90:59  get_uvw_pres(u3,v3,w3,p,ex,ey,sx,sy,maxn,ex_max,dx,dy,dz,dt,u_inlet) */
int tpls_get_uvw_pres(double *u3, double *v3, double *w3, const double *p, int ex, int ey, int sx,
                      int sy, int maxn, int ex_max, double dx, double dy, double dz, double dt,
                      const double *u_inlet);
/* The pressure solve main_ns_hybrid.F90:896-913 runs through PETSc:
 * KSPSolve with the callbacks copy_pressure_in_to_petsc (initial guess = pres),
 * compute_rhs (RHS_p, u3(0,:,:), u_inlet) and compute_matrix
 * (pressure_solver.F90:205-441), then KSPGetIterationNumber,
 * KSPGetResidualNorm, KSPGetConvergedReason and copy_pressure_out_of_petsc
 * (pressure_solver.F90:88-145, 444-486).  This entry point is that whole
 * sequence: pres holds the initial guess on entry and the solution with its
 * ghost planes on return. */
int tpls_pressure_solve(double *pres, const double *RHS_p, const double *u3,
                        const double *u_inlet, int maxl, int maxm, int maxn, double dx, double dy,
                        double dz, double dt, double rtol, double abstol, double dtol, int maxits,
                        int *its, double *rnorm, int *reason);

/* ---- the halo exchanges of one rank (mpi.f90) ----
 * mpi.f90:119-170 exchange2d, :215-270 exchange2d_aug1, :294-369
 * exchange2d_aug2 with a single rank whose y neighbours are itself and whose
 * x neighbours are MPI_PROC_NULL: the y halo rows take the opposite interior
 * rows over x = sx-1..ex+1 of every z-plane.  kind 0: exchange2d on
 * (sx-1:ex+1, sy-1:ey+1, sz:ez), rows sy-1 <- ey and ey+1 <- sy; kind 1:
 * exchange2d_aug1 on (sx-2:ex+2, sy-2:ey+2, sz:ez), the same rows; kind 2:
 * exchange2d_aug2 on that shape, rows sy-2 <- ey-1 and ey+2 <- sy+1.  An
 * array attached to the mirror whose device copy holds data the host has not
 * seen is exchanged on the device.  TPLS_ERR_DECOMP unless sx = sy = 1. */
int tpls_exchange2d(double *uu, int ex, int ey, int ez, int sx, int sy, int sz, int kind);

/* ---- device mirrors of host arrays for the entry points above ----
 * The reference's procedures run on host arrays that main_ns_hybrid.F90
 * allocates once (main_ns_hybrid.F90:416-549) and reuses every step.  An
 * attached host array keeps a device copy that every per-routine entry point
 * (and tpls_exchange2d) uses in place of a copy over the host link.  Host
 * pages are protected so that data moves only where needed: a host write
 * marks its pages for upload before the next device use; a device write
 * makes the host pages fetch on their first host access.  So the caller's
 * code reads and writes the arrays as before.
 * Rules: an attached array keeps its address and stays allocated until it is
 * detached (detach before deallocate); system calls that write into it
 * (read(2), Fortran unformatted READ into the array) need it detached; the
 * partial pages at either end of an array are copied on every device use.
 * tpls_mirror_flush makes every host copy current (before output, for
 * instance).  tpls_mirror_counters: out[0] bytes uploaded, out[1] bytes
 * downloaded, out[2] protection faults resolved, out[3] uploads, out[4]
 * downloads, out[5] arrays attached.  Attaching and detaching are not
 * meant to race with entry points on other threads. */
int tpls_mirror_attach(void *host, long long bytes);
int tpls_mirror_detach(void *host);
int tpls_mirror_detach_all(void);
int tpls_mirror_flush(void);
int tpls_mirror_set_granule(long long bytes);
int tpls_mirror_counters(long long *out);

#ifdef __cplusplus
}
#endif
#endif
