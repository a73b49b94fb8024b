// Kernel launchers of the TPLS stencil core (gfx950).  Each launcher enqueues
// on the given stream and returns the launch status.  Index conventions are
// the reference's (see tpls_dev.hpp); every launcher assumes the single-rank
// domain sx = 1, ex = maxl-1, sy = 1, ey = maxm-1.
#pragma once

#include "tpls_dev.hpp"

namespace tpls {

struct Phys {
    double dx, dy, dz, dt, Re, mu_plus, mu_minus, height, sw, scap;
};

// Grid reductions: the producing kernel writes per-block partials into
// `part`; a one-block finishing kernel sums them in a fixed order into `out`
// (deterministic for a fixed launch shape).
struct Reduce {
    double* part;
    double* out;
};

// ---- advection (weno.f90, advect_phi.f90) ----
template <typename T>
hipError_t launch_advect_phi(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi2,
                             FieldView<T> u2, FieldView<T> v2, FieldView<T> w2, FieldView<T> conv2_phi);
template <typename T>
hipError_t launch_weno_faces(hipStream_t s, const Grid& g, FieldView<T> phi, FieldView<T> u,
                             FieldView<T> v, FieldView<T> w, FieldView<T> dfx, FieldView<T> dfy,
                             FieldView<T> dfz);

// ---- level set (two_phase_levelset.f90, levelset.f90) ----
template <typename T>
hipError_t launch_tpls_update(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi2,
                              FieldView<T> c0, FieldView<T> c1, FieldView<T> c2, FieldView<T> phi3);
template <typename T>
hipError_t launch_phi_bc(hipStream_t s, const Grid& g, double dz, double height, FieldView<T> phi,
                         int inlet_outlet);
template <typename T>
hipError_t launch_levelset_status(hipStream_t s, const Grid& g, double h, FieldView<T> phi,
                                  FieldView<T> D, FieldView<int> ist, FieldView<T> sign);
template <typename T>
hipError_t launch_levelset_iter(hipStream_t s, const Grid& g, double h, FieldView<T> phi,
                                FieldView<T> D, FieldView<int> ist, FieldView<T> sign,
                                FieldView<T> pr, FieldView<T> pr_new);
template <typename T>
hipError_t launch_levelset_err(hipStream_t s, const Grid& g, FieldView<T> a, FieldView<T> b,
                               Reduce red);
// The status and iteration with the cell state packed into one int per cell
// (ist, the sign of phi and the iteration's branch); same values, less traffic.
template <typename T>
hipError_t launch_levelset_status_code(hipStream_t s, const Grid& g, double h, FieldView<T> phi,
                                       FieldView<T> D, FieldView<int> code);
template <typename T>
hipError_t launch_levelset_iter_code(hipStream_t s, const Grid& g, double h, FieldView<T> D,
                                     FieldView<int> code, FieldView<T> pr, FieldView<T> pr_new);
template <typename T>
hipError_t launch_csf(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi,
                      FieldView<T> fx, FieldView<T> fy, FieldView<T> fz);
template <typename T>
hipError_t launch_csf_bc(hipStream_t s, const Grid& g, FieldView<T> fx, FieldView<T> fy,
                         FieldView<T> fz);

// ---- diffuse interface (cahn_hilliard_solver.f90) ----
// Scalar factors of dim, in FP64: the explicit part's coefficients
// nmda*tao/2 and epn^2*nmda/2, the sweep's dt*nmda/2/Pe, tao, alfa1 and
// diagonal.
struct DimCoef {
    double epn, epn2, Pe, dt, c_tao, c_epn, c_rel, tao, alfa1, diagonal;
};
void dim_coefficients(double dx, double dy, double dz, double dt, double epn, double Pe, DimCoef* c);
// inlet profile phiNew(0,.,k), k = 0..maxn (maxn+1 values)
void dim_inlet(int maxn, double dz, double height, double epn, double* prof);
// One dim call: phi3 = (last iterate of max_iteration_dim sweeps) - 0.5, with
// ConvPrevPhi updated.  zero: phiNew starts from zero (the first step) instead
// of phi2.  pn, pn2, d2, src: work arrays of the p shape (pn2 only for the
// fused sweep); inlet: maxn+1 values on the device; res: if non-null, the
// last sweep's largest |change|.  fused: each sweep as one out-of-place pass
// (identical results) instead of the Laplacian and the in-place update; ks:
// levels per block of that pass (0: sized to the resident blocks).
template <typename T>
hipError_t launch_dim(hipStream_t s, const Grid& g, const Phys& ph, const DimCoef& c, int its, int zero,
                      FieldView<T> phi2, FieldView<T> conv2, FieldView<T> cprev, FieldView<T> phi3,
                      FieldView<T> pn, FieldView<T> pn2, FieldView<T> d2, FieldView<T> src, const T* inlet,
                      double* res, int fused, int ks = 0);
// The body of dim on given arrays: phi = phi_in + shift, phiOld = phio +
// shift; pn (and pn2 when fused) hold the initial iterate, boundary planes
// included; the last iterate ends in pn, or in pn2 after an odd number of
// fused sweeps.
template <typename T>
hipError_t launch_dim_core(hipStream_t s, const Grid& g, const Phys& ph, const DimCoef& c, int its, double shift,
                           FieldView<T> phi, FieldView<T> phio, FieldView<T> conv2, FieldView<T> cprev,
                           FieldView<T> pn, FieldView<T> pn2, FieldView<T> d2, FieldView<T> src, const T* inlet,
                           double* res, int fused, int ks = 0);

// ---- momentum (momentum_allflux.f90, main_ns_hybrid.F90:629-757) ----
// shift: 0, or 1 for the diffuse interface method, whose viscosity is that
// of phi2 - 0.5 (main_ns_hybrid.F90:650-651, 664-668, 689-693)
template <typename T>
hipError_t launch_viscosity(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi,
                            FieldView<T> mu, int ilo, int ihi, int shift = 0);
// mode bit 1: write conv, bit 2: write diff, bit 4: fused AB3/CN right-hand
// side (needs conv0/1, csf faces and histories)
template <typename T>
hipError_t launch_momentum(hipStream_t s, const Grid& g, const Phys& ph, int comp, int mode,
                           FieldView<T> u2, FieldView<T> v2, FieldView<T> w2, FieldView<T> mu,
                           FieldView<T> conv2, FieldView<T> diff, FieldView<T> f_csf,
                           FieldView<T> conv0, FieldView<T> conv1, FieldView<T> csf0,
                           FieldView<T> csf1, FieldView<T> csf2, FieldView<T> rhs);

// ---- Helmholtz sweeps (jacobi_iteration_allflux.f90, sor_iteration_allflux.f90) ----
// Launch shapes of the sweeps: rows per 64-wide tile of the Jacobi and
// even-parity SOR sweeps and of the odd-parity (column) SOR sweep, and the
// odd sweep's prefetch depth.  Every choice computes the same values.
struct Tiling {
    int rows = 4, rows_odd = 4, depth_odd = 4;
    // both SOR sweeps of an iteration in one out-of-place pass (bit 0: odd
    // iteration_time, bit 1: even; else the two in-place sweeps), with tiles
    // of rb_rows (even parity) / rb_rows_odd rows
    int fused = 1, rb_rows = 8, rb_rows_odd = 8;
    // Jacobi sweep with staged planes and tiles of jst rows (0: the unstaged
    // sweep with `rows`)
    int jst = 4;
    // levels per block of the k-marching staged sweeps (0: sized to the
    // resident blocks by march_chunk)
    int ks = 0;
};
// Tiling::fused bit 2: an even iteration's Jacobi step and first colour sweep
// in one pass (launch_jsor0) where the even sweeps are split and staged
constexpr int FUSED_JSOR = 4;
template <typename T>
hipError_t launch_jacobi(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xold,
                         FieldView<T> rhs, FieldView<T> mu, FieldView<T> x, bool xold_is_zero,
                         const Tiling& tl = Tiling{});
template <typename T>
hipError_t launch_sor(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> x,
                      FieldView<T> rhs, FieldView<T> mu, int flg, int iteration_time,
                      const Tiling& tl = Tiling{});
// both sweeps (flg = 0, then 1) of one SOR iteration, out of place: xin is
// x after the Jacobi step and its boundary conditions, xout receives every
// updated cell
template <typename T>
hipError_t launch_sor_pair(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xin,
                           FieldView<T> xout, FieldView<T> rhs, FieldView<T> mu, int iteration_time,
                           const Tiling& tl = Tiling{});
// launch_sor_pair of u, v and w (xin[q], xout[q], rhs[q]) at an odd
// iteration_time as one launch, which keeps more blocks on each CU than the
// one-component passes (same values).  Where the one-component pass takes
// 8-row tiles: sor_pair3_fits.
inline bool sor_pair3_fits(const Grid& g, const Tiling& tl) { return tl.rb_rows_odd == 8 && g.P % 8 == 0; }
template <typename T>
hipError_t launch_sor_pair3(hipStream_t s, const Grid& g, const Phys& ph, const FieldView<T>* xin,
                            const FieldView<T>* xout, const FieldView<T>* rhs, FieldView<T> mu, const Tiling& tl);
// outer-loop boundary conditions (main_ns_hybrid.F90:764-776, 826-836,
// 864-879; the same statements close do_jacobi_v / do_jacobi_w)
template <typename T>
hipError_t launch_helmholtz_bc(hipStream_t s, const Grid& g, const Phys& ph, int comp,
                               FieldView<T> x, FieldView<T> pres);
// an even iteration_time's Jacobi step from xs, its boundary conditions and
// the first SOR colour sweep (flg 0), into xd in one pass: xd then holds what
// launch_jacobi, launch_helmholtz_bc and launch_sor(flg 0) leave in it.
// Needs a period divisible by 8 (jsor0_fits).
inline bool jsor0_fits(const Grid& g) { return g.P % 8 == 0; }
template <typename T>
hipError_t launch_jsor0(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xs, FieldView<T> xd,
                        FieldView<T> rhs, FieldView<T> mu, FieldView<T> pres, const Tiling& tl = Tiling{});
template <typename T>
hipError_t launch_difference(hipStream_t s, const Grid& g, int nz, FieldView<T> a, FieldView<T> b,
                             Reduce red);
// V = sum of H_eps(phi - shift) dx dy dz over the interior, FP64, into
// red.out[0]; shift 0.5 for the diffuse interface method
template <typename T>
hipError_t launch_volume(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi, Reduce red,
                         double shift);

// ---- pressure (pressure.f90, pressure_solver.F90) ----
template <typename T>
hipError_t launch_source_pres(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> u3,
                              FieldView<T> v3, FieldView<T> w3, FieldView<double> rhs_p);
template <typename T>
hipError_t launch_uvw_pres(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> p,
                           const T* u_inlet, FieldView<T> u3, FieldView<T> v3, FieldView<T> w3);

// Conjugate-gradient state on the device (PETSc 3.6 KSPCG with pc none and a
// constant null space, restated on the reduced symmetric system).
// With the multigrid preconditioner z = M r - mean(M r); sr = sum r and
// mz = mean(M r) feed beta = z.r = sum(M r * r) - mz * sr.
struct CgState {
    double beta, betaold, alpha, m, bnorm, rnorm0, ttol, dp, dpi;
    double rtol, abstol, dtol, mb, sr, mz;
    int its, i, reason, done, maxits, pad;
};
// All CG vectors use the p-grid layout and touch interior cells only.
hipError_t launch_cg_init(hipStream_t s, CgState* st, double rtol, double abstol, double dtol, int maxits);
template <typename T>
hipError_t launch_cg_setup_b(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> u3,
                             FieldView<T> v3, FieldView<T> w3, const T* u_inlet,
                             FieldView<double> rhs_p, int use_rhs, FieldView<double> b, Reduce red,
                             CgState* st);
template <typename T>
hipError_t launch_cg_setup_r(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x,
                             FieldView<double> r, Reduce red2, CgState* st);
// p <- (z - mean) + beta/beta_old p with z = r (mean st->m) or, when
// use_mz, z = M r (mean st->mz)
template <typename Z>
hipError_t launch_cg_p(hipStream_t s, const Grid& g, const Phys& ph, FieldView<Z> z,
                       FieldView<double> p_old, FieldView<double> p_new, int use_mz, Reduce red,
                       CgState* st);
hipError_t launch_cg_x(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> p,
                       FieldView<double> x, FieldView<double> r, Reduce red2, CgState* st);
template <typename T>
hipError_t launch_pres_ghosts(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x,
                              FieldView<T> u3, const T* u_inlet, FieldView<T> pres);

// ---- multigrid preconditioner of the pressure solve ----
// One level of the hierarchy: the reduced operator on nx x ny x nz cells
// (Neumann in x and z, periodic in y) and its vectors, addressed with
// Fortran indices 1..n.  Level 0 is the pressure grid itself.
// One level: x (the smoothed iterate) stored as X, b (its right-hand side) as
// B; the level's arithmetic is done in X.
template <typename X, typename B>
struct MgLevelT {
    int nx, ny, nz;
    double cx, cy, cz, diag;
    FieldView<X> x;
    FieldView<B> b;
};
using MgLevel = MgLevelT<double, double>;
using MgLevel32 = MgLevelT<float, float>;
constexpr int MG_MAXLEV = 16;
constexpr int MG_COARSE_MAX = 2048;  // cells of the coarsest level (one block, LDS)
// L: extents, coefficients and the FP64 vectors.  With fp32 set the cycle
// runs on the FP32 vectors of F instead (level 0 reads its right-hand side,
// the CG residual, from L[0].b in FP64); the coefficients are L's.
struct MgHierarchy {
    int nlev, nu;
    MgLevel L[MG_MAXLEV];
    int fp32;
    MgLevel32 F[MG_MAXLEV];
    // eigenbasis tables of the coarsest level (mg_spectral_tables), or null:
    // the coarsest solve then sweeps
    const double* spec;
    // a second vector of level 0's shape for the cycle's arithmetic (FP64:
    // x2, FP32: x2f), or null: with it each smoothing pair of level 0 runs as
    // one out-of-place pass (where mg_pair_fits)
    FieldView<double> x2{nullptr, 0, 0};
    FieldView<float> x2f{nullptr, 0, 0};
    // the same for level 1 (x21, x21f: level 1's shape), or null
    FieldView<double> x21{nullptr, 0, 0};
    FieldView<float> x21f{nullptr, 0, 0};
};
// Levels whose smoothing pairs run as one pass: whole 8-row tiles in y and
// enough cells (2 M) to fill the device with 64 x 8 tiles.
inline bool mg_pair_fits(int nx, int ny, int nz) { return ny % 8 == 0 && long(nx) * ny * nz >= (1L << 21); }
// The coarsest level's separable operator cx Lx + cy Ly + cz Lz (+ the
// shift diag - 2 (cx + cy + cz)) in its eigenbasis: Qx (nx x nx, row i,
// column = mode), Qy, Qz, then the unit eigenvalues lx, ly, lz.  Returns the
// table length, or 0 when an extent exceeds MG_SPEC_MAX.
constexpr int MG_SPEC_MAX = 128;
int mg_spectral_tables(int nx, int ny, int nz, double* tab);
// Extents and coefficients of every level (vectors left unset); returns the
// number of levels.
int mg_plan(int nx, int ny, int nz, double cx, double cy, double cz, double diag, MgLevel* L);
// z = M r into L[0].x from L[0].b, then partials of z.r and sum z and the
// scalar update of st (mz, beta); every kernel returns at once when st->done.
// g is the level-0 grid (maxl = nx+1, P = ny, maxn = nz+1).
hipError_t launch_mg_vcycle(hipStream_t s, const Grid& g, const MgHierarchy& h, Reduce red, CgState* st);

// ---- utilities ----
template <typename T>
hipError_t launch_fill(hipStream_t s, T* p, long n, T v);
template <typename T>
hipError_t launch_copy(hipStream_t s, T* dst, const T* src, long n);

// ---- y-halo rows of one rank with y-periodic neighbours (mpi.f90) ----
// On a contiguous array of nx * ny * nz doubles (x fastest): for every pair q
// < npairs, row dst[q] of every z-plane takes row src[q], over x in
// [i0, i0 + ni).  Rows are 0-based; no destination row is a source row.
constexpr int MAX_ROW_PAIRS = 6;
struct RowPairs {
    int n;
    int dst[MAX_ROW_PAIRS], src[MAX_ROW_PAIRS];
};
hipError_t launch_copy_rows(hipStream_t s, double* a, long nx, long ny, long nz, int i0, int ni, const RowPairs& rp);

// Launch-shape helpers shared by the .hip files.
constexpr int BX = 64;  // one wavefront along x
constexpr int BY = 4;
constexpr int REDUCE_BLOCKS = 1024;  // grid of the grid-stride reduction kernels

// Number of blocks of the CG kernels (= partials per reduction).
long cg_num_blocks(const Grid& g);

}  // namespace tpls
