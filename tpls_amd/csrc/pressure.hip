// Pressure: divergence and projection (pressure.f90:16-134) and the Poisson
// solve that PETSc KSP performs in the reference (pressure_solver.F90:205-486,
// configure_mpi_petsc.F90:143-183, main_ns_hybrid.F90:896-913).
//
// The solve is the conjugate-gradient method of PETSc 3.6 KSPCG with no
// preconditioner and the constant null space attached (the null space is
// removed from the right-hand side and from every preconditioned residual
// z = r - mean(r)), run on the symmetric operator left after eliminating the
// two-term Neumann ghost rows of compute_matrix:
//   p(0) = p(1) + dy*b_in,  p(maxl) = p(maxl-1),  p(k=0) = p(1),
//   p(k=maxn) = p(maxn-1),
// with the inlet flux b_in moved to the right-hand side.  Stopping test:
// |z| <= max(rtol*|b|, abstol) (KSPConvergedDefault with a nonzero initial
// guess); divergence if |z| >= dtol*|b|.
//
// An iteration is two streaming kernels, each followed by a one-block kernel
// that finishes its reduction and the scalar recurrences:
//   cg_p:  p <- z + beta/beta_old * p (double-buffered; each plane of the new
//          p is shared through LDS for the stencil), w = A p kept in
//          registers, partials of p.w;
//   cg_pf: alpha = beta / p.w;
//   cg_x:  x <- x + alpha p, r <- r - alpha A p (A p recomputed from p),
//          partials of sum r and sum r^2;
//   cg_xf: mean(r), beta = |z|^2, |z| and the convergence test, which raises
//          st->done.
// Every CG kernel returns at once when st->done is set, so the host enqueues
// iterations ahead of the convergence test.  HBM traffic per iteration:
// r, p_old read and p written (24 B) + p, x, r read and x, r written (40 B)
// = 64 B per unknown.
//
// The streaming kernels walk tiles of 64 x 4 columns x 4 levels (one
// wavefront spans a 512-byte row segment; the 4-level march reuses the z
// neighbours in registers), grid-stride over at most CG_BLOCKS blocks.
#include "reduce.hpp"
#include "tiles.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

// ---- get_source_pres (pressure.f90:38-50): rhs_p = div(u3)/dt ----
// Divergence of the provisional velocity (pressure.f90:42-43), evaluated in
// double precision for every storage type.
template <typename T>
__device__ __forceinline__ double div_at(const FieldView<T>& u3, const FieldView<T>& v3, const FieldView<T>& w3, int i,
                                         int j, int jm, int k, double dx, double dy, double dz)
{
    return ((double(u3(i, jm, k - 1)) - double(u3(i - 1, jm, k - 1))) / dx) +
           ((double(v3(i, j, k - 1)) - double(v3(i, jm, k - 1))) / dy) +
           ((double(w3(i, jm, k)) - double(w3(i, jm, k - 1))) / dz);
}

template <typename T>
__global__ void __launch_bounds__(BX* BY) k_source_pres(Grid g, double dx, double dy, double dz, double dt,
                                                        FieldView<T> u3, FieldView<T> v3, FieldView<T> w3,
                                                        FieldView<double> rhs)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    rhs(i, j, k) = div_at(u3, v3, w3, i, j, wrapj(j - 1, g.P), k, dx, dy, dz) / dt;
}

// ---- get_uvw_pres (pressure.f90:89-131): projection on i in 0..maxl ----
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_uvw_pres(Grid g, T dtdx, T dtdy, T dtdz, FieldView<T> p, const T* u_inlet,
                                                     FieldView<T> u, FieldView<T> v, FieldView<T> w)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    const int L = g.maxl;
    if (i > L || j > g.P) return;
    const int jp = wrapj(j + 1, g.P);
    if (k <= g.maxn - 2) {
        if (i == 0)
            u(i, j, k) = u_inlet[k];
        else if (i == L)
            u(i, j, k) = T(0);
        else
            u(i, j, k) = u(i, j, k) - dtdx * (p(i + 1, jp, k + 1) - p(i, jp, k + 1));
        v(i, j, k) = v(i, j, k) - dtdy * (p(i, jp, k + 1) - p(i, j, k + 1));
    }
    w(i, j, k) = w(i, j, k) - dtdz * (p(i, jp, k + 1) - p(i, jp, k));
}

// ---- CG ----
struct Op {
    int nx, ny, nz;           // interior extents: maxl-1, maxm-1, maxn-1
    double cx, cy, cz, diag;  // pressure_solver.F90:394-421 (TPLS-x uses 1/dy^2)
};

static Op make_op(const Grid& g, const Phys& ph)
{
    Op o;
    o.nx = g.maxl - 1;
    o.ny = g.P;
    o.nz = g.maxn - 1;
    o.cz = 1.0 / (ph.dz * ph.dz);
    o.cx = 1.0 / (ph.dy * ph.dy);
    o.cy = 1.0 / (ph.dx * ph.dx);
    o.diag = 2.0 * (1.0 / (ph.dx * ph.dx) + 1.0 / (ph.dy * ph.dy) + 1.0 / (ph.dz * ph.dz));
    return o;
}

__device__ __forceinline__ double apply_op(const Op& o, double c, double xm, double xp, double ym, double yp, double zm,
                                           double zp)
{
    return o.diag * c - o.cz * zm - o.cx * xm - o.cy * ym - o.cy * yp - o.cx * xp - o.cz * zp;
}

long cg_num_blocks(const Grid& g) { return cg_tiles(g).nblocks; }

// b = -RHS_p with the inlet Neumann flux folded in (pressure_solver.F90:272-286),
// written into the r buffer; partials of sum b.  RHS_p is either read
// (use_rhs) or formed on the fly as div(u3)/dt.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_cg_setup_b(Grid g, Tiles tl, double dx, double dy, double dz, double dt,
                                                       FieldView<T> u3, FieldView<T> v3, FieldView<T> w3,
                                                       const T* u_inlet, FieldView<double> rhs_p, int use_rhs,
                                                       FieldView<double> b, double* part)
{
    const int nx = g.maxl - 1, nz = g.maxn - 1;
    double acc[1] = {0.0};
    for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
        int i, j, k0;
        tile_origin(tl, t, i, j, k0);
        if (i > nx || j > g.P) continue;
        const int jm = wrapj(j - 1, g.P);
        for (int k = k0; k < k0 + CG_KC && k <= nz; ++k) {
            double val = use_rhs ? -rhs_p(i, j, k) : -(div_at(u3, v3, w3, i, j, jm, k, dx, dy, dz) / dt);
            if (i == 1) {
                double b_in = (dx / dt) * (double(u_inlet[k - 1]) - double(u3(0, jm, k - 1))) / dy;
                val = val + b_in / dy;
            }
            b(i, j, k) = val;
            acc[0] += val;
        }
    }
    block_partials<1>(acc, part, gridDim.x, blockIdx.x);
}

// Reset the device state of a solve (a kernel rather than a host copy so that
// the set-up can be captured in a graph).
__global__ void k_cg_init(CgState* st, double rtol, double abstol, double dtol, int maxits)
{
    if (threadIdx.x != 0) return;
    CgState s{};
    s.rtol = rtol;
    s.abstol = abstol;
    s.dtol = dtol;
    s.maxits = maxits;
    *st = s;
}

__global__ void __launch_bounds__(256) k_cg_setup_bf(const double* part, int nb, double N, CgState* st)
{
    double tot[1];
    finish_sums<1>(part, nb, tot);
    if (tid256() == 0) st->mb = tot[0] / N;
}

// b <- b - mean(b); r = b - A x; partials of |b|^2, sum r, sum r^2.
__global__ void __launch_bounds__(BX* BY) k_cg_setup_r(Op o, Tiles tl, FieldView<double> x, FieldView<double> r,
                                                       double* part, const CgState* st)
{
    const double mb = st->mb;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
        int i, j, k0;
        tile_origin(tl, t, i, j, k0);
        if (i > o.nx || j > o.ny) continue;
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        for (int k = k0; k < k0 + CG_KC && k <= o.nz; ++k) {
            double c = x(i, j, k);
            double ax = apply_op(o, c, i > 1 ? x(i - 1, j, k) : c, i < o.nx ? x(i + 1, j, k) : c, x(i, jm, k),
                                 x(i, jp, k), k > 1 ? x(i, j, k - 1) : c, k < o.nz ? x(i, j, k + 1) : c);
            double bb = r(i, j, k) - mb;
            double rr = bb - ax;
            r(i, j, k) = rr;
            acc[0] += bb * bb;
            acc[1] += rr;
            acc[2] += rr * rr;
        }
    }
    block_partials<3>(acc, part, gridDim.x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_cg_setup_rf(const double* part, int nb, double N, CgState* st)
{
    double tot[3];
    finish_sums<3>(part, nb, tot);
    if (tid256() != 0) return;
    double bnorm = sqrt(tot[0]);
    double m = tot[1] / N;
    double beta = tot[2] - tot[1] * m;
    // a NaN in b or in the initial guess reaches beta and must reach dp
    double dp = sqrt(beta > 0.0 || isnan(beta) ? beta : 0.0);
    double rnorm0 = bnorm > 0.0 ? bnorm : dp;
    double ttol = st->rtol * rnorm0 > st->abstol ? st->rtol * rnorm0 : st->abstol;
    st->bnorm = bnorm;
    st->m = m;
    st->sr = tot[1];
    st->beta = beta;
    st->betaold = 0.0;
    st->dp = dp;
    st->rnorm0 = rnorm0;
    st->ttol = ttol;
    st->its = 0;
    st->i = 0;
    st->reason = 0;
    st->done = 0;
    if (isnan(dp) || isnan(bnorm)) {
        st->reason = -9;
        st->done = 1;
    } else if (dp <= ttol) {
        st->reason = dp < st->abstol ? 3 : 2;
        st->done = 1;
    }
}

// p <- z + (beta/beta_old) p  (p <- z on the first iteration), w = A p, p.w.
// Each plane of the new p over the block's 64 x 4 columns is formed once and
// shared through LDS; the x/y neighbours outside the block are formed by the
// edge lanes / rows.  FIRST is the first iteration (p = z exactly).
template <bool FIRST, typename Z>
__device__ __forceinline__ void cg_p_tiles(const Op& o, const Tiles& tl, const FieldView<Z>& r,
                                           const FieldView<double>& po, const FieldView<double>& pn, double m,
                                           double bt, double& acc)
{
    __shared__ double sp[BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto P = [&](int ii, int jj, int kk) -> double {
        double z = double(r(ii, jj, kk)) - m;
        return FIRST ? z : z + bt * po(ii, jj, kk);
    };
    for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
        int i, j, k0;
        tile_origin(tl, t, i, j, k0);
        const bool act = i <= o.nx && j <= o.ny;
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        double pc = 0.0, pm = 0.0;
        if (act) {
            pc = P(i, j, k0);
            pm = k0 > 1 ? P(i, j, k0 - 1) : pc;
        }
        for (int q = 0; q < CG_KC; ++q) {
            const int k = k0 + q;
            if (k > o.nz) break;  // uniform over the block
            double pp = 0.0;
            if (act) pp = k < o.nz ? P(i, j, k + 1) : pc;
            __syncthreads();
            sp[ty][tx] = pc;
            __syncthreads();
            if (act) {
                double xm = i == 1 ? pc : (tx > 0 ? sp[ty][tx - 1] : P(i - 1, j, k));
                double xp = i == o.nx ? pc : (tx < BX - 1 ? sp[ty][tx + 1] : P(i + 1, j, k));
                double ym = ty > 0 ? sp[ty - 1][tx] : P(i, jm, k);
                double yp = (ty < BY - 1 && j < o.ny) ? sp[ty + 1][tx] : P(i, jp, k);
                double w = apply_op(o, pc, xm, xp, ym, yp, k > 1 ? pm : pc, pp);
                pn(i, j, k) = pc;
                acc += pc * w;
            }
            pm = pc;
            pc = pp;
        }
    }
}

// r: the preconditioned residual z (FP64, or FP32 from the FP32 V-cycle), or r
template <typename Z>
__global__ void __launch_bounds__(BX* BY) k_cg_p(Op o, Tiles tl, FieldView<Z> r, FieldView<double> po,
                                                 FieldView<double> pn, int use_mz, double* part, const CgState* st)
{
    if (st->done) return;
    const double m = use_mz ? st->mz : st->m;
    double acc[1] = {0.0};
    if (st->i == 0)
        cg_p_tiles<true>(o, tl, r, po, pn, m, 0.0, acc[0]);
    else
        cg_p_tiles<false>(o, tl, r, po, pn, m, st->beta / st->betaold, acc[0]);
    block_partials<1>(acc, part, gridDim.x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_cg_pf(const double* part, int nb, CgState* st)
{
    if (st->done) return;
    double tot[1];
    finish_sums<1>(part, nb, tot);
    if (tid256() != 0) return;
    double dpi = tot[0];
    st->dpi = dpi;
    if (!(dpi > 0.0)) {
        st->reason = isnan(dpi) ? -9 : -10;
        st->its = st->i + 1;
        st->done = 1;
    } else {
        st->alpha = st->beta / dpi;
    }
}

// x <- x + alpha p, r <- r - alpha A p, partials of sum r, sum r^2.
__global__ void __launch_bounds__(BX* BY) k_cg_x(Op o, Tiles tl, FieldView<double> p, FieldView<double> x,
                                                 FieldView<double> r, double* part, const CgState* st)
{
    if (st->done) return;
    const double a = st->alpha;
    double acc[2] = {0.0, 0.0};
    for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
        int i, j, k0;
        tile_origin(tl, t, i, j, k0);
        if (i > o.nx || j > o.ny) continue;
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        double pc = p(i, j, k0);
        double pm = k0 > 1 ? p(i, j, k0 - 1) : pc;
#pragma unroll
        for (int q = 0; q < CG_KC; ++q) {
            const int k = k0 + q;
            if (k > o.nz) break;
            double pp = k < o.nz ? p(i, j, k + 1) : pc;
            double w = apply_op(o, pc, i > 1 ? p(i - 1, j, k) : pc, i < o.nx ? p(i + 1, j, k) : pc, p(i, jm, k),
                                p(i, jp, k), k > 1 ? pm : pc, pp);
            x(i, j, k) = x(i, j, k) + a * pc;
            double rn = r(i, j, k) - a * w;
            r(i, j, k) = rn;
            acc[0] += rn;
            acc[1] += rn * rn;
            pm = pc;
            pc = pp;
        }
    }
    block_partials<2>(acc, part, gridDim.x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_cg_xf(const double* part, int nb, double N, CgState* st)
{
    if (st->done) return;
    double tot[2];
    finish_sums<2>(part, nb, tot);
    if (tid256() != 0) return;
    double m = tot[0] / N;
    double beta = tot[1] - tot[0] * m;
    double dp = sqrt(beta > 0.0 ? beta : 0.0);
    st->its = st->i + 1;
    st->m = m;
    st->sr = tot[0];
    st->betaold = st->beta;
    st->beta = beta;
    st->dp = dp;
    int reason = 0;
    if (isnan(dp) || isnan(beta))
        reason = -9;
    else if (dp <= st->ttol)
        reason = dp < st->abstol ? 3 : 2;
    else if (dp >= st->dtol * st->rnorm0)
        reason = -4;
    if (!reason) {
        st->i = st->i + 1;
        if (st->i >= st->maxits) reason = -3;
    }
    if (reason) {
        st->reason = reason;
        st->done = 1;
    }
}

// Ghost planes of the solution (copy_pressure_out_of_petsc,
// pressure_solver.F90:444-486, with the eliminated Neumann rows), written
// with the interior into pres.  One thread per cell, i in 0..maxl,
// k in 0..maxn; the z ghost planes copy k = 1 and k = maxn-1.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_pres_out(Grid g, double dx, double dy, double dt, FieldView<double> x,
                                                     FieldView<T> u3, const T* u_inlet, FieldView<T> pres)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    const int L = g.maxl, n = g.maxn;
    if (i > L || j > g.P) return;
    const int ks = k == 0 ? 1 : (k == n ? n - 1 : k);
    double v;
    if (i == 0) {
        const int jm = wrapj(j - 1, g.P);
        double b_in = (dx / dt) * (double(u_inlet[ks - 1]) - double(u3(0, jm, ks - 1))) / dy;
        v = x(1, j, ks) + dy * b_in;
    } else if (i == L) {
        v = x(L - 1, j, ks);
    } else {
        v = x(i, j, ks);
    }
    pres(i, j, k) = T(v);
}

// Sum of squared differences over i in 1..maxl-1, j in 1..P, k in 0..nz-1
// (levelset.f90:218-228 and get_difference, jacobi_iteration_allflux.f90:297-308).
template <typename T>
__global__ void __launch_bounds__(256) k_sqdiff(Grid g, int nz, FieldView<T> a, FieldView<T> b, double* part)
{
    const int nx = g.maxl - 1;
    const long n = (long)nx * g.P * nz;
    double acc[1] = {0.0};
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        int i = int(q % nx) + 1;
        long rr = q / nx;
        int j = int(rr % g.P) + 1;
        int k = int(rr / g.P);
        double d = double(a(i, j, k)) - double(b(i, j, k));
        acc[0] += d * d;
    }
    block_partials<1>(acc, part, gridDim.x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_final1(const double* part, int nb, double* out)
{
    double tot[1];
    finish_sums<1>(part, nb, tot);
    if (tid256() == 0) out[0] = tot[0];
}

// Liquid volume V = sum over interior cells of H_eps(phi) dx dy dz in FP64,
// with the smoothed Heaviside of the viscosity (momentum_allflux.f90:499-512).
// The reference has no mass diagnostic; this one measures volume drift.
template <typename T>
__global__ void __launch_bounds__(256) k_volume(Grid g, double sw, double dv, double shift, FieldView<T> phi,
                                                double* part)
{
    const int nx = g.maxl - 1, nz = g.maxn - 1;
    const long n = (long)nx * g.P * nz;
    double acc[1] = {0.0};
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int i = int(q % nx) + 1;
        const long rr = q / nx;
        const int j = int(rr % g.P) + 1;
        const int k = int(rr / g.P) + 1;
        const double pv = double(phi(i, j, k)) - shift;
        double H;
        if (pv < -sw)
            H = 0.0;
        else if (pv > sw)
            H = 1.0;
        else
            H = 0.5 + pv / (2.0 * sw) + (1.0 / (2.0 * TPLS_PI_SP)) * sin(TPLS_PI_SP * pv / sw);
        acc[0] += H * dv;
    }
    block_partials<1>(acc, part, gridDim.x, blockIdx.x);
}

// ---- launchers ----
template <typename T>
hipError_t launch_source_pres(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> u3, FieldView<T> v3,
                              FieldView<T> w3, FieldView<double> rhs_p)
{
    dim3 grd((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn - 1);
    hipLaunchKernelGGL(k_source_pres<T>, grd, dim3(BX, BY), 0, s, g, ph.dx, ph.dy, ph.dz, ph.dt, u3, v3, w3, rhs_p);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_uvw_pres(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> p, const T* u_inlet,
                           FieldView<T> u3, FieldView<T> v3, FieldView<T> w3)
{
    dim3 grd((g.maxl + 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn);
    hipLaunchKernelGGL(k_uvw_pres<T>, grd, dim3(BX, BY), 0, s, g, T(ph.dt / ph.dx), T(ph.dt / ph.dy), T(ph.dt / ph.dz),
                       p, u_inlet, u3, v3, w3);
    return hipGetLastError();
}

static double n_unknowns(const Grid& g) { return double(g.maxl - 1) * g.P * (g.maxn - 1); }

hipError_t launch_cg_init(hipStream_t s, CgState* st, double rtol, double abstol, double dtol, int maxits)
{
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(64), 0, s, st, rtol, abstol, dtol, maxits);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_setup_b(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> u3, FieldView<T> v3,
                             FieldView<T> w3, const T* u_inlet, FieldView<double> rhs_p, int use_rhs,
                             FieldView<double> b, Reduce red, CgState* st)
{
    Tiles tl = cg_tiles(g);
    hipLaunchKernelGGL(k_cg_setup_b<T>, dim3(tl.nblocks), dim3(BX, BY), 0, s, g, tl, ph.dx, ph.dy, ph.dz, ph.dt, u3, v3,
                       w3, u_inlet, rhs_p, use_rhs, b, red.part);
    hipLaunchKernelGGL(k_cg_setup_bf, dim3(1), dim3(256), 0, s, red.part, tl.nblocks, n_unknowns(g), st);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_setup_r(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x, FieldView<double> r,
                             Reduce red, CgState* st)
{
    Tiles tl = cg_tiles(g);
    hipLaunchKernelGGL(k_cg_setup_r, dim3(tl.nblocks), dim3(BX, BY), 0, s, make_op(g, ph), tl, x, r, red.part, st);
    hipLaunchKernelGGL(k_cg_setup_rf, dim3(1), dim3(256), 0, s, red.part, tl.nblocks, n_unknowns(g), st);
    return hipGetLastError();
}

template <typename Z>
hipError_t launch_cg_p(hipStream_t s, const Grid& g, const Phys& ph, FieldView<Z> z, FieldView<double> p_old,
                       FieldView<double> p_new, int use_mz, Reduce red, CgState* st)
{
    Tiles tl = cg_tiles(g);
    hipLaunchKernelGGL(k_cg_p<Z>, dim3(tl.nblocks), dim3(BX, BY), 0, s, make_op(g, ph), tl, z, p_old, p_new, use_mz,
                       red.part, st);
    hipLaunchKernelGGL(k_cg_pf, dim3(1), dim3(256), 0, s, red.part, tl.nblocks, st);
    return hipGetLastError();
}
template hipError_t launch_cg_p<double>(hipStream_t, const Grid&, const Phys&, FieldView<double>, FieldView<double>,
                                        FieldView<double>, int, Reduce, CgState*);
template hipError_t launch_cg_p<float>(hipStream_t, const Grid&, const Phys&, FieldView<float>, FieldView<double>,
                                       FieldView<double>, int, Reduce, CgState*);

hipError_t launch_cg_x(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> p, FieldView<double> x,
                       FieldView<double> r, Reduce red, CgState* st)
{
    Tiles tl = cg_tiles(g);
    hipLaunchKernelGGL(k_cg_x, dim3(tl.nblocks), dim3(BX, BY), 0, s, make_op(g, ph), tl, p, x, r, red.part, st);
    hipLaunchKernelGGL(k_cg_xf, dim3(1), dim3(256), 0, s, red.part, tl.nblocks, n_unknowns(g), st);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pres_ghosts(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x, FieldView<T> u3,
                              const T* u_inlet, FieldView<T> pres)
{
    dim3 grd((g.maxl + 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn + 1);
    hipLaunchKernelGGL(k_pres_out<T>, grd, dim3(BX, BY), 0, s, g, ph.dx, ph.dy, ph.dt, x, u3, u_inlet, pres);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_levelset_err(hipStream_t s, const Grid& g, FieldView<T> a, FieldView<T> b, Reduce red)
{
    hipLaunchKernelGGL(k_sqdiff<T>, dim3(REDUCE_BLOCKS), dim3(256), 0, s, g, g.maxn + 1, a, b, red.part);
    hipLaunchKernelGGL(k_final1, dim3(1), dim3(256), 0, s, red.part, REDUCE_BLOCKS, red.out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_difference(hipStream_t s, const Grid& g, int nz, FieldView<T> a, FieldView<T> b, Reduce red)
{
    hipLaunchKernelGGL(k_sqdiff<T>, dim3(REDUCE_BLOCKS), dim3(256), 0, s, g, nz, a, b, red.part);
    hipLaunchKernelGGL(k_final1, dim3(1), dim3(256), 0, s, red.part, REDUCE_BLOCKS, red.out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_volume(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi, Reduce red, double shift)
{
    hipLaunchKernelGGL(k_volume<T>, dim3(REDUCE_BLOCKS), dim3(256), 0, s, g, ph.sw, ph.dx * ph.dy * ph.dz, shift, phi,
                       red.part);
    hipLaunchKernelGGL(k_final1, dim3(1), dim3(256), 0, s, red.part, REDUCE_BLOCKS, red.out);
    return hipGetLastError();
}

// ---- utilities ----
template <typename T>
__global__ void k_fill(T* p, long n, T v)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) p[q] = v;
}
template <typename T>
__global__ void k_copy(T* d, const T* s, long n)
{
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) d[q] = s[q];
}

template <typename T>
hipError_t launch_fill(hipStream_t s, T* p, long n, T v)
{
    long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_copy(hipStream_t s, T* dst, const T* src, long n)
{
    long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_copy<T>, dim3(nb), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

#define TPLS_INST_PR(T)                                                                                             \
    template hipError_t launch_source_pres<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, FieldView<T>,    \
                                              FieldView<T>, FieldView<double>);                                     \
    template hipError_t launch_uvw_pres<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, const T*,           \
                                           FieldView<T>, FieldView<T>, FieldView<T>);                               \
    template hipError_t launch_cg_setup_b<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, FieldView<T>,     \
                                             FieldView<T>, const T*, FieldView<double>, int, FieldView<double>,     \
                                             Reduce, CgState*);                                                     \
    template hipError_t launch_cg_setup_r<T>(hipStream_t, const Grid&, const Phys&, FieldView<double>,              \
                                             FieldView<double>, Reduce, CgState*);                                  \
    template hipError_t launch_pres_ghosts<T>(hipStream_t, const Grid&, const Phys&, FieldView<double>,             \
                                              FieldView<T>, const T*, FieldView<T>);                                \
    template hipError_t launch_levelset_err<T>(hipStream_t, const Grid&, FieldView<T>, FieldView<T>, Reduce);       \
    template hipError_t launch_difference<T>(hipStream_t, const Grid&, int, FieldView<T>, FieldView<T>, Reduce);    \
    template hipError_t launch_volume<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, Reduce, double);      \
    template hipError_t launch_fill<T>(hipStream_t, T*, long, T);                                                   \
    template hipError_t launch_copy<T>(hipStream_t, T*, const T*, long);
TPLS_INST_PR(double)
TPLS_INST_PR(float)
template hipError_t launch_fill<int>(hipStream_t, int*, long, int);

}  // namespace tpls
