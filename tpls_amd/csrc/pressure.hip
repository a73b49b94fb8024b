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
// Each iteration is two kernels, each a single pass with one fused grid
// reduction:
//   cg_p: p <- z + beta/beta_old * p (double-buffered, recomputed at the
//         stencil neighbours), w = A p kept in registers, partial p.w;
//         the last block forms alpha.
//   cg_x: x <- x + alpha p, r <- r - alpha A p (A p recomputed from p),
//         partial sums of r and r^2; the last block forms mean(r), beta, |z|
//         and runs the convergence test, raising st->done.
// Every CG kernel returns at once when st->done is set, so the host can
// enqueue iterations ahead of the convergence test.  HBM traffic per
// iteration: r, p_old read and p written (24 B) + p, x, r read and x, r
// written (40 B) = 64 B per unknown.
#include "reduce.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

// ---- get_source_pres (pressure.f90:38-50): rhs_p = div(u3)/dt ----
template <typename T>
__device__ __forceinline__ double div_at(const FieldView<T>& u3, const FieldView<T>& v3, const FieldView<T>& w3, int i,
                                         int j, int jm, int k, double dx, double dy, double dz);

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

// Launch shape of the CG kernels: 64 x 4 columns per block, each thread
// marching kc consecutive levels.
struct CgShape {
    dim3 grid;
    int kc;
};

static CgShape cg_shape(const Grid& g)
{
    const int nx = g.maxl - 1, nz = g.maxn - 1;
    const int bxy = ((nx + BX - 1) / BX) * ((g.P + BY - 1) / BY);
    int nch = (2048 + bxy - 1) / bxy;
    if (nch > nz) nch = nz;
    if (nch < 1) nch = 1;
    const int kc = (nz + nch - 1) / nch;
    nch = (nz + kc - 1) / kc;
    CgShape s;
    s.grid = dim3((nx + BX - 1) / BX, (g.P + BY - 1) / BY, nch);
    s.kc = kc;
    return s;
}

long cg_num_blocks(const Grid& g)
{
    CgShape sh = cg_shape(g);
    return long(sh.grid.x) * sh.grid.y * sh.grid.z;
}

__device__ __forceinline__ int tid256() { return threadIdx.x + BX * threadIdx.y; }
__device__ __forceinline__ int bid3() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ __forceinline__ int nblk3() { return gridDim.x * gridDim.y * gridDim.z; }

// Reduction of up to 3 values over the grid; returns true in the last block
// with the totals in thread 0's tot[].
__device__ __forceinline__ bool grid_sums(int nval, double* v, const Reduce& red, double* tot)
{
    __shared__ double sm[8];
    double s[3];
    for (int q = 0; q < nval; ++q) s[q] = block_sum256(v[q], sm);
    if (!publish_partials(red, nval, s, nblk3(), bid3())) return false;
    for (int q = 0; q < nval; ++q) tot[q] = sum_partials(red, q, nblk3(), sm);
    return true;
}

// Divergence of the provisional velocity, get_source_pres (pressure.f90:42-43),
// evaluated in double precision for every storage type.
template <typename T>
__device__ __forceinline__ double div_at(const FieldView<T>& u3, const FieldView<T>& v3, const FieldView<T>& w3, int i,
                                         int j, int jm, int k, double dx, double dy, double dz)
{
    return ((double(u3(i, jm, k - 1)) - double(u3(i - 1, jm, k - 1))) / dx) +
           ((double(v3(i, j, k - 1)) - double(v3(i, jm, k - 1))) / dy) +
           ((double(w3(i, jm, k)) - double(w3(i, jm, k - 1))) / dz);
}

// b = -RHS_p with the inlet Neumann flux folded in (pressure_solver.F90:272-286),
// written into the r buffer; the last block stores mean(b).  RHS_p is either
// read (use_rhs) or formed on the fly as div(u3)/dt.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_cg_setup_b(Grid g, double dx, double dy, double dz, double dt, int kc,
                                                       FieldView<T> u3, FieldView<T> v3, FieldView<T> w3,
                                                       const T* u_inlet, FieldView<double> rhs_p, int use_rhs,
                                                       FieldView<double> b, Reduce red, CgState* st)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k0 = blockIdx.z * kc + 1;
    const int nz = g.maxn - 1;
    double acc = 0.0;
    if (i <= g.maxl - 1 && j <= g.P) {
        const int jm = wrapj(j - 1, g.P);
        for (int k = k0; k < k0 + kc && k <= nz; ++k) {
            double val = use_rhs ? -rhs_p(i, j, k) : -(div_at(u3, v3, w3, i, j, jm, k, dx, dy, dz) / dt);
            if (i == 1) {
                double b_in = (dx / dt) * (double(u_inlet[k - 1]) - double(u3(0, jm, k - 1))) / dy;
                val = val + b_in / dy;
            }
            b(i, j, k) = val;
            acc += val;
        }
    }
    double tot[1];
    if (grid_sums(1, &acc, red, tot) && tid256() == 0) {
        const double N = double(g.maxl - 1) * g.P * (g.maxn - 1);
        st->mb = tot[0] / N;
        *red.counter = 0u;
    }
}

// b <- b - mean(b); r = b - A x; sums |b|^2, sum r, sum r^2; initial test.
__global__ void __launch_bounds__(BX* BY) k_cg_setup_r(Grid g, Op o, int kc, FieldView<double> x, FieldView<double> r,
                                                       Reduce red, CgState* st)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k0 = blockIdx.z * kc + 1;
    const double mb = st->mb;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i <= o.nx && j <= o.ny) {
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        for (int k = k0; k < k0 + kc && k <= o.nz; ++k) {
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
    double tot[3];
    if (grid_sums(3, acc, red, tot) && tid256() == 0) {
        const double N = double(o.nx) * o.ny * o.nz;
        double bnorm = sqrt(tot[0]);
        double m = tot[1] / N;
        double beta = tot[2] - tot[1] * m;
        double dp = sqrt(beta > 0.0 ? beta : 0.0);
        double rnorm0 = bnorm > 0.0 ? bnorm : dp;
        double ttol = st->rtol * rnorm0 > st->abstol ? st->rtol * rnorm0 : st->abstol;
        st->bnorm = bnorm;
        st->m = m;
        st->beta = beta;
        st->betaold = 0.0;
        st->dp = dp;
        st->rnorm0 = rnorm0;
        st->ttol = ttol;
        st->its = 0;
        st->i = 0;
        st->reason = 0;
        st->done = 0;
        if (isnan(dp)) {
            st->reason = -9;
            st->done = 1;
        } else if (dp <= ttol) {
            st->reason = dp < st->abstol ? 3 : 2;
            st->done = 1;
        }
        *red.counter = 0u;
    }
}

// p <- z + (beta/beta_old) p  (p <- z on the first iteration), w = A p, p.w
__global__ void __launch_bounds__(BX* BY) k_cg_p(Grid g, Op o, int kc, FieldView<double> r, FieldView<double> po,
                                                 FieldView<double> pn, Reduce red, CgState* st)
{
    if (st->done) return;
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k0 = blockIdx.z * kc + 1;
    const double m = st->m;
    const bool first = st->i == 0;
    const double bt = first ? 0.0 : st->beta / st->betaold;
    double acc = 0.0;
    if (i <= o.nx && j <= o.ny) {
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        auto P = [&](int ii, int jj, int kk) -> double {
            double z = r(ii, jj, kk) - m;
            return first ? z : z + bt * po(ii, jj, kk);
        };
        const int k1 = min(k0 + kc - 1, o.nz);
        double pc = P(i, j, k0);
        double pm = k0 > 1 ? P(i, j, k0 - 1) : pc;
        for (int k = k0; k <= k1; ++k) {
            double pp = k < o.nz ? P(i, j, k + 1) : pc;
            double w = apply_op(o, pc, i > 1 ? P(i - 1, j, k) : pc, i < o.nx ? P(i + 1, j, k) : pc, P(i, jm, k),
                                P(i, jp, k), k > 1 ? pm : pc, pp);
            pn(i, j, k) = pc;
            acc += pc * w;
            pm = pc;
            pc = pp;
        }
    }
    double tot[1];
    if (grid_sums(1, &acc, red, tot) && tid256() == 0) {
        double dpi = tot[0];
        st->dpi = dpi;
        if (!(dpi > 0.0)) {
            st->reason = isnan(dpi) ? -9 : -10;
            st->its = st->i + 1;
            st->done = 1;
        } else {
            st->alpha = st->beta / dpi;
        }
        *red.counter = 0u;
    }
}

// x <- x + alpha p, r <- r - alpha A p, sum r, sum r^2, convergence test.
__global__ void __launch_bounds__(BX* BY) k_cg_x(Grid g, Op o, int kc, FieldView<double> p, FieldView<double> x,
                                                 FieldView<double> r, Reduce red, CgState* st)
{
    if (st->done) return;
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k0 = blockIdx.z * kc + 1;
    const double a = st->alpha;
    double acc[2] = {0.0, 0.0};
    if (i <= o.nx && j <= o.ny) {
        const int jm = wrapj(j - 1, o.ny), jp = wrapj(j + 1, o.ny);
        const int k1 = min(k0 + kc - 1, o.nz);
        double pc = p(i, j, k0);
        double pm = k0 > 1 ? p(i, j, k0 - 1) : pc;
        for (int k = k0; k <= k1; ++k) {
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
    double tot[2];
    if (grid_sums(2, acc, red, tot) && tid256() == 0) {
        const double N = double(o.nx) * o.ny * o.nz;
        double m = tot[0] / N;
        double beta = tot[1] - tot[0] * m;
        double dp = sqrt(beta > 0.0 ? beta : 0.0);
        st->its = st->i + 1;
        st->m = m;
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
        *red.counter = 0u;
    }
}

// Ghost planes of the solution (copy_pressure_out_of_petsc,
// pressure_solver.F90:444-486, with the eliminated Neumann rows).  One thread
// per (i, j) column, i in 0..maxl; pres may alias x.
template <typename T>
__global__ void k_pres_ghosts(Grid g, T dx, T dy, T dt, FieldView<double> x, FieldView<T> u3, const T* u_inlet,
                              FieldView<T> pres)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    const int L = g.maxl, n = g.maxn;
    if (i > L) return;
    const int jm = wrapj(j - 1, g.P);
    for (int k = 1; k <= n - 1; ++k) {
        double v;
        if (i == 0) {
            double b_in = (double(dx) / double(dt)) * (double(u_inlet[k - 1]) - double(u3(0, jm, k - 1))) / double(dy);
            v = x(1, j, k) + double(dy) * b_in;
        } else if (i == L) {
            v = x(L - 1, j, k);
        } else {
            v = x(i, j, k);
        }
        pres(i, j, k) = T(v);
        if (k == 1) pres(i, j, 0) = T(v);
        if (k == n - 1) pres(i, j, n) = T(v);
    }
}

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

template <typename T>
hipError_t launch_cg_setup_b(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> u3, FieldView<T> v3,
                             FieldView<T> w3, const T* u_inlet, FieldView<double> rhs_p, int use_rhs,
                             FieldView<double> b, Reduce red, CgState* st)
{
    CgShape sh = cg_shape(g);
    hipLaunchKernelGGL(k_cg_setup_b<T>, sh.grid, dim3(BX, BY), 0, s, g, ph.dx, ph.dy, ph.dz, ph.dt, sh.kc, u3, v3, w3,
                       u_inlet, rhs_p, use_rhs, b, red, st);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_setup_r(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x, FieldView<double> r,
                             Reduce red, CgState* st)
{
    CgShape sh = cg_shape(g);
    hipLaunchKernelGGL(k_cg_setup_r, sh.grid, dim3(BX, BY), 0, s, g, make_op(g, ph), sh.kc, x, r, red, st);
    return hipGetLastError();
}

hipError_t launch_cg_p(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> r, FieldView<double> p_old,
                       FieldView<double> p_new, Reduce red, CgState* st)
{
    CgShape sh = cg_shape(g);
    hipLaunchKernelGGL(k_cg_p, sh.grid, dim3(BX, BY), 0, s, g, make_op(g, ph), sh.kc, r, p_old, p_new, red, st);
    return hipGetLastError();
}

hipError_t launch_cg_x(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> p, FieldView<double> x,
                       FieldView<double> r, Reduce red, CgState* st)
{
    CgShape sh = cg_shape(g);
    hipLaunchKernelGGL(k_cg_x, sh.grid, dim3(BX, BY), 0, s, g, make_op(g, ph), sh.kc, p, x, r, red, st);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pres_ghosts(hipStream_t s, const Grid& g, const Phys& ph, FieldView<double> x, FieldView<T> u3,
                              const T* u_inlet, FieldView<T> pres)
{
    hipLaunchKernelGGL(k_pres_ghosts<T>, dim3((g.maxl + 1 + 63) / 64, g.P), dim3(64), 0, s, g, T(ph.dx), T(ph.dy),
                       T(ph.dt), x, u3, u_inlet, pres);
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
    template hipError_t launch_fill<T>(hipStream_t, T*, long, T);                                                   \
    template hipError_t launch_copy<T>(hipStream_t, T*, const T*, long);
TPLS_INST_PR(double)
TPLS_INST_PR(float)
template hipError_t launch_fill<int>(hipStream_t, int*, long, int);

}  // namespace tpls
