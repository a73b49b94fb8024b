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
#include "group.hpp"
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
struct OpCgSetupB {
    struct Args {
        Grid g;
        Tiles tl;
        double dx;
        double dy;
        double dz;
        double dt;
        FieldView<T> u3;
        FieldView<T> v3;
        FieldView<T> w3;
        const T* u_inlet;
        FieldView<double> rhs_p;
        int use_rhs;
        FieldView<double> b;
        double* part;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

template <typename T>
__device__ __forceinline__ void OpCgSetupB<T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    Tiles tl = args_.tl;
    double dx = args_.dx;
    double dy = args_.dy;
    double dz = args_.dz;
    double dt = args_.dt;
    FieldView<T> u3 = args_.u3;
    FieldView<T> v3 = args_.v3;
    FieldView<T> w3 = args_.w3;
    const T* u_inlet = args_.u_inlet;
    FieldView<double> rhs_p = args_.rhs_p;
    int use_rhs = args_.use_rhs;
    FieldView<double> b = args_.b;
    double* part = args_.part;
    const int nx = g.maxl - 1, nz = g.maxn - 1;
    double acc[1] = {0.0};
    for (int t = bk_.x; t < tl.ntiles; t += bk_.n) {
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
    block_partials<1>(acc, part, bk_.n, bk_.x);
}

template <typename T>
__global__ void __launch_bounds__(BX* BY) k_cg_setup_b(Grid g, Tiles tl, double dx, double dy, double dz, double dt,
                                                       FieldView<T> u3, FieldView<T> v3, FieldView<T> w3,
                                                       const T* u_inlet, FieldView<double> rhs_p, int use_rhs,
                                                       FieldView<double> b, double* part)
{
    OpCgSetupB<T>::run({g, tl, dx, dy, dz, dt, u3, v3, w3, u_inlet, rhs_p, use_rhs, b, part}, blk_here());
}

// Reset the device state of a solve (a kernel rather than a host copy so that
// the set-up can be captured in a graph).
struct OpCgInit {
    struct Args {
        CgState* st;
        double rtol;
        double abstol;
        double dtol;
        int maxits;
    };
    static constexpr int THREADS = 64;
    static dim3 block() { return dim3(64); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

__device__ __forceinline__ void OpCgInit::run(const Args& args_, const Blk& bk_)
{
    CgState* st = args_.st;
    double rtol = args_.rtol;
    double abstol = args_.abstol;
    double dtol = args_.dtol;
    int maxits = args_.maxits;
    if (threadIdx.x != 0) return;
    CgState s{};
    s.rtol = rtol;
    s.abstol = abstol;
    s.dtol = dtol;
    s.maxits = maxits;
    *st = s;
}

__global__ void k_cg_init(CgState* st, double rtol, double abstol, double dtol, int maxits)
{
    OpCgInit::run({st, rtol, abstol, dtol, maxits}, blk_here());
}

struct OpCgSetupBf {
    struct Args {
        const double* part;
        int nb;
        double N;
        CgState* st;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

__device__ __forceinline__ void OpCgSetupBf::run(const Args& args_, const Blk& bk_)
{
    const double* part = args_.part;
    int nb = args_.nb;
    double N = args_.N;
    CgState* st = args_.st;
    double tot[1];
    finish_sums<1>(part, nb, tot);
    if (tid256() == 0) st->mb = tot[0] / N;
}

__global__ void __launch_bounds__(256) k_cg_setup_bf(const double* part, int nb, double N, CgState* st)
{
    OpCgSetupBf::run({part, nb, N, st}, blk_here());
}

// b <- b - mean(b); r = b - A x; partials of |b|^2, sum r, sum r^2.
struct OpCgSetupR {
    struct Args {
        Op o;
        Tiles tl;
        FieldView<double> x;
        FieldView<double> r;
        double* part;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

__device__ __forceinline__ void OpCgSetupR::run(const Args& args_, const Blk& bk_)
{
    Op o = args_.o;
    Tiles tl = args_.tl;
    FieldView<double> x = args_.x;
    FieldView<double> r = args_.r;
    double* part = args_.part;
    const CgState* st = args_.st;
    const double mb = st->mb;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = bk_.x; t < tl.ntiles; t += bk_.n) {
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
    block_partials<3>(acc, part, bk_.n, bk_.x);
}

__global__ void __launch_bounds__(BX* BY) k_cg_setup_r(Op o, Tiles tl, FieldView<double> x, FieldView<double> r,
                                                       double* part, const CgState* st)
{
    OpCgSetupR::run({o, tl, x, r, part, st}, blk_here());
}

struct OpCgSetupRf {
    struct Args {
        const double* part;
        int nb;
        double N;
        CgState* st;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

__device__ __forceinline__ void OpCgSetupRf::run(const Args& args_, const Blk& bk_)
{
    const double* part = args_.part;
    int nb = args_.nb;
    double N = args_.N;
    CgState* st = args_.st;
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

__global__ void __launch_bounds__(256) k_cg_setup_rf(const double* part, int nb, double N, CgState* st)
{
    OpCgSetupRf::run({part, nb, N, st}, blk_here());
}

// p <- z + (beta/beta_old) p  (p <- z on the first iteration), w = A p, p.w.
// Each plane of the new p over the block's 64 x 4 columns is formed once and
// shared through LDS; the x/y neighbours outside the block are formed by the
// edge lanes / rows.  FIRST is the first iteration (p = z exactly).
template <bool FIRST, typename Z>
__device__ __forceinline__ void cg_p_tiles(const Op& o, const Tiles& tl, const FieldView<Z>& r,
                                           const FieldView<double>& po, const FieldView<double>& pn, double m,
                                           double bt, double& acc, const Blk& bk)
{
    __shared__ double sp[BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto P = [&](int ii, int jj, int kk) -> double {
        double z = double(r(ii, jj, kk)) - m;
        return FIRST ? z : z + bt * po(ii, jj, kk);
    };
    for (int t = bk.x; t < tl.ntiles; t += bk.n) {
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
struct CgP {
    struct Args {
        Op o;
        Tiles tl;
        FieldView<Z> r;
        FieldView<double> po, pn;
        int use_mz;
        double* part;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk)
    {
        const CgState* st = a.st;
        if (st->done) return;
        const double m = a.use_mz ? st->mz : st->m;
        double acc[1] = {0.0};
        if (st->i == 0)
            cg_p_tiles<true>(a.o, a.tl, a.r, a.po, a.pn, m, 0.0, acc[0], bk);
        else
            cg_p_tiles<false>(a.o, a.tl, a.r, a.po, a.pn, m, st->beta / st->betaold, acc[0], bk);
        block_partials<1>(acc, a.part, bk.n, bk.x);
    }
};

template <typename Z>
__global__ void __launch_bounds__(BX* BY) k_cg_p(Op o, Tiles tl, FieldView<Z> r, FieldView<double> po,
                                                 FieldView<double> pn, int use_mz, double* part, const CgState* st)
{
    CgP<Z>::run({o, tl, r, po, pn, use_mz, part, st}, blk_here());
}

struct CgPf {
    struct Args {
        const double* part;
        int nb;
        CgState* st;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& a, const Blk&);
};

__global__ void __launch_bounds__(256) k_cg_pf(const double* part, int nb, CgState* st)
{
    CgPf::run({part, nb, st}, blk_here());
}

__device__ __forceinline__ void CgPf::run(const Args& a, const Blk&)
{
    const double* part = a.part;
    const int nb = a.nb;
    CgState* st = a.st;
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
struct CgX {
    struct Args {
        Op o;
        Tiles tl;
        FieldView<double> p, x, r;
        double* part;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& args, const Blk& bk);
};

__global__ void __launch_bounds__(BX* BY) k_cg_x(Op o, Tiles tl, FieldView<double> p, FieldView<double> x,
                                                 FieldView<double> r, double* part, const CgState* st)
{
    CgX::run({o, tl, p, x, r, part, st}, blk_here());
}

__device__ __forceinline__ void CgX::run(const Args& args, const Blk& bk)
{
    const Op& o = args.o;
    const Tiles& tl = args.tl;
    const FieldView<double>& p = args.p;
    const FieldView<double>& x = args.x;
    const FieldView<double>& r = args.r;
    const CgState* st = args.st;
    if (st->done) return;
    const double a = st->alpha;
    double acc[2] = {0.0, 0.0};
    for (int t = bk.x; t < tl.ntiles; t += bk.n) {
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
    block_partials<2>(acc, args.part, bk.n, bk.x);
}

struct CgXf {
    struct Args {
        const double* part;
        int nb;
        double N;
        CgState* st;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& a, const Blk&);
};

__global__ void __launch_bounds__(256) k_cg_xf(const double* part, int nb, double N, CgState* st)
{
    CgXf::run({part, nb, N, st}, blk_here());
}

__device__ __forceinline__ void CgXf::run(const Args& a, const Blk&)
{
    const double* part = a.part;
    const int nb = a.nb;
    const double N = a.N;
    CgState* st = a.st;
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
struct OpSqdiff {
    struct Args {
        Grid g;
        int nz;
        FieldView<T> a;
        FieldView<T> b;
        double* part;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

template <typename T>
__device__ __forceinline__ void OpSqdiff<T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nz = args_.nz;
    FieldView<T> a = args_.a;
    FieldView<T> b = args_.b;
    double* part = args_.part;
    const int nx = g.maxl - 1;
    const long n = (long)nx * g.P * nz;
    double acc[1] = {0.0};
    for (long q = bk_.x * 256L + threadIdx.x; q < n; q += (long)bk_.n * 256) {
        int i = int(q % nx) + 1;
        long rr = q / nx;
        int j = int(rr % g.P) + 1;
        int k = int(rr / g.P);
        double d = double(a(i, j, k)) - double(b(i, j, k));
        acc[0] += d * d;
    }
    block_partials<1>(acc, part, bk_.n, bk_.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_sqdiff(Grid g, int nz, FieldView<T> a, FieldView<T> b, double* part)
{
    OpSqdiff<T>::run({g, nz, a, b, part}, blk_here());
}

struct OpFinal1 {
    struct Args {
        const double* part;
        int nb;
        double* out;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
};

__device__ __forceinline__ void OpFinal1::run(const Args& args_, const Blk& bk_)
{
    const double* part = args_.part;
    int nb = args_.nb;
    double* out = args_.out;
    double tot[1];
    finish_sums<1>(part, nb, tot);
    if (tid256() == 0) out[0] = tot[0];
}

__global__ void __launch_bounds__(256) k_final1(const double* part, int nb, double* out)
{
    OpFinal1::run({part, nb, out}, blk_here());
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

// ---- the CG iteration for many cases in lock-step (group.hpp) ----
struct CgGather {
    struct Args {
        const CgState* src;
        CgState* dst;
    };
    static constexpr int THREADS = 64;
    static dim3 block() { return dim3(64); }
    static __device__ __forceinline__ void run(const Args& a, const Blk&)
    {
        if (threadIdx.x == 0) *a.dst = *a.src;
    }
};

int cg_group_iteration(GroupPlan& P, const std::vector<CgGroupCase>& cs, int cur)
{
    GroupItems<CgP<double>::Args> pd;
    GroupItems<CgP<float>::Args> pf32;
    GroupItems<CgPf::Args> pfin;
    GroupItems<CgX::Args> xs;
    GroupItems<CgXf::Args> xfin;
    for (const CgGroupCase& c : cs) {
        const Tiles tl = cg_tiles(c.g);
        const Op o = make_op(c.g, c.ph);
        const FieldView<double> po = c.pb[cur], pn = c.pb[cur ^ 1];
        if (c.z32) pf32.add({o, tl, c.zf, po, pn, c.use_mz, c.part, c.st}, dim3(tl.nblocks));
        else pd.add({o, tl, c.z, po, pn, c.use_mz, c.part, c.st}, dim3(tl.nblocks));
        pfin.add({c.part, tl.nblocks, c.st}, dim3(1));
        xs.add({o, tl, pn, c.x, c.r, c.part, c.st}, dim3(tl.nblocks));
        xfin.add({c.part, tl.nblocks, n_unknowns(c.g), c.st}, dim3(1));
    }
    P.add<CgP<double>>(pd.v);
    P.add<CgP<float>>(pf32.v);
    P.add<CgPf>(pfin.v);
    P.add<CgX>(xs.v);
    P.add<CgXf>(xfin.v);
    return 0;
}

template <typename T>
void group_difference(GroupPlan& P, const std::vector<const HhGroupCase<T>*>& on, const std::vector<FieldView<T>>& a,
                      const std::vector<FieldView<T>>& b, int q)
{
    GroupItems<typename OpSqdiff<T>::Args> sq;
    GroupItems<OpFinal1::Args> fin;
    for (size_t k = 0; k < on.size(); ++k) {
        const Grid& g = on[k]->g;
        sq.add({g, q == 2 ? g.maxn : g.maxn - 1, a[k], b[k], on[k]->part}, dim3(REDUCE_BLOCKS));
        fin.add({on[k]->part, REDUCE_BLOCKS, on[k]->out + q}, dim3(1));
    }
    P.add<OpSqdiff<T>>(sq.v);
    P.add<OpFinal1>(fin.v);
}
template void group_difference<double>(GroupPlan&, const std::vector<const HhGroupCase<double>*>&,
                                       const std::vector<FieldView<double>>&, const std::vector<FieldView<double>>&,
                                       int);
template void group_difference<float>(GroupPlan&, const std::vector<const HhGroupCase<float>*>&,
                                      const std::vector<FieldView<float>>&, const std::vector<FieldView<float>>&, int);

template <typename T>
void cg_group_setup(GroupPlan& P, const std::vector<SetupGroupCase<T>>& cs)
{
    GroupItems<OpCgInit::Args> init;
    GroupItems<typename OpCgSetupB<T>::Args> sb;
    GroupItems<OpCgSetupBf::Args> sbf;
    GroupItems<OpCgSetupR::Args> sr;
    GroupItems<OpCgSetupRf::Args> srf;
    for (const SetupGroupCase<T>& c : cs) {
        const Tiles tl = cg_tiles(c.g);
        init.add({c.st, c.rtol, c.abstol, c.dtol, c.maxits}, dim3(1));
        sb.add({c.g, tl, c.ph.dx, c.ph.dy, c.ph.dz, c.ph.dt, c.u3, c.v3, c.w3, c.inlet, c.rhs_p, 0, c.r, c.part},
               dim3(tl.nblocks));
        sbf.add({c.part, tl.nblocks, n_unknowns(c.g), c.st}, dim3(1));
        sr.add({make_op(c.g, c.ph), tl, c.x, c.r, c.part, c.st}, dim3(tl.nblocks));
        srf.add({c.part, tl.nblocks, n_unknowns(c.g), c.st}, dim3(1));
    }
    P.add<OpCgInit>(init.v);
    P.add<OpCgSetupB<T>>(sb.v);
    P.add<OpCgSetupBf>(sbf.v);
    P.add<OpCgSetupR>(sr.v);
    P.add<OpCgSetupRf>(srf.v);
}
template void cg_group_setup<double>(GroupPlan&, const std::vector<SetupGroupCase<double>>&);
template void cg_group_setup<float>(GroupPlan&, const std::vector<SetupGroupCase<float>>&);

void cg_group_gather(GroupPlan& P, const std::vector<const CgState*>& src, CgState* dst)
{
    GroupItems<CgGather::Args> g;
    for (size_t q = 0; q < src.size(); ++q) g.add({src[q], dst + q}, dim3(1));
    P.add<CgGather>(g.v);
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
