// Geometric multigrid preconditioner of the pressure solve: one symmetric
// V-cycle z = M r per conjugate-gradient iteration (the class of
// configuration PETSc selects with -ksp_type cg -pc_type mg; the reference
// takes its KSP and PC from the options database, configure_mpi_petsc.F90:158).
//
// Definition (identical to the oracle's, operation for operation):
//   levels      the reduced operator on nx x ny x nz cells, halved in every
//               direction while nx and nz are even, ny is a multiple of 4
//               (the red-black colouring across the periodic y seam needs an
//               even ny on every level) and all extents are >= 4; the
//               coefficients scale by 1/4 per level;
//   smoothing   red-black Gauss-Seidel, red = (i+j+k) even; nu sweeps red
//               then black from a zero guess before the coarse correction,
//               nu sweeps black then red after it (the adjoint order, so M is
//               symmetric);
//   transfer    trilinear cell-centred prolongation (3/4, 1/4 per direction,
//               reflected at the Neumann ends, wrapped in y) and its
//               transpose / 8 as restriction;
//   coarsest    the exact solve on the mean-free subspace (the pseudo-
//               inverse of the singular operator) by fast diagonalisation:
//               the operator is cx Lx + cy Ly + cz Lz with 1-D Laplacians
//               Neumann in x and z (cosine modes) and periodic in y
//               (Fourier modes), so z = Q diag(1/lambda) Q^T b with the
//               constant mode dropped, three transforms each way in one
//               block's LDS; a level with an extent above MG_SPEC_MAX
//               instead takes 20 + nmax^2/2 symmetric Gauss-Seidel sweeps
//               (red, black, black, red) from zero on the mean-free
//               right-hand side, mean-free result.
// Within a colour every update reads only the other colour, so the sweeps are
// order-independent and the kernels reproduce the oracle's values; FMA
// contraction is off in this file for the same reason.
//
// Precision: the kernels are templates on the storage type of a level's x,
// which is also the type of its arithmetic.  The FP64 hierarchy is the
// definition above.  The FP32 hierarchy (cases stored in FP32) runs the same
// cycle on FP32 vectors and arithmetic; its finest level reads the CG
// residual in FP64 and the CG reads z back into FP64, so the CG's inner
// products, residual and stopping test stay FP64.
//
// Kernels (one thread per cell unless noted):
//   k_mg_pre     the first pre-smoothing sweep pair from x = 0, fused: red
//                cells take b/d, black cells form their red neighbours'
//                values b_n/d_n on the fly (one thread per red-black pair);
//   k_mg_smooth  one colour sweep;
//   k_mg_rr      the residual b - A x and its restriction to the next level,
//                fused (the residual never leaves LDS);
//   k_mg_prolong x += P e;
//   k_mg_spec    the coarsest solve by fast diagonalisation, one block;
//   k_mg_coarse  the coarsest solve by sweeps (large coarsest levels);
//   k_mg_last    the final red sweep on the finest level, fused with the
//                partials of z.r and sum z (grid-stride tiles, at most
//                CG_BLOCKS partials);
//   k_mg_zf      one block: mz = mean z, beta = z.r - mz sum r and the
//                PETSc breakdown checks of a preconditioned CG step.
#include <cmath>
#include <type_traits>

#include "group.hpp"
#include "reduce.hpp"
#include "tiles.hpp"
#include "tpls_kernels.hpp"

#pragma clang fp contract(off)

namespace tpls {

// Arithmetic of a level is done in C, the storage type of its x (double for
// the FP64 hierarchy, which is the oracle's operation for operation; float
// for the FP32 one, whose divisions by the diagonal use the hardware
// reciprocal: sdiv).
template <typename C, typename L>
__host__ __device__ __forceinline__ C mg_deff(const L& o, int i, int k)
{
    C d = C(o.diag);
    if (k == 1) d = d - C(o.cz);
    if (k == o.nz) d = d - C(o.cz);
    if (i == 1) d = d - C(o.cx);
    if (i == o.nx) d = d - C(o.cx);
    return d;
}

// sum of c * neighbour over the interior neighbours of (i,j,k); v(i,j,k)
// gives the neighbour values
template <typename C, typename L, typename V>
__device__ __forceinline__ C mg_nsum(const L& o, const V& v, int i, int j, int k)
{
    const int jm = j == 1 ? o.ny : j - 1, jp = j == o.ny ? 1 : j + 1;
    const C zm = k > 1 ? C(v(i, j, k - 1)) : C(0);
    const C zp = k < o.nz ? C(v(i, j, k + 1)) : C(0);
    const C xm = i > 1 ? C(v(i - 1, j, k)) : C(0);
    const C xp = i < o.nx ? C(v(i + 1, j, k)) : C(0);
    const C ym = C(v(i, jm, k)), yp = C(v(i, jp, k));
    return C(o.cz) * (zm + zp) + C(o.cx) * (xm + xp) + C(o.cy) * (ym + yp);
}

// 1-D trilinear pieces: fine cell f (1-based) takes 3/4 of its parent
// (f+1)/2 and 1/4 of the neighbour toward f, reflected to the parent at a
// Neumann end or wrapped when periodic.
__device__ __forceinline__ int mg_nbr(int f, int nc, bool periodic)
{
    const int I = (f + 1) / 2;
    int n = (f & 1) ? I - 1 : I + 1;
    if (n < 1) n = periodic ? nc : I;
    if (n > nc) n = periodic ? 1 : I;
    return n;
}

// first pre-smoothing pair (red then black) from x = 0.  One thread per two
// cells along x, one of each colour (nx is even on every smoothed level), so
// no lane waits on the other colour's branch.  bc (when bc.p is set) receives
// b converted to the cycle's type, which is all the later kernels of the
// level use of it.
template <typename X, typename B>
struct MgPre {
    struct Args {
        MgLevelT<X, B> o;
        FieldView<X> bc;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

// The pre-smoothing pair of the cells (i0, j, k) and (i0+1, j, k), i0 odd.
template <typename X, typename B>
__device__ __forceinline__ void mg_pre_cells(const MgLevelT<X, B>& o, const FieldView<X>& bc, int i0, int j, int k)
{
    using C = X;
    const int ir = ((i0 + j + k) & 1) == 0 ? i0 : i0 + 1;
    const int ib = ir == i0 ? i0 + 1 : i0;
    auto red = [&](int ii, int jj, int kk) { return sdiv(C(o.b(ii, jj, kk)) + C(0), mg_deff<C>(o, ii, kk)); };
    o.x(ir, j, k) = red(ir, j, k);
    o.x(ib, j, k) = sdiv(C(o.b(ib, j, k)) + mg_nsum<C>(o, red, ib, j, k), mg_deff<C>(o, ib, k));
    if (bc.p) {
        bc(ir, j, k) = C(o.b(ir, j, k));
        bc(ib, j, k) = C(o.b(ib, j, k));
    }
}

template <typename X, typename B>
__device__ __forceinline__ void MgPre<X, B>::run(const Args& a, const Blk& bk)
{
    const MgLevelT<X, B>& o = a.o;
    if (a.st->done) return;
    const int i0 = 2 * (bk.x * BX + int(threadIdx.x)) + 1;
    const int j = bk.y * BY + int(threadIdx.y) + 1, k = bk.z + 1;
    if (i0 > o.nx || j > o.ny) return;
    mg_pre_cells(o, a.bc, i0, j, k);
}

template <typename X, typename B>
__global__ void __launch_bounds__(BX* BY) k_mg_pre(MgLevelT<X, B> o, FieldView<X> bc, const CgState* st)
{
    MgPre<X, B>::run({o, bc, st}, blk_here());
}

// The cell of colour `colour` among (i0, j, k) and (i0+1, j, k), i0 odd.
template <typename X, typename B>
__device__ __forceinline__ void mg_smooth_cell(const MgLevelT<X, B>& o, int colour, int i0, int j, int k)
{
    using C = X;
    const int i = ((i0 + j + k) & 1) == colour ? i0 : i0 + 1;
    o.x(i, j, k) = sdiv(C(o.b(i, j, k)) + mg_nsum<C>(o, o.x, i, j, k), mg_deff<C>(o, i, k));
}

// One colour sweep.  A thread owns the pair (i0, i0+1) along x and updates
// the one of the sweep's colour (nx is even on every smoothed level), so
// every lane of a wavefront has a cell to update.
template <typename X, typename B>
struct MgSmooth {
    struct Args {
        MgLevelT<X, B> o;
        int colour;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk)
    {
        const MgLevelT<X, B>& o = a.o;
        if (a.st->done) return;
        const int i0 = 2 * (bk.x * BX + int(threadIdx.x)) + 1;
        const int j = bk.y * BY + int(threadIdx.y) + 1, k = bk.z + 1;
        if (i0 > o.nx || j > o.ny) return;
        mg_smooth_cell(o, a.colour, i0, j, k);
    }
};

template <typename X, typename B>
__global__ void __launch_bounds__(BX* BY) k_mg_smooth(MgLevelT<X, B> o, int colour, const CgState* st)
{
    MgSmooth<X, B>::run({o, colour, st}, blk_here());
}

template <typename L>
static dim3 pair_grid(const L& o) { return dim3((o.nx / 2 + BX - 1) / BX, (o.ny + BY - 1) / BY, o.nz); }

// Both sweeps of a smoothing pair, colour f then 1 - f, in one k-march and
// out of place: xout receives what k_mg_smooth(f) then k_mg_smooth(1 - f)
// leave in x, bit for bit (each cell reads the values it reads there; within
// a colour the updates are order-free).  A block owns MP_TX x MP_TY columns
// and a chunk of levels; a lane owns a pair of columns (a, a+1) of the tile
// and its one-cell ring, one row.  At step m the lane updates its pair's
// colour-f cell at level m from the old x (planes staged in LDS as in the
// Helmholtz marches) and publishes it in LDS; the pair's colour-(1-f) cell
// at level m-1 is in the same column, so the lane updates it next from the
// colour-f values of level m-1 in LDS (in-plane) and its own colour-f values
// of levels m-2 and m (z).  Absent neighbours are 0 as in mg_nsum.  The
// first colour is formed one level below and above the chunk, so chunks
// overlap only in reads.  ZERO: the old x is zero (nothing is loaded): the
// pass is then the first pre-smoothing pair, k_mg_pre's values bit for bit
// (b + 0 where k_mg_pre adds 0, and the same first-colour values around
// each second-colour cell).
constexpr int MP_TX = 64, MP_TY = 8;
template <typename X, typename B, bool ZERO = false>
struct MgPair {
    static constexpr int PW = (MP_TX + 2) / 2, NP = PW * (MP_TY + 2), NT = (NP + 63) / 64 * 64;
    // levels of old x and b loaded ahead: 8 measured 1.5 % faster per FP32
    // step, 4 faster in FP64
    static constexpr int D = sizeof(X) == 4 ? 8 : 4;
    struct Args {
        MgLevelT<X, B> o;
        FieldView<X> xout;
        int f;
        int nbx, nby, ks;
        const CgState* st;
        // ZERO only: when set, receives b converted to X at every cell the
        // pass writes (as k_mg_pre's bc)
        FieldView<X> bc;
    };
    static constexpr int THREADS = NT;
    static dim3 block() { return dim3(NT); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

template <typename X, typename B, bool ZERO>
__device__ __forceinline__ void MgPair<X, B, ZERO>::run(const Args& a, const Blk& bk)
{
    using C = X;
    const MgLevelT<X, B>& o = a.o;
    if (a.st->done) return;
    constexpr int SR = MP_TY + 4, SW = MP_TX + 4, SN = SR * SW, NS = (SN + NT - 1) / NT;
    __shared__ C xs[4][SR][SW];
    __shared__ C fp[2][MP_TY + 2][MP_TX + 2];
    const int t = xcd_tile(bk.x, bk.n);
    const int bx = t % a.nbx, rest = t / a.nbx, by = rest % a.nby, bz = rest / a.nby;
    const int i0 = 1 + bx * MP_TX, j0 = 1 + by * MP_TY;
    const int nx = o.nx, ny = o.ny, nz = o.nz, f = a.f;
    const int kb = 1 + bz * a.ks, ke = min(kb + a.ks - 1, nz);
    auto wrap = [&](int jj) { return jj < 1 ? jj + ny : (jj > ny ? jj - ny : jj); };
    const int e = int(threadIdx.x);
    const bool lane = e < NP;
    const int r = lane ? e / PW : 1, h = lane ? e % PW : 0;
    const int j = wrap(j0 - 1 + r);
    const int ia = i0 - 1 + 2 * h;  // the pair's first column; region column 2h
    const bool row_in = lane && r >= 1 && r <= MP_TY;
    // staged old x: rows j0-2.. (wrapped), columns i0-2.. and planes clamped
    // into the arrays (a clamped value is read by no updated cell)
    long ox[NS];
    int sl[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int q = min(e + n * NT, SN - 1);
        const int er = q / SW, ec = q % SW;
        sl[n] = q;
        ox[n] = min(max(i0 - 2 + ec, 0), nx + 1) + long(wrap(j0 - 2 + er)) * o.x.sy;
    }
    auto stage_load = [&](C* v, int st) {
        const long kx = long(min(max(st, 1), nz)) * o.x.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) v[n] = ZERO ? C(0) : o.x.p[ox[n] + kx];
    };
    auto stage_store = [&](const C* v, int st) {
        C* px = &xs[st & 3][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) px[sl[n]] = v[n];
    };
    // b of the pair's two cells at a level (clamped into the interior)
    const int ib0 = min(max(ia, 1), nx), ib1 = min(max(ia + 1, 1), nx);
    auto b_at = [&](C* v, int m) {
        const int mm = min(max(m, 1), nz);
        v[0] = C(o.b(ib0, j, mm));
        v[1] = C(o.b(ib1, j, mm));
    };
    const int m0 = kb - 1;
    {
        C v[NS];
#pragma unroll
        for (int st = m0 - 1; st <= m0 + 1; ++st) {
            stage_load(v, st);
            stage_store(v, st);
        }
    }
    C rx[D][NS], rb[D][2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        stage_load(rx[d], m0 + 2 + d);
        b_at(rb[d], m0 + d);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const C cx = C(o.cx), cy = C(o.cy), cz = C(o.cz);
    C fm1 = C(0), fm2 = C(0);  // this lane's colour-f values of the last two steps
    C bprev[2] = {C(0), C(0)};
    const int send = m0 + (ke + 2 - m0 + D - 1) / D * D;
    for (int sb = m0; sb < send; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int m = sb + d, sx = m & 3;
            stage_store(rx[d], m + 2);
            stage_load(rx[d], m + 2 + D);
            const C bc[2] = {rb[d][0], rb[d][1]};
            b_at(rb[d], m + D);
            // colour f at level m, column cf (region column c)
            const int off = (ia + j + m + f) & 1;
            const int cf = ia + off, c = 2 * h + off;
            const C xm = cf > 1 ? C(xs[sx][r + 1][c]) : C(0);
            const C xp = cf < nx ? C(xs[sx][r + 1][c + 2]) : C(0);
            const C ym = C(xs[sx][r][c + 1]), yp = C(xs[sx][r + 2][c + 1]);
            const C zm = m > 1 ? C(xs[(m - 1) & 3][r + 1][c + 1]) : C(0);
            const C zp = m < nz ? C(xs[(m + 1) & 3][r + 1][c + 1]) : C(0);
            const C fv = sdiv(bc[off] + (cz * (zm + zp) + cx * (xm + xp) + cy * (ym + yp)), mg_deff<C>(o, cf, m));
            if (lane) fp[m & 1][r][c] = fv;
            // colour 1 - f at level k = m - 1, the same column
            const int k = m - 1;
            if (row_in && k >= kb && k <= ke) {
                const int q = k & 1;
                // (a ring column's own update is discarded: clamp its slots)
                const C sxm = cf > 1 ? C(fp[q][r][max(c - 1, 0)]) : C(0);
                const C sxp = cf < nx ? C(fp[q][r][min(c + 1, MP_TX + 1)]) : C(0);
                const C szm = k > 1 ? fm2 : C(0);
                const C szp = k < nz ? fv : C(0);
                const C sv = sdiv(bprev[off] + (cz * (szm + szp) + cx * (sxm + sxp) +
                                                cy * (C(fp[q][r - 1][c]) + C(fp[q][r + 1][c]))),
                                  mg_deff<C>(o, cf, k));
                // the pair's other column held colour f at level k: fm1
                const int co = ia + (off ^ 1);
                if (cf >= i0 && cf < i0 + MP_TX && cf <= nx) a.xout(cf, j, k) = sv;
                if (co >= i0 && co < i0 + MP_TX && co <= nx) a.xout(co, j, k) = fm1;
                if (ZERO && a.bc.p) {
                    if (cf >= i0 && cf < i0 + MP_TX && cf <= nx) a.bc(cf, j, k) = bprev[off];
                    if (co >= i0 && co < i0 + MP_TX && co <= nx) a.bc(co, j, k) = bprev[off ^ 1];
                }
            }
            fm2 = fm1;
            fm1 = fv;
            bprev[0] = bc[0];
            bprev[1] = bc[1];
            __syncthreads();
        }
    }
}

template <typename X, typename B, bool ZERO = false>
__global__ void __launch_bounds__((MgPair<X, B, ZERO>::NT)) k_mg_pair(typename MgPair<X, B, ZERO>::Args a)
{
    MgPair<X, B, ZERO>::run(a, blk_here());
}

// The pair sweeps of a level as one k_mg_pair launch into xout when there is
// one (the case allocates it where mg_pair_fits, or as tpls_params.mg_pair
// asks) and the level has whole tiles in y, else two k_mg_smooth launches in
// place.  Returns the vector holding the result.
template <typename X, typename B, bool ZERO = false>
static FieldView<X> mg_pair(hipStream_t s, const MgLevelT<X, B>& o, FieldView<X> xout, int f, const CgState* st,
                            FieldView<X> bc = FieldView<X>{nullptr, 0, 0})
{
    if (xout.p && o.ny % MP_TY == 0) {
        using Op = MgPair<X, B, ZERO>;
        static const int slots = resident_blocks(k_mg_pair<X, B, ZERO>, Op::NT);
        const int nbx = (o.nx + MP_TX - 1) / MP_TX, nby = o.ny / MP_TY;
        const int ks = march_chunk(nbx * nby, o.nz, slots, 32);
        const int nbz = (o.nz + ks - 1) / ks;
        hipLaunchKernelGGL((k_mg_pair<X, B, ZERO>), dim3(nbx * nby * nbz), dim3(Op::NT), 0, s,
                           typename Op::Args{o, xout, f, nbx, nby, ks, st, bc});
        return xout;
    }
    if (ZERO) {
        // the first pre-smoothing pair into xout as k_mg_pre forms it
        MgLevelT<X, B> o2 = o;
        o2.x = xout.p ? xout : o.x;
        hipLaunchKernelGGL((k_mg_pre<X, B>), pair_grid(o2), dim3(BX, BY), 0, s, o2, bc, st);
        return o2.x;
    }
    const dim3 blk(BX, BY);
    hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, f, st);
    hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, 1 - f, st);
    return o.x;
}

// Weights of the fine cells 2I-2+q (q = 0..3) feeding coarse cell I under the
// transpose of the interpolation: the fine cell's parent contributes 3/4 and
// its neighbour toward it 1/4 (mg_nbr); w[q] = 0 for a candidate outside a
// Neumann end.  Sums over q = 0..3 skip zero weights.  A periodic candidate
// wraps but keeps its position q.
template <typename C>
__device__ __forceinline__ void mg_rweights_q(int I, int nc, int nf, bool periodic, C* w)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int f = 2 * I - 2 + q;
        if (f < 1) f = periodic ? f + nf : 0;
        if (f > nf) f = periodic ? f - nf : 0;
        const int par = (f + 1) / 2, nb = mg_nbr(f, nc, periodic);
        w[q] = f == 0 ? C(0) : (par == I ? C(0.75) : C(0)) + (nb == I ? C(0.25) : C(0));
    }
}

// Residual and restriction fused: b_c = P^T (b - A x) / 8 without writing the
// fine residual.  A block owns RR_CX x RR_CY coarse columns and rk coarse
// levels and marches the fine planes 2K-2 .. 2K+1 that feed them.  Each fine
// plane's residual over the block's footprint (2 RR_CX + 2 by 2 RR_CY + 2
// cells, y wrapped) is formed in LDS with k_mg_resid's arithmetic (x * d -
// (cz (zm + zp) + cx (xm + xp) + cy (ym + yp)), absent neighbours 0), summed
// along x (si, LDS) and y (sj, registers of the coarse cell's thread), and
// every second plane closes a coarse level along z.  The sums nest and run
// in the order of the restriction's definition, so the coarse values are the
// oracle's.  The inputs of the next plane are loaded while the current one
// is summed.
constexpr int RR_CX = 32, RR_CY = 4, RR_T = 256;
constexpr int RR_FX = 2 * RR_CX + 2, RR_FY = 2 * RR_CY + 2, RR_N = RR_FX * RR_FY;
constexpr int RR_C = (RR_N + RR_T - 1) / RR_T;  // footprint cells per thread

template <typename C>
struct RrIn {
    C b;  // the cell's right-hand side (its x neighbours come from the staged planes)
};

template <typename XF, typename BF, typename XC>
struct MgRr {
    struct Args {
        MgLevelT<XF, BF> f;
        MgLevelT<XC, XC> c;
        int ntx, nty, rk;
        const CgState* st;
    };
    static constexpr int THREADS = RR_T;
    static dim3 block() { return dim3(RR_T); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& blk);
};

template <typename XF, typename BF, typename XC>
__device__ __forceinline__ void MgRr<XF, BF, XC>::run(const Args& a, const Blk& blk)
{
    using C = XF;
    const MgLevelT<XF, BF>& f = a.f;
    const MgLevelT<XC, XC>& c = a.c;
    const int ntx = a.ntx, nty = a.nty, rk = a.rk;
    if (a.st->done) return;
    // x planes over the footprint and a one-cell ring, a ring of four in LDS
    // (plane fk in slot fk & 3): each value is loaded once per block and plane
    constexpr int SX = RR_FX + 2, SY = RR_FY + 2, SXN = SX * SY, NSX = (SXN + RR_T - 1) / RR_T;
    __shared__ C XS[4][SY][SX];
    __shared__ C T[RR_FY][RR_FX];
    __shared__ C SI[RR_FY][RR_CX];
    const int tid = threadIdx.x;
    const int bk = blk.x, tx = bk % ntx, r0 = bk / ntx, ty = r0 % nty, tz = r0 / nty;
    const int I0 = 1 + tx * RR_CX, J0 = 1 + ty * RR_CY, Kb = 1 + tz * rk, Ke = min(Kb + rk - 1, c.nz);
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;  // footprint origin (fine, unwrapped)
    // fine row J (unwrapped) in 1..ny: the footprint can span more than one
    // period on small levels
    auto wrapf = [&](int J) { return ((J - 1) % f.ny + f.ny) % f.ny + 1; };
    // this thread's footprint cells: fine i and wrapped j (footprint row r is
    // fine row fj0 + r, staged row r + 1; column q is fine i fi0 + q, staged
    // column q + 1)
    int ci[RR_C], cj[RR_C];
    bool live[RR_C];
#pragma unroll
    for (int n = 0; n < RR_C; ++n) {
        const int e = min(tid + n * RR_T, RR_N - 1);
        const int row = e / RR_FX, col = e % RR_FX;
        ci[n] = fi0 + col;
        cj[n] = wrapf(fj0 + row);
        live[n] = tid + n * RR_T < RR_N && ci[n] >= 1 && ci[n] <= f.nx;
    }
    // staged x: columns fi0-1.. clamped into the array (a clamped value is
    // read by no live cell), rows wrapped, planes clamped to 1..nz
    long xo[NSX];
    int xsl[NSX];
#pragma unroll
    for (int n = 0; n < NSX; ++n) {
        const int q = min(tid + n * RR_T, SXN - 1);
        const int er = q / SX, ec = q % SX;
        xsl[n] = q;
        xo[n] = min(max(fi0 - 1 + ec, 0), f.nx + 1) + long(wrapf(fj0 - 1 + er)) * f.x.sy;
    }
    auto xload = [&](C* v, int fk) {
        const long kz = long(min(max(fk, 1), f.nz)) * f.x.sz;
#pragma unroll
        for (int n = 0; n < NSX; ++n) v[n] = f.x.p[xo[n] + kz];
    };
    auto xstore = [&](const C* v, int fk) {
        C* px = &XS[fk & 3][0][0];
#pragma unroll
        for (int n = 0; n < NSX; ++n) px[xsl[n]] = v[n];
    };
    auto load = [&](RrIn<C>* in, int fk) {
#pragma unroll
        for (int n = 0; n < RR_C; ++n)
            if (live[n]) in[n].b = C(f.b(ci[n], cj[n], fk));
    };
    // this thread's coarse cell in the y and z sums
    const int cI = tid % RR_CX, cJ = tid / RR_CX;
    const int I = I0 + cI, J = J0 + cJ;
    const bool own = cJ < RR_CY && I <= c.nx && J <= c.ny;
    C wj[4];
    mg_rweights_q(J, c.ny, f.ny, true, wj);
    const int f_lo = max(2 * Kb - 2, 1), f_hi = min(2 * Ke + 1, f.nz);
    RrIn<C> in[RR_C];
    load(in, f_lo);
    C xr[NSX];
    for (int fk = f_lo - 1; fk <= f_lo + 1; ++fk) {
        xload(xr, fk);
        xstore(xr, fk);
    }
    xload(xr, f_lo + 2);
    __syncthreads();
    C s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // sj of fine planes 2K-2 .. 2K+1
    for (int fk = f_lo; fk <= f_hi; ++fk) {
        // the residual of plane fk over the footprint
        const int sc = fk & 3, sm = (fk - 1) & 3, sp = (fk + 1) & 3;
#pragma unroll
        for (int n = 0; n < RR_C; ++n) {
            const int e = tid + n * RR_T;
            if (e >= RR_N) continue;
            const int row = e / RR_FX, col = e % RR_FX;
            C t = 0;  // outside 1..nx: weight 0
            if (live[n]) {
                const int i = ci[n];
                const C xc = XS[sc][row + 1][col + 1];
                const C zm = fk > 1 ? C(XS[sm][row + 1][col + 1]) : C(0);
                const C zp = fk < f.nz ? C(XS[sp][row + 1][col + 1]) : C(0);
                const C xm = i > 1 ? C(XS[sc][row + 1][col]) : C(0);
                const C xp = i < f.nx ? C(XS[sc][row + 1][col + 2]) : C(0);
                const C ym = XS[sc][row][col + 1], yp = XS[sc][row + 2][col + 1];
                const C nsum = C(f.cz) * (zm + zp) + C(f.cx) * (xm + xp) + C(f.cy) * (ym + yp);
                t = in[n].b - (mg_deff<C>(f, i, fk) * xc - nsum);
            }
            T[row][col] = t;
        }
        // plane fk+2 to its slot (the one plane fk-2 had, read last at fk-1)
        xstore(xr, fk + 2);
        xload(xr, fk + 3);
        if (fk < f_hi) load(in, fk + 1);
        __syncthreads();
        for (int e = tid; e < RR_FY * RR_CX; e += RR_T) {
            const int row = e / RR_CX, cx = e % RR_CX;
            C wi[4];
            mg_rweights_q(I0 + cx, c.nx, f.nx, false, wi);
            C si = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (wi[q] != C(0)) si += wi[q] * T[row][2 * cx + q];
            SI[row][cx] = si;
        }
        __syncthreads();
        C sj = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (wj[q] != C(0)) sj += wj[q] * SI[2 * cJ + q < RR_FY ? 2 * cJ + q : 0][cI];
        // plane fk is candidate 2K + (fk & 1) of coarse level K = fk / 2 and
        // candidate 2K - 2 + (fk & 1) of level K + 1
        if (fk & 1) s3 = sj; else s2 = sj;
        const int K = fk / 2;
        if ((fk & 1) || fk == f.nz) {
            // plane 2K+1 done (or the last plane): close coarse level K
            if (own && K >= Kb && K <= Ke) {
                C wk[4];
                mg_rweights_q(K, c.nz, f.nz, false, wk);
                const C sv[4] = {s0, s1, s2, s3};
                C sk = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (wk[q] != C(0)) sk += wk[q] * sv[q];
                c.b(I, J, K) = C(0.125) * sk;
            }
            s0 = s2;
            s1 = s3;
        }
    }
}

template <typename XF, typename BF, typename XC>
__global__ void __launch_bounds__(RR_T) k_mg_rr(MgLevelT<XF, BF> f, MgLevelT<XC, XC> c, int ntx, int nty, int rk,
                                                const CgState* st)
{
    MgRr<XF, BF, XC>::run({f, c, ntx, nty, rk, st}, blk_here());
}

// x_f += P e_c (z first, then y, then x)
// x += P e for the two fine cells 2I-1, 2I of coarse column I (fine nx is
// twice the coarse one): they share their parent's terms (the four values at
// I and the partial sums built from them) and differ only in the x neighbour
// (I-1 for the odd cell, I+1 for the even one), so a thread loads 12 coarse
// values for two cells instead of 16.  Each cell's sum is formed in the same
// order as for one cell alone.
template <typename XF, typename BF, typename XC>
struct MgProlong {
    struct Args {
        MgLevelT<XF, BF> f;
        MgLevelT<XC, XC> c;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

// x += P e for the fine cells (2I-1, j, k) and (2I, j, k).
template <typename XF, typename BF, typename XC>
__device__ __forceinline__ void mg_prolong_cells(const MgLevelT<XF, BF>& f, const MgLevelT<XC, XC>& c, int I, int j,
                                                 int k)
{
    using C = XF;
    const int J = (j + 1) / 2, K = (k + 1) / 2;
    const int Jn = mg_nbr(j, c.ny, true), Kn = mg_nbr(k, c.nz, false);
    const C q3 = C(0.75), q1 = C(0.25);
    const C a00 = q3 * C(c.x(I, J, K)) + q1 * C(c.x(I, J, Kn));
    const C a01 = q3 * C(c.x(I, Jn, K)) + q1 * C(c.x(I, Jn, Kn));
    const C b0 = q3 * a00 + q1 * a01;
    const int i1 = 2 * I - 1, i2 = 2 * I;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = h ? i2 : i1;
        const int In = mg_nbr(i, c.nx, false);
        const C a10 = q3 * C(c.x(In, J, K)) + q1 * C(c.x(In, J, Kn));
        const C a11 = q3 * C(c.x(In, Jn, K)) + q1 * C(c.x(In, Jn, Kn));
        const C b1 = q3 * a10 + q1 * a11;
        f.x(i, j, k) = f.x(i, j, k) + (q3 * b0 + q1 * b1);
    }
}

template <typename XF, typename BF, typename XC>
__device__ __forceinline__ void MgProlong<XF, BF, XC>::run(const Args& a, const Blk& bk)
{
    const MgLevelT<XF, BF>& f = a.f;
    const MgLevelT<XC, XC>& c = a.c;
    if (a.st->done) return;
    const int I = bk.x * BX + int(threadIdx.x) + 1;
    const int j = bk.y * BY + int(threadIdx.y) + 1, k = bk.z + 1;
    if (I > c.nx || j > f.ny) return;
    mg_prolong_cells(f, c, I, j, k);
}

template <typename XF, typename BF, typename XC>
__global__ void __launch_bounds__(BX* BY) k_mg_prolong(MgLevelT<XF, BF> f, MgLevelT<XC, XC> c, const CgState* st)
{
    MgProlong<XF, BF, XC>::run({f, c, st}, blk_here());
}

// Levels of at most MG1_CELLS cells are visited in one block (k_mg_bottom):
// on the way down the first pre-smoothing pair and the nu - 1 pairs (red,
// black) after it, on the way up the prolongation and the nu pairs (black,
// red).  Each sweep covers the level's column pairs block-stride with
// mg_pre_cells / mg_smooth_cell / mg_prolong_cells, a barrier between
// sweeps: every cell computes what k_mg_pre, k_mg_smooth and k_mg_prolong
// compute (within a colour the updates are order-free).
constexpr int MG1_T = 1024;
constexpr long MG1_CELLS = 8192;
template <typename L>
static bool mg_one_block(const L& o) { return long(o.nx) * o.ny * o.nz <= MG1_CELLS; }

template <typename X, typename B, typename F>
__device__ __forceinline__ void mg1_pairs(const MgLevelT<X, B>& o, F&& cell)
{
    const int hx = o.nx / 2, n = hx * o.ny * o.nz;
    for (int q = int(threadIdx.x); q < n; q += MG1_T) {
        const int r = q / hx;
        cell(2 * (q % hx) + 1, r % o.ny + 1, r / o.ny + 1);
    }
    __syncthreads();
}

// The visits of one such level, in the block.
template <typename X, typename B>
__device__ __forceinline__ void mg1_down(const MgLevelT<X, B>& o, int nu)
{
    const FieldView<X> nobc{nullptr, 0, 0};
    mg1_pairs(o, [&](int i0, int j, int k) { mg_pre_cells(o, nobc, i0, j, k); });
    for (int q = 1; q < nu; ++q) {
        mg1_pairs(o, [&](int i0, int j, int k) { mg_smooth_cell(o, 0, i0, j, k); });
        mg1_pairs(o, [&](int i0, int j, int k) { mg_smooth_cell(o, 1, i0, j, k); });
    }
}

template <typename X, typename B, typename XC>
__device__ __forceinline__ void mg1_up(const MgLevelT<X, B>& o, const MgLevelT<XC, XC>& c, int nu)
{
    // the prolongation: one thread per coarse column I and fine (j, k)
    const int n = c.nx * o.ny * o.nz;
    for (int q = int(threadIdx.x); q < n; q += MG1_T) {
        const int r = q / c.nx;
        mg_prolong_cells(o, c, q % c.nx + 1, r % o.ny + 1, r / o.ny + 1);
    }
    __syncthreads();
    for (int q = 0; q < nu; ++q) {
        mg1_pairs(o, [&](int i0, int j, int k) { mg_smooth_cell(o, 1, i0, j, k); });
        mg1_pairs(o, [&](int i0, int j, int k) { mg_smooth_cell(o, 0, i0, j, k); });
    }
}

// The residual-restriction f -> c in the block: every fine cell's residual
// into LDS (r), then each coarse cell's b_c = P^T r / 8 with k_mg_rr's sums
// (along x, then y, then z, each over the candidates q = 0..3 with nonzero
// weight, in order), so the coarse values are k_mg_rr's bit for bit.
template <typename X>
__device__ __forceinline__ void mg1_rr(const MgLevelT<X, X>& f, const MgLevelT<X, X>& c, X* r)
{
    using C = X;
    const int nf = f.nx * f.ny * f.nz;
    for (int q = int(threadIdx.x); q < nf; q += MG1_T) {
        const int i = q % f.nx + 1, rr = q / f.nx, j = rr % f.ny + 1, k = rr / f.ny + 1;
        const int jm = j == 1 ? f.ny : j - 1, jp = j == f.ny ? 1 : j + 1;
        const C xc = f.x(i, j, k);
        const C zm = k > 1 ? C(f.x(i, j, k - 1)) : C(0);
        const C zp = k < f.nz ? C(f.x(i, j, k + 1)) : C(0);
        const C xm = i > 1 ? C(f.x(i - 1, j, k)) : C(0);
        const C xp = i < f.nx ? C(f.x(i + 1, j, k)) : C(0);
        const C ym = f.x(i, jm, k), yp = f.x(i, jp, k);
        const C nsum = C(f.cz) * (zm + zp) + C(f.cx) * (xm + xp) + C(f.cy) * (ym + yp);
        r[q] = C(f.b(i, j, k)) - (mg_deff<C>(f, i, k) * xc - nsum);
    }
    __syncthreads();
    const int nc = c.nx * c.ny * c.nz;
    for (int q = int(threadIdx.x); q < nc; q += MG1_T) {
        const int I = q % c.nx + 1, rr = q / c.nx, J = rr % c.ny + 1, K = rr / c.ny + 1;
        C wi[4], wj[4], wk[4];
        mg_rweights_q(I, c.nx, f.nx, false, wi);
        mg_rweights_q(J, c.ny, f.ny, true, wj);
        mg_rweights_q(K, c.nz, f.nz, false, wk);
        C sk = 0;
#pragma unroll
        for (int qz = 0; qz < 4; ++qz) {
            if (wk[qz] == C(0)) continue;
            const int fk = 2 * K - 2 + qz;
            C sj = 0;
#pragma unroll
            for (int qy = 0; qy < 4; ++qy) {
                if (wj[qy] == C(0)) continue;
                int fj = 2 * J - 2 + qy;
                if (fj < 1) fj += f.ny;
                if (fj > f.ny) fj -= f.ny;
                C si = 0;
#pragma unroll
                for (int qx = 0; qx < 4; ++qx)
                    if (wi[qx] != C(0)) si += wi[qx] * r[(2 * I - 2 + qx - 1) + f.nx * ((fj - 1) + f.ny * (fk - 1))];
                sj += wj[qy] * si;
            }
            sk += wk[qz] * sj;
        }
        c.b(I, J, K) = C(0.125) * sk;
    }
    __syncthreads();
}

// Coarsest level: one block, a single wavefront for up to 64 cells (its
// barriers cost next to nothing), else 256 threads; x and b live in LDS
// (n <= MG_COARSE_MAX).  Each thread's cells, their colours and diagonals are
// worked out once, not in every pass.  The two means are summed by one
// thread in index order.
constexpr int MGC_T = 256, MGC_CPT = MG_COARSE_MAX / MGC_T;

template <typename X>
struct MgCoarse {
    struct Args {
        MgLevelT<X, X> o;
        const CgState* st;
    };
    static constexpr int THREADS = MGC_T;
    static dim3 block() { return dim3(MGC_T); }
    static __device__ __forceinline__ void run(const Args& a, const Blk&);
};

template <typename X>
__device__ __forceinline__ void MgCoarse<X>::run(const Args& a, const Blk&)
{
    using C = X;
    const MgLevelT<X, X>& o = a.o;
    if (a.st->done) return;
    __shared__ C sx[MG_COARSE_MAX], sb[MG_COARSE_MAX];
    __shared__ C mean;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = o.nx * o.ny * o.nz;
    int ci[MGC_CPT], cj[MGC_CPT], ck[MGC_CPT], col[MGC_CPT];
    C cd[MGC_CPT];
#pragma unroll
    for (int c = 0; c < MGC_CPT; ++c) {
        const int q = min(tid + c * nt, n - 1);
        ci[c] = q % o.nx + 1;
        const int r = q / o.nx;
        cj[c] = r % o.ny + 1;
        ck[c] = r / o.ny + 1;
        col[c] = tid + c * nt < n ? (ci[c] + cj[c] + ck[c]) & 1 : 2;  // 2: no cell
        cd[c] = mg_deff<C>(o, ci[c], ck[c]);
    }
#pragma unroll
    for (int c = 0; c < MGC_CPT; ++c)
        if (col[c] != 2) {
            const int q = tid + c * nt;
            sb[q] = o.b(ci[c], cj[c], ck[c]);
            sx[q] = 0;
        }
    __syncthreads();
    if (tid == 0) {
        C s = 0;
        for (int q = 0; q < n; ++q) s += sb[q];
        mean = s / C(n);
    }
    __syncthreads();
    const C mb = mean;
    for (int q = tid; q < n; q += nt) sb[q] = sb[q] - mb;
    __syncthreads();
    const int nmax = max(o.nx, max(o.ny, o.nz));
    const int sweeps = 20 + nmax * nmax / 2;
    auto lds = [&](int i, int j, int k) { return sx[(i - 1) + o.nx * ((j - 1) + o.ny * (k - 1))]; };
    for (int s = 0; s < sweeps; ++s) {
        for (int pass = 0; pass < 4; ++pass) {
            const int colour = (pass == 1 || pass == 2) ? 1 : 0;
#pragma unroll
            for (int c = 0; c < MGC_CPT; ++c)
                if (col[c] == colour) {
                    const int q = tid + c * nt;
                    sx[q] = (sb[q] + mg_nsum<C>(o, lds, ci[c], cj[c], ck[c])) / cd[c];
                }
            __syncthreads();
        }
    }
    if (tid == 0) {
        C s = 0;
        for (int q = 0; q < n; ++q) s += sx[q];
        mean = s / C(n);
    }
    __syncthreads();
    const C mx = mean;
#pragma unroll
    for (int c = 0; c < MGC_CPT; ++c)
        if (col[c] != 2) o.x(ci[c], cj[c], ck[c]) = sx[tid + c * nt] - mx;
}

template <typename X>
__global__ void __launch_bounds__(MGC_T) k_mg_coarse(MgLevelT<X, X> o, const CgState* st)
{
    MgCoarse<X>::run({o, st}, blk_here());
}

// Coarsest level by fast diagonalisation, in FP64 whatever the level's
// storage: b into LDS, b^ = Qz^T Qy^T Qx^T b (one pass per direction, each
// output summed over its line in index order), b^ / lambda with the constant
// mode dropped, then z = Qx Qy Qz b^.  tab: mg_spectral_tables.
template <typename X, typename B>
struct MgSpec {
    struct Args {
        MgLevelT<X, B> o;
        const double* tab;
        const CgState* st;
    };
    static constexpr int THREADS = MGC_T;
    static dim3 block() { return dim3(MGC_T); }
    static __device__ __forceinline__ void run(const Args& a, const Blk&);
};

template <typename X, typename B>
__device__ __forceinline__ void MgSpec<X, B>::run(const Args& a, const Blk&)
{
    const MgLevelT<X, B>& o = a.o;
    const double* tab = a.tab;
    if (a.st->done) return;
    __shared__ double s0[MG_COARSE_MAX], s1[MG_COARSE_MAX];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nx = o.nx, ny = o.ny, nz = o.nz, n = nx * ny * nz;
    const double* Qx = tab;
    const double* Qy = Qx + nx * nx;
    const double* Qz = Qy + ny * ny;
    const double* lx = Qz + nz * nz;
    const double* ly = lx + nx;
    const double* lz = ly + ny;
    for (int q = tid; q < n; q += nt) {
        const int r = q / nx;
        s0[q] = double(o.b(q % nx + 1, r % ny + 1, r / ny + 1));
    }
    __syncthreads();
    // one direction: out = Q^T in (fwd) or Q in along direction d
    auto pass = [&](const double* in, double* out, int d, bool fwd) {
        const int nd = d == 0 ? nx : (d == 1 ? ny : nz);
        const int stride = d == 0 ? 1 : (d == 1 ? nx : nx * ny);
        const double* Q = d == 0 ? Qx : (d == 1 ? Qy : Qz);
        for (int q = tid; q < n; q += nt) {
            const int c = (q / stride) % nd, base = q - c * stride;
            double acc = 0.0;
            for (int m = 0; m < nd; ++m) acc += (fwd ? Q[m * nd + c] : Q[c * nd + m]) * in[base + m * stride];
            out[q] = acc;
        }
        __syncthreads();
    };
    pass(s0, s1, 0, true);
    pass(s1, s0, 1, true);
    pass(s0, s1, 2, true);
    const double shift = o.diag - 2.0 * (o.cx + o.cy + o.cz);
    for (int q = tid; q < n; q += nt) {
        const int r = q / nx;
        const double ev = shift + o.cx * lx[q % nx] + o.cy * ly[r % ny] + o.cz * lz[r / ny];
        s1[q] = q == 0 ? 0.0 : s1[q] / ev;
    }
    __syncthreads();
    pass(s1, s0, 2, false);
    pass(s0, s1, 1, false);
    pass(s1, s0, 0, false);
    for (int q = tid; q < n; q += nt) {
        const int r = q / nx;
        o.x(q % nx + 1, r % ny + 1, r / ny + 1) = X(s0[q]);
    }
}

template <typename X, typename B>
__global__ void __launch_bounds__(MGC_T) k_mg_spec(MgLevelT<X, B> o, const double* tab, const CgState* st)
{
    MgSpec<X, B>::run({o, tab, st}, blk_here());
}

// The bottom of the V-cycle in one block: levels L[0] .. L[n-2] (each of at
// most MG1_CELLS cells) down (mg1_down, mg1_rr) and up (mg1_up) as the
// multi-block kernels would run them, and the coarsest level L[n-1] as
// k_mg_spec (tab set) or k_mg_coarse: one launch instead of about
// (2 nu + 1 + 2 nu + 1) (n - 1) + 1.
constexpr int MG1_MAXLEV = 6;
template <typename X>
struct MgBottom {
    MgLevelT<X, X> L[MG1_MAXLEV];
    int n, nu;
};

template <typename X>
__global__ void __launch_bounds__(MG1_T) k_mg_bottom(MgBottom<X> a, const double* tab, const CgState* st)
{
    if (st->done) return;
    __shared__ X r[MG1_CELLS];
    const int n = a.n;
    for (int l = 0; l + 1 < n; ++l) {
        mg1_down(a.L[l], a.nu);
        mg1_rr(a.L[l], a.L[l + 1], r);
    }
    if (tab) MgSpec<X, X>::run({a.L[n - 1], tab, st}, Blk{0, 0, 0, 1});
    else MgCoarse<X>::run({a.L[n - 1], st}, Blk{0, 0, 0, 1});
    __syncthreads();
    for (int l = n - 2; l >= 0; --l) mg1_up(a.L[l], a.L[l + 1], a.nu);
}

// Final red sweep on the finest level with the partials of z.r (r = b) and
// sum z over all cells, in FP64 from the stored z.  With update = 0 the sweep
// is skipped (a one-level hierarchy, where the coarsest solve already
// produced z).
template <typename X>
struct MgLast {
    struct Args {
        MgLevelT<X, double> o;
        Tiles tl;
        int update;
        double* part;
        const CgState* st;
    };
    static constexpr int THREADS = BX * BY;
    static dim3 block() { return dim3(BX, BY); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
};

template <typename X>
__device__ __forceinline__ void MgLast<X>::run(const Args& a, const Blk& bk)
{
    using C = X;
    const MgLevelT<X, double>& o = a.o;
    const Tiles& tl = a.tl;
    const int update = a.update;
    if (a.st->done) return;
    double acc[2] = {0.0, 0.0};
    for (int t = bk.x; t < tl.ntiles; t += bk.n) {
        int i, j, k0;
        tile_origin(tl, t, i, j, k0);
        if (i > o.nx || j > o.ny) continue;
        for (int k = k0; k < k0 + CG_KC && k <= o.nz; ++k) {
            const double bv = o.b(i, j, k);
            X z;
            if (update && ((i + j + k) & 1) == 0) {
                z = sdiv(C(bv) + mg_nsum<C>(o, o.x, i, j, k), mg_deff<C>(o, i, k));
                o.x(i, j, k) = z;
            } else {
                z = o.x(i, j, k);
            }
            acc[0] += double(z) * bv;
            acc[1] += double(z);
        }
    }
    block_partials<2>(acc, a.part, bk.n, bk.x);
}

template <typename X>
__global__ void __launch_bounds__(BX* BY) k_mg_last(MgLevelT<X, double> o, Tiles tl, int update, double* part,
                                                   const CgState* st)
{
    MgLast<X>::run({o, tl, update, part, st}, blk_here());
}

struct MgZf {
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

__device__ __forceinline__ void MgZf::run(const Args& a, const Blk&)
{
    const double* part = a.part;
    const int nb = a.nb;
    const double N = a.N;
    CgState* st = a.st;
    if (st->done) return;
    double tot[2];
    finish_sums<2>(part, nb, tot);
    if (tid256() != 0) return;
    const double mz = tot[1] / N;
    const double beta = tot[0] - mz * st->sr;
    st->mz = mz;
    st->beta = beta;
    // breakdown checks at the top of the next iteration
    if (beta == 0.0) {
        st->reason = 3;
        st->its = st->i + 1;
        st->done = 1;
    } else if (beta < 0.0) {
        st->reason = -8;
        st->its = st->i + 1;
        st->done = 1;
    }
}

__global__ void __launch_bounds__(256) k_mg_zf(const double* part, int nb, double N, CgState* st)
{
    MgZf::run({part, nb, N, st}, blk_here());
}

// ---- host side ----
int mg_plan(int nx, int ny, int nz, double cx, double cy, double cz, double diag, MgLevel* L)
{
    int l = 0;
    for (;;) {
        MgLevel& o = L[l];
        o = MgLevel{};
        o.nx = nx;
        o.ny = ny;
        o.nz = nz;
        o.cx = cx;
        o.cy = cy;
        o.cz = cz;
        o.diag = diag;
        ++l;
        if (l == MG_MAXLEV || nx % 2 || ny % 4 || nz % 2 || nx < 4 || ny < 4 || nz < 4) break;
        nx /= 2;
        ny /= 2;
        nz /= 2;
        cx *= 0.25;
        cy *= 0.25;
        cz *= 0.25;
        diag *= 0.25;
    }
    return l;
}


// The V-cycle on a hierarchy whose level 0 is `top` (x stored as X, b the
// FP64 CG residual) and whose levels 1.. are `lev[1..]` (all X).
template <typename X>
static hipError_t vcycle(hipStream_t s, const Grid& g, const MgLevelT<X, double>& top, const MgLevelT<X, X>* lev,
                         int nlev, int nu, const double* spec, Reduce red, CgState* st,
                         const MgLevelT<X, X>* topx = nullptr, FieldView<X> x2 = FieldView<X>{nullptr, 0, 0},
                         FieldView<X> x21 = FieldView<X>{nullptr, 0, 0})
{
    const dim3 blk(BX, BY);
    const int last = nlev - 1;
    const FieldView<X> nobc{nullptr, 0, 0};
    auto rr = [&](const auto& o, const MgLevelT<X, X>& oc) {
        using B = std::decay_t<decltype(*o.b.p)>;
        // coarse levels per block: as many as keep >= 1024 blocks, up to 8
        // (each block recomputes the two fine planes below its first level)
        const int ntx = (oc.nx + RR_CX - 1) / RR_CX, nty = (oc.ny + RR_CY - 1) / RR_CY;
        int rk = 8;
        while (rk > 1 && long(ntx) * nty * ((oc.nz + rk - 1) / rk) < 1024) rk /= 2;
        const int ntz = (oc.nz + rk - 1) / rk;
        hipLaunchKernelGGL((k_mg_rr<X, B, X>), dim3(ntx * nty * ntz), dim3(RR_T), 0, s, o, oc, ntx, nty, rk, st);
    };
    // The smoothing pairs of the way down (red, black) on o, then the
    // residual-restriction to oc.  With xb (the finest level's second
    // vector) each pair is one out-of-place pass (mg_pair), alternating
    // between o.x and xb.  Returns the vector the level's values end in.
    auto pairs_rr = [&](auto o, const MgLevelT<X, X>& oc, FieldView<X> xb) {
        using B = std::decay_t<decltype(*o.b.p)>;
        const FieldView<X> xa = o.x;
        for (int q = 1; q < nu; ++q) {
            if (xb.p) {
                o.x = mg_pair(s, o, o.x.p == xa.p ? xb : xa, 0, st);
            } else {
                hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, 0, st);
                hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, 1, st);
            }
        }
        rr(o, oc);
        return o.x;
    };
    auto down = [&](const auto& o, const MgLevelT<X, X>& oc) {
        using B = std::decay_t<decltype(*o.b.p)>;
        hipLaunchKernelGGL((k_mg_pre<X, B>), pair_grid(o), blk, 0, s, o, nobc, st);
        pairs_rr(o, oc, nobc);
    };
    // The way up: prolongation, then the pairs (black, red), the red of the
    // last dropped on the finest level (k_mg_last does it).  o.x holds the
    // level's values; with xb the pairs alternate between xa and xb as above.
    auto up = [&](auto o, const MgLevelT<X, X>& oc, bool finest, FieldView<X> xa, FieldView<X> xb) {
        using B = std::decay_t<decltype(*o.b.p)>;
        hipLaunchKernelGGL((k_mg_prolong<X, B, X>), dim3((oc.nx + BX - 1) / BX, (o.ny + BY - 1) / BY, o.nz), blk, 0,
                           s, o, oc, st);
        for (int q = 0; q < nu; ++q) {
            // the finest level's last red runs in k_mg_last, or, with xb, in
            // the last pair pass
            const bool pair = !finest || q < nu - 1 || xb.p;
            if (xb.p && pair) {
                o.x = mg_pair(s, o, o.x.p == xa.p ? xb : xa, 1, st);
                continue;
            }
            hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, 1, st);
            if (pair) hipLaunchKernelGGL((k_mg_smooth<X, B>), pair_grid(o), blk, 0, s, o, 0, st);
        }
        return o.x;
    };
    // Where the finest level's values are.  With x2 its sweeps after the
    // first pass are 2 nu - 1 pair passes (nu - 1 down, nu up, the last red
    // among them), each moving the values to the other vector: the first
    // pass writes x2, so that they end in the level's own vector.
    const bool pairs = last > 0 && x2.p;
    FieldView<X> cur = pairs ? x2 : top.x;
    MgLevelT<X, double> top0 = top;
    top0.x = cur;
    if (last > 0 && topx) {
        // the finest level with b converted once by its first pass (a pair
        // pass from x = 0 where there is a second vector)
        if (pairs) mg_pair<X, double, true>(s, top0, x2, 0, st, topx->b);
        else hipLaunchKernelGGL((k_mg_pre<X, double>), pair_grid(top0), blk, 0, s, top0, topx->b, st);
        MgLevelT<X, X> o = *topx;
        o.x = cur;
        cur = pairs_rr(o, lev[1], pairs ? top.x : nobc);
    } else if (last > 0) {
        // with the second vector the first pre-smoothing pair is a pair pass
        // from x = 0 into it (k_mg_pre's values; nothing of x is read)
        if (pairs) mg_pair<X, double, true>(s, top0, x2, 0, st);
        else hipLaunchKernelGGL((k_mg_pre<X, double>), pair_grid(top0), blk, 0, s, top0, nobc, st);
        cur = pairs_rr(top0, lev[1], pairs ? top.x : nobc);
    }
    // Level 1 with its second vector (x21) runs its pairs as level 0 does:
    // the first pass writes x21, and the 2 nu - 1 pair passes after it end
    // in the level's own vector, which level 0's prolongation reads.
    const bool pairs1 = last > 1 && x21.p;
    FieldView<X> cur1 = lev[1].x;
    // levels lb .. last, each of at most MG1_CELLS cells (at least one of
    // them smoothed), run as one single-block launch (k_mg_bottom)
    int lb = last;
    while (lb > 1 && mg_one_block(lev[lb - 1])) --lb;
    const bool bottom = last > 1 && lb < last && last - lb + 1 <= MG1_MAXLEV;
    const int lmid = bottom ? lb : last;
    for (int l = 1; l < lmid; ++l) {
        if (l == 1 && pairs1) {
            MgLevelT<X, X> o = lev[1];
            o.x = x21;
            mg_pair<X, X, true>(s, o, x21, 0, st);
            cur1 = pairs_rr(o, lev[2], lev[1].x);
        } else {
            down(lev[l], lev[l + 1]);
        }
    }
    if (bottom) {
        MgBottom<X> mb{};
        for (int l = lb; l <= last; ++l) mb.L[l - lb] = lev[l];
        mb.n = last - lb + 1;
        mb.nu = nu;
        hipLaunchKernelGGL((k_mg_bottom<X>), dim3(1), dim3(MG1_T), 0, s, mb, spec, st);
    } else if (last > 0) {
        const MgLevelT<X, X>& c = lev[last];
        const long nc = long(c.nx) * c.ny * c.nz;
        if (spec)
            hipLaunchKernelGGL((k_mg_spec<X, X>), dim3(1), dim3(nc <= 64 ? 64 : MGC_T), 0, s, c, spec, st);
        else
            hipLaunchKernelGGL((k_mg_coarse<X>), dim3(1), dim3(nc <= 64 ? 64 : MGC_T), 0, s, c, st);
    } else {
        // one level: the coarsest solve runs on level 0 itself (its b is the
        // FP64 residual only on the FP64 hierarchy, where the types agree)
        if constexpr (std::is_same_v<X, double>) {
            const long nc = long(top.nx) * top.ny * top.nz;
            if (spec)
                hipLaunchKernelGGL((k_mg_spec<X, double>), dim3(1), dim3(nc <= 64 ? 64 : MGC_T), 0, s, top, spec, st);
            else
                hipLaunchKernelGGL((k_mg_coarse<X>), dim3(1), dim3(nc <= 64 ? 64 : MGC_T), 0, s, top, st);
        } else {
            return hipErrorInvalidValue;
        }
    }
    for (int l = lmid - 1; l >= 1; --l) {
        if (l == 1 && pairs1) {
            MgLevelT<X, X> o = lev[1];
            o.x = cur1;
            if (up(o, lev[2], false, x21, lev[1].x).p != lev[1].x.p) return hipErrorInvalidValue;
        } else {
            up(lev[l], lev[l + 1], false, nobc, nobc);
        }
    }
    if (last > 0 && topx) {
        MgLevelT<X, X> o = *topx;
        o.x = cur;
        cur = up(o, lev[1], true, x2, pairs ? top.x : nobc);
    } else if (last > 0) {
        MgLevelT<X, double> o = top;
        o.x = cur;
        cur = up(o, lev[1], true, x2, pairs ? top.x : nobc);
    }
    if (cur.p != top.x.p) return hipErrorInvalidValue;
    // the last red sweep (unless a pair pass did it) and the partials of z.r
    const Tiles tl = cg_tiles(g);
    hipLaunchKernelGGL((k_mg_last<X>), dim3(tl.nblocks), blk, 0, s, top, tl, last > 0 && !pairs ? 1 : 0, red.part,
                       st);
    const double N = double(top.nx) * top.ny * top.nz;
    hipLaunchKernelGGL(k_mg_zf, dim3(1), dim3(256), 0, s, red.part, tl.nblocks, N, st);
    return hipGetLastError();
}

hipError_t launch_mg_vcycle(hipStream_t s, const Grid& g, const MgHierarchy& h, Reduce red, CgState* st)
{
    if (!h.fp32) return vcycle<double>(s, g, h.L[0], h.L, h.nlev, h.nu, h.spec, red, st, nullptr, h.x2, h.x21);
    // the FP32 vectors with L's extents and coefficients
    MgLevel32 lev[MG_MAXLEV];
    for (int l = 0; l < h.nlev; ++l) {
        const MgLevel& o = h.L[l];
        lev[l] = MgLevel32{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[l].x, h.F[l].b};
    }
    const MgLevel& o = h.L[0];
    const MgLevelT<float, double> top{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[0].x, o.b};
    // F[0].b: the FP64 residual in FP32, written by the first pass
    const MgLevel32 topx{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[0].x, h.F[0].b};
    return vcycle<float>(s, g, top, lev, h.nlev, h.nu, h.spec, red, st, h.F[0].b.p ? &topx : nullptr, h.x2f,
                         h.x21f);
}

// ---- the same V-cycle for many cases in lock-step, one launch per kernel
// and level across the cases (group.hpp) ----
namespace {

// rr's grid and coarse levels per block, as vcycle() chooses them
struct RrShape {
    int ntx, nty, rk, nblk;
};
template <typename L>
RrShape rr_shape(const L& oc)
{
    RrShape r;
    r.ntx = (oc.nx + RR_CX - 1) / RR_CX;
    r.nty = (oc.ny + RR_CY - 1) / RR_CY;
    r.rk = 8;
    while (r.rk > 1 && long(r.ntx) * r.nty * ((oc.nz + r.rk - 1) / r.rk) < 1024) r.rk /= 2;
    r.nblk = r.ntx * r.nty * ((oc.nz + r.rk - 1) / r.rk);
    return r;
}

// One case's levels in the cycle's arithmetic X, as launch_mg_vcycle forms them.
template <typename X>
struct CycleView {
    MgLevelT<X, double> top;
    bool hx;               // FP32 finest level with its b converted once (topx)
    MgLevelT<X, X> topx;
    MgLevelT<X, X> lev[MG_MAXLEV];
    int last, nu;
    const double* spec;
    const MgGroupCase* c;
    FieldView<X> x2;  // level 0's second vector, or null: its pairs as one pass each
};

template <typename X>
CycleView<X> cycle_view(const MgGroupCase& gc)
{
    const MgHierarchy& h = *gc.h;
    CycleView<X> v{};
    v.last = h.nlev - 1;
    v.nu = h.nu;
    v.spec = h.spec;
    v.c = &gc;
    if constexpr (std::is_same_v<X, double>) {
        v.top = h.L[0];
        v.hx = false;
        for (int l = 0; l < h.nlev; ++l) v.lev[l] = h.L[l];
    } else {
        for (int l = 0; l < h.nlev; ++l) {
            const MgLevel& o = h.L[l];
            v.lev[l] = MgLevel32{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[l].x, h.F[l].b};
        }
        const MgLevel& o = h.L[0];
        v.top = MgLevelT<float, double>{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[0].x, o.b};
        v.hx = h.F[0].b.p != nullptr;
        v.topx = MgLevel32{o.nx, o.ny, o.nz, o.cx, o.cy, o.cz, o.diag, h.F[0].x, h.F[0].b};
    }
    if constexpr (std::is_same_v<X, double>)
        v.x2 = h.x2;
    else
        v.x2 = h.x2f;
    if (v.last == 0 || v.top.ny % MP_TY != 0) v.x2 = FieldView<X>{nullptr, 0, 0};
    return v;
}

// Level 0's vector after m pair passes of a case whose level 0 runs them:
// its first pass writes the second vector, each pair pass moves to the other.
template <typename X>
FieldView<X> pair_buf(const CycleView<X>& v, int m)
{
    return m % 2 == 0 ? v.x2 : v.top.x;
}

// Pair passes m0 .. m0+count-1 of level 0 (colour f first) for a set of cases.
template <typename X, typename B, typename Get>
void group_pairs(GroupPlan& P, const std::vector<const CycleView<X>*>& vs, int f, int m0, int count, Get&& level_of)
{
    if (vs.empty()) return;
    static const int slots = resident_blocks(k_mg_pair<X, B>, MgPair<X, B>::NT);
    for (int q = 0; q < count; ++q) {
        GroupItems<typename MgPair<X, B>::Args> it;
        for (const CycleView<X>* v : vs) {
            MgLevelT<X, B> o = level_of(*v);
            o.x = pair_buf(*v, m0 + q);
            const int nbx = (o.nx + MP_TX - 1) / MP_TX, nby = o.ny / MP_TY;
            const int ks = march_chunk(nbx * nby, o.nz, std::max(1, slots / int(vs.size())), 32);
            it.add({o, pair_buf(*v, m0 + q + 1), f, nbx, nby, ks, v->c->st}, dim3(nbx * nby * ((o.nz + ks - 1) / ks)));
        }
        P.add<MgPair<X, B>>(it.v);
    }
}

// Smoothing sweeps of one level for a set of cases: colours in the order
// vcycle() runs them (down: nu-1 pairs red, black after k_mg_pre; up: nu
// pairs black, red, the last red dropped on the finest level).
template <typename X, typename B, typename Get>
void group_sweeps(GroupPlan& P, const std::vector<const CycleView<X>*>& vs, int pairs, bool up, bool finest,
                  Get&& level_of)
{
    for (int q = 0; q < pairs; ++q)
        for (int h = 0; h < 2; ++h) {
            const int colour = up ? 1 - h : h;
            if (up && finest && h == 1 && q == pairs - 1) continue;
            GroupItems<typename MgSmooth<X, B>::Args> it;
            for (const CycleView<X>* v : vs) {
                const MgLevelT<X, B>& o = level_of(*v);
                it.add({o, colour, v->c->st}, pair_grid(o));
            }
            P.add<MgSmooth<X, B>>(it.v);
        }
}

template <typename X>
void group_cycle(GroupPlan& P, const std::vector<CycleView<X>>& all)
{
    const FieldView<X> nobc{nullptr, 0, 0};
    int lmax = 0;
    for (const auto& v : all) lmax = std::max(lmax, v.last);
    const int nu = all.front().nu;
    // descent: level l of every case that has a coarser level below it, then
    // the coarsest solve of every case whose coarsest level is l
    for (int l = 0; l <= lmax; ++l) {
        std::vector<const CycleView<X>*> dn, dnx, co;
        for (const auto& v : all) {
            if (l < v.last) (l == 0 && v.hx ? dnx : dn).push_back(&v);
            if (l == v.last) co.push_back(&v);
        }
        if (l == 0) {
            // cases whose level 0 runs pair passes (p) and the others
            std::vector<const CycleView<X>*> dnp, dnxp, dns, dnxs;
            for (const auto* v : dn) (v->x2.p ? dnp : dns).push_back(v);
            for (const auto* v : dnx) (v->x2.p ? dnxp : dnxs).push_back(v);
            GroupItems<typename MgPre<X, double>::Args> pre;
            for (const auto* v : dn) {
                MgLevelT<X, double> t = v->top;
                if (v->x2.p) t.x = pair_buf(*v, 0);
                pre.add({t, nobc, v->c->st}, pair_grid(t));
            }
            for (const auto* v : dnx) {
                MgLevelT<X, double> t = v->top;
                if (v->x2.p) t.x = pair_buf(*v, 0);
                pre.add({t, v->topx.b, v->c->st}, pair_grid(t));
            }
            P.add<MgPre<X, double>>(pre.v);
            group_sweeps<X, double>(P, dns, nu - 1, false, false, [](const CycleView<X>& v) -> const MgLevelT<X, double>& { return v.top; });
            group_sweeps<X, X>(P, dnxs, nu - 1, false, false, [](const CycleView<X>& v) -> const MgLevelT<X, X>& { return v.topx; });
            group_pairs<X, double>(P, dnp, 0, 0, nu - 1, [](const CycleView<X>& v) { return v.top; });
            group_pairs<X, X>(P, dnxp, 0, 0, nu - 1, [](const CycleView<X>& v) { return v.topx; });
            GroupItems<typename MgRr<X, double, X>::Args> rr;
            for (const auto* v : dn) {
                const RrShape r = rr_shape(v->lev[1]);
                MgLevelT<X, double> t = v->top;
                if (v->x2.p) t.x = pair_buf(*v, nu - 1);
                rr.add({t, v->lev[1], r.ntx, r.nty, r.rk, v->c->st}, dim3(r.nblk));
            }
            P.add<MgRr<X, double, X>>(rr.v);
            GroupItems<typename MgRr<X, X, X>::Args> rrx;
            for (const auto* v : dnx) {
                const RrShape r = rr_shape(v->lev[1]);
                MgLevelT<X, X> t = v->topx;
                if (v->x2.p) t.x = pair_buf(*v, nu - 1);
                rrx.add({t, v->lev[1], r.ntx, r.nty, r.rk, v->c->st}, dim3(r.nblk));
            }
            P.add<MgRr<X, X, X>>(rrx.v);
        } else if (!dn.empty()) {
            GroupItems<typename MgPre<X, X>::Args> pre;
            for (const auto* v : dn) pre.add({v->lev[l], nobc, v->c->st}, pair_grid(v->lev[l]));
            P.add<MgPre<X, X>>(pre.v);
            group_sweeps<X, X>(P, dn, nu - 1, false, false, [l](const CycleView<X>& v) -> const MgLevelT<X, X>& { return v.lev[l]; });
            GroupItems<typename MgRr<X, X, X>::Args> rr;
            for (const auto* v : dn) {
                const RrShape r = rr_shape(v->lev[l + 1]);
                rr.add({v->lev[l], v->lev[l + 1], r.ntx, r.nty, r.rk, v->c->st}, dim3(r.nblk));
            }
            P.add<MgRr<X, X, X>>(rr.v);
        }
        if (!co.empty()) {
            if (l == 0) {
                // a one-level hierarchy: the coarsest solve on level 0 (FP64 only)
                if constexpr (std::is_same_v<X, double>) {
                    GroupItems<typename MgSpec<X, double>::Args> sp;
                    GroupItems<typename MgCoarse<X>::Args> cs;
                    for (const auto* v : co) {
                        if (v->spec) sp.add({v->top, v->spec, v->c->st}, dim3(1));
                        else cs.add({v->top, v->c->st}, dim3(1));
                    }
                    P.add<MgSpec<X, double>>(sp.v);
                    P.add<MgCoarse<X>>(cs.v);
                }
            } else {
                GroupItems<typename MgSpec<X, X>::Args> sp;
                GroupItems<typename MgCoarse<X>::Args> cs;
                for (const auto* v : co) {
                    if (v->spec) sp.add({v->lev[l], v->spec, v->c->st}, dim3(1));
                    else cs.add({v->lev[l], v->c->st}, dim3(1));
                }
                P.add<MgSpec<X, X>>(sp.v);
                P.add<MgCoarse<X>>(cs.v);
            }
        }
    }
    // ascent
    for (int l = lmax - 1; l >= 0; --l) {
        std::vector<const CycleView<X>*> up, upx;
        for (const auto& v : all)
            if (l < v.last) (l == 0 && v.hx ? upx : up).push_back(&v);
        if (l == 0) {
            GroupItems<typename MgProlong<X, double, X>::Args> pr;
            for (const auto* v : up) {
                MgLevelT<X, double> t = v->top;
                if (v->x2.p) t.x = pair_buf(*v, nu - 1);
                pr.add({t, v->lev[1], v->c->st}, dim3((v->lev[1].nx + BX - 1) / BX, (t.ny + BY - 1) / BY, t.nz));
            }
            P.add<MgProlong<X, double, X>>(pr.v);
            GroupItems<typename MgProlong<X, X, X>::Args> prx;
            for (const auto* v : upx) {
                MgLevelT<X, X> t = v->topx;
                if (v->x2.p) t.x = pair_buf(*v, nu - 1);
                prx.add({t, v->lev[1], v->c->st}, dim3((v->lev[1].nx + BX - 1) / BX, (t.ny + BY - 1) / BY, t.nz));
            }
            P.add<MgProlong<X, X, X>>(prx.v);
            std::vector<const CycleView<X>*> upp, upxp, ups, upxs;
            for (const auto* v : up) (v->x2.p ? upp : ups).push_back(v);
            for (const auto* v : upx) (v->x2.p ? upxp : upxs).push_back(v);
            group_sweeps<X, double>(P, ups, nu, true, true, [](const CycleView<X>& v) -> const MgLevelT<X, double>& { return v.top; });
            group_sweeps<X, X>(P, upxs, nu, true, true, [](const CycleView<X>& v) -> const MgLevelT<X, X>& { return v.topx; });
            // all nu pairs, the last red among them: 2 nu - 1 passes end in top.x
            group_pairs<X, double>(P, upp, 1, nu - 1, nu, [](const CycleView<X>& v) { return v.top; });
            group_pairs<X, X>(P, upxp, 1, nu - 1, nu, [](const CycleView<X>& v) { return v.topx; });
        } else if (!up.empty()) {
            GroupItems<typename MgProlong<X, X, X>::Args> pr;
            for (const auto* v : up)
                pr.add({v->lev[l], v->lev[l + 1], v->c->st},
                       dim3((v->lev[l + 1].nx + BX - 1) / BX, (v->lev[l].ny + BY - 1) / BY, v->lev[l].nz));
            P.add<MgProlong<X, X, X>>(pr.v);
            group_sweeps<X, X>(P, up, nu, true, false, [l](const CycleView<X>& v) -> const MgLevelT<X, X>& { return v.lev[l]; });
        }
    }
    // the final red sweep with z.r and sum z, and the scalar update
    GroupItems<typename MgLast<X>::Args> la;
    GroupItems<MgZf::Args> zf;
    for (const auto& v : all) {
        const Tiles tl = cg_tiles(v.c->g);
        la.add({v.top, tl, v.last > 0 && !v.x2.p ? 1 : 0, v.c->part, v.c->st}, dim3(tl.nblocks));
        zf.add({v.c->part, tl.nblocks, double(v.top.nx) * v.top.ny * v.top.nz, v.c->st}, dim3(1));
    }
    P.add<MgLast<X>>(la.v);
    P.add<MgZf>(zf.v);
}

}  // namespace

int mg_group_vcycle(GroupPlan& P, const std::vector<MgGroupCase>& cs)
{
    // cases of one precision and sweep count run in lock-step; classes in turn
    for (int fp32 = 0; fp32 < 2; ++fp32)
        for (int nu = 1; nu <= 16; ++nu) {
            std::vector<const MgGroupCase*> sel;
            for (const auto& c : cs)
                if ((c.h->fp32 != 0) == (fp32 != 0) && c.h->nu == nu) sel.push_back(&c);
            if (sel.empty()) continue;
            if (fp32) {
                std::vector<CycleView<float>> v;
                for (const auto* c : sel) {
                    if (c->h->nlev < 2) return -1;  // as launch_mg_vcycle: no one-level FP32 cycle
                    v.push_back(cycle_view<float>(*c));
                }
                group_cycle<float>(P, v);
            } else {
                std::vector<CycleView<double>> v;
                for (const auto* c : sel) v.push_back(cycle_view<double>(*c));
                group_cycle<double>(P, v);
            }
        }
    for (const auto& c : cs)
        if (c.h->nu < 1 || c.h->nu > 16) return -1;
    return 0;
}

// 1-D eigenbases.  Neumann (cell-centred, ends coupled to themselves): mode k
// cos(pi k (2i+1) / 2n), normalised, eigenvalue 2 - 2 cos(pi k / n).
// Periodic: the constant, then cos and sin of 2 pi k j / n for k < n/2, then
// (-1)^j for even n; eigenvalues 0, 2 - 2 cos(2 pi k / n) (twice), 4.
static void mg_neumann_basis(int n, double* Q, double* lam)
{
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k) {
        lam[k] = 2.0 - 2.0 * std::cos(pi * double(k) / double(n));
        const double w = k == 0 ? std::sqrt(1.0 / double(n)) : std::sqrt(2.0 / double(n));
        for (int i = 0; i < n; ++i) Q[i * n + k] = w * std::cos(pi * double(k) * double(2 * i + 1) / (2.0 * double(n)));
    }
}

static void mg_periodic_basis(int n, double* Q, double* lam)
{
    const double pi = 3.14159265358979323846;
    const double w0 = std::sqrt(1.0 / double(n)), w = std::sqrt(2.0 / double(n));
    for (int j = 0; j < n; ++j) Q[j * n] = w0;
    lam[0] = 0.0;
    int col = 1;
    for (int k = 1; 2 * k < n; ++k, col += 2) {
        lam[col] = lam[col + 1] = 2.0 - 2.0 * std::cos(2.0 * pi * double(k) / double(n));
        for (int j = 0; j < n; ++j) {
            const double a = 2.0 * pi * double(k) * double(j) / double(n);
            Q[j * n + col] = w * std::cos(a);
            Q[j * n + col + 1] = w * std::sin(a);
        }
    }
    if (n % 2 == 0) {
        for (int j = 0; j < n; ++j) Q[j * n + col] = (j % 2) ? -w0 : w0;
        lam[col] = 4.0;
    }
}

int mg_spectral_tables(int nx, int ny, int nz, double* tab)
{
    if (nx > MG_SPEC_MAX || ny > MG_SPEC_MAX || nz > MG_SPEC_MAX) return 0;
    const int len = nx * nx + ny * ny + nz * nz + nx + ny + nz;
    if (!tab) return len;
    double* Qx = tab;
    double* Qy = Qx + nx * nx;
    double* Qz = Qy + ny * ny;
    double* lx = Qz + nz * nz;
    double* ly = lx + nx;
    double* lz = ly + ny;
    mg_neumann_basis(nx, Qx, lx);
    mg_periodic_basis(ny, Qy, ly);
    mg_neumann_basis(nz, Qz, lz);
    return len;
}

}  // namespace tpls
