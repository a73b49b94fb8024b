// Crank-Nicolson viscous Helmholtz sweeps: jacobi_iteration_allflux.f90:21-282,
// sor_iteration_allflux.f90:19-377 and the outer-loop boundary conditions of
// main_ns_hybrid.F90:758-880.
//
// Component ranges (comp 0/1/2 = u/v/w): u on i in 1..maxl-2 (the outlet face
// maxl-1 is a boundary condition), k in 0..maxn-2 with the wall ghost
// -2 x0 + x1/3; v on i in 1..maxl-1, same k; w on i in 1..maxl-1,
// k in 1..maxn-2 between the walls w = 0.
//
// No FMA contraction anywhere in this file, including the shared helpers
// (helmholtz(), MuFaces) included below: every update then rounds exactly as
// the reference's Fortran does, operation by operation, and the in-place
// sweeps, the fused pass and the CPU restatement agree bit for bit.
#pragma clang fp contract(off)

#include <algorithm>
#include <memory>
#include <type_traits>

#include "group.hpp"
#include "tiles.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

// Raw viscosity values behind the six face viscosities of one cell: a, b
// and the two extra values of each of the four averaged faces, in the
// averaging order of jacobi_iteration_allflux.f90:75-82 (u), 153-160 (v),
// 231-238 (w).  Loading them apart from the averaging lets the column sweep
// issue the loads of later levels before it needs them.
template <typename C>
struct MuRaw {
    C a, b, q[4][2];
};

template <int COMP, typename C, typename T>
__device__ __forceinline__ MuRaw<C> mu_raw(const FieldView<T>& mu, int i, int j, int k, int P)
{
    const int j1 = wrapj(j + 1, P), j0 = wrapj(j, P), j2 = wrapj(j + 2, P);
    MuRaw<C> r;
    if (COMP == 0) {
        r.a = mu(i, j1, k + 1);
        r.b = mu(i + 1, j1, k + 1);
        r.q[0][0] = mu(i, j2, k + 1), r.q[0][1] = mu(i + 1, j2, k + 1);   // yp
        r.q[1][0] = mu(i, j0, k + 1), r.q[1][1] = mu(i + 1, j0, k + 1);   // ym
        r.q[2][0] = mu(i, j1, k + 2), r.q[2][1] = mu(i + 1, j1, k + 2);   // zp
        r.q[3][0] = mu(i, j1, k), r.q[3][1] = mu(i + 1, j1, k);           // zm
    } else if (COMP == 1) {
        r.a = mu(i, j1, k + 1);
        r.b = mu(i, j0, k + 1);
        r.q[0][0] = mu(i + 1, j1, k + 1), r.q[0][1] = mu(i + 1, j0, k + 1);  // xp
        r.q[1][0] = mu(i - 1, j1, k + 1), r.q[1][1] = mu(i - 1, j0, k + 1);  // xm
        r.q[2][0] = mu(i, j1, k + 2), r.q[2][1] = mu(i, j0, k + 2);          // zp
        r.q[3][0] = mu(i, j1, k), r.q[3][1] = mu(i, j0, k);                  // zm
    } else {
        r.a = mu(i, j1, k);
        r.b = mu(i, j1, k + 1);
        r.q[0][0] = mu(i + 1, j1, k), r.q[0][1] = mu(i + 1, j1, k + 1);  // xp
        r.q[1][0] = mu(i - 1, j1, k), r.q[1][1] = mu(i - 1, j1, k + 1);  // xm
        r.q[2][0] = mu(i, j2, k), r.q[2][1] = mu(i, j2, k + 1);          // yp
        r.q[3][0] = mu(i, j0, k), r.q[3][1] = mu(i, j0, k + 1);          // ym
    }
    return r;
}

// The face viscosities from the raw values, with mu_faces_u/v/w's arithmetic.
template <int COMP, typename C>
__device__ __forceinline__ MuFaces<C> mu_cook(const MuRaw<C>& r)
{
    auto avg = [&](int f) { return (r.a + r.b + r.q[f][0] + r.q[f][1]) * C(0.25); };
    MuFaces<C> m;
    if (COMP == 0) {
        m.xp = r.b, m.xm = r.a, m.yp = avg(0), m.ym = avg(1), m.zp = avg(2), m.zm = avg(3);
    } else if (COMP == 1) {
        m.xp = avg(0), m.xm = avg(1), m.yp = r.a, m.ym = r.b, m.zp = avg(2), m.zm = avg(3);
    } else {
        m.xp = avg(0), m.xm = avg(1), m.yp = avg(2), m.ym = avg(3), m.zp = r.b, m.zm = r.a;
    }
    return m;
}

template <int COMP>
__device__ __forceinline__ int imax_of(int maxl) { return COMP == 0 ? maxl - 2 : maxl - 1; }
template <int COMP>
__device__ __forceinline__ int kmin_of() { return COMP == 2 ? 1 : 0; }

// The RHS of a staged sweep, read once per cell with no halo: in FP64 a
// streaming (non-temporal) load, which leaves L2 to the x and viscosity rows
// that neighbouring tiles share (at 512x128x256 the Jacobi 110.7 -> 106.4
// us, the even colour sweep 105 -> 100 us; in FP32 the Jacobi measured 3 %
// slower, so FP32 loads normally).
template <typename C>
__device__ __forceinline__ C rhs_load(const C* p)
{
    if constexpr (sizeof(C) == 8)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Wall-ghost neighbours in z for u and v (jacobi_iteration_allflux.f90:63-73).
template <int COMP, typename C, typename T>
__device__ __forceinline__ void z_nbrs(const FieldView<T>& x, int i, int j, int k, int maxn, C& zp, C& zm)
{
    if (COMP == 2) {
        zp = x(i, j, k + 1);
        zm = x(i, j, k - 1);
        return;
    }
    zm = (k == 0) ? C(-2) * C(x(i, j, 0)) + C(x(i, j, 1)) / C(3) : C(x(i, j, k - 1));
    zp = (k == maxn - 2) ? C(-2) * C(x(i, j, maxn - 2)) + C(x(i, j, maxn - 3)) / C(3) : C(x(i, j, k + 1));
}

// One cell per thread over tiles of 64 x RY columns, one level per tile.
template <int COMP, int RY>
__device__ __forceinline__ bool cell_of(const Grid& g, int nbx, int nby, const Blk& bk, int& i, int& j, int& k)
{
    const int t = xcd_tile(bk.x, bk.n);
    const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
    i = bx * BX + threadIdx.x;
    j = by * RY + threadIdx.y + 1;
    k = bz + kmin_of<COMP>();
    return j <= g.P;
}

template <int COMP, int RY, typename T>
struct OpJacobi {
    struct Args {
        Grid g;
        int nbx;
        int nby;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> xo;
        FieldView<T> rhs;
        FieldView<T> mu;
        FieldView<T> x;
        int xo_zero;
    };
    static constexpr int THREADS = BX * RY;
    static dim3 block() { return dim3(BX, RY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int RY, typename T>
__device__ __forceinline__ void OpJacobi<COMP, RY, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    int nby = args_.nby;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> xo = args_.xo;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    FieldView<T> x = args_.x;
    int xo_zero = args_.xo_zero;
    using C = T;
    int i, j, k;
    if (!cell_of<COMP, RY>(g, nbx, nby, bk_, i, j, k)) return;
    i += 1;
    if (i > imax_of<COMP>(g.maxl)) return;
    MuFaces<C> m = mu_cook<COMP, C>(mu_raw<COMP, C>(mu, i, j, k, g.P));
    C xp = 0, xm = 0, yp = 0, ym = 0, zp = 0, zm = 0;
    if (!xo_zero) {
        xp = xo(i + 1, j, k);
        xm = xo(i - 1, j, k);
        yp = xo(i, wrapj(j + 1, g.P), k);
        ym = xo(i, wrapj(j - 1, g.P), k);
        z_nbrs<COMP, C>(xo, i, j, k, g.maxn, zp, zm);
    }
    x(i, j, k) = helmholtz<C>(m, ax2, ay2, az2, xp, xm, yp, ym, zp, zm, C(rhs(i, j, k)));
}

template <int COMP, int RY, typename T>
__global__ void __launch_bounds__(BX* RY) k_jacobi(Grid g, int nbx, int nby, T ax2, T ay2, T az2, FieldView<T> xo,
                                                   FieldView<T> rhs, FieldView<T> mu, FieldView<T> x, int xo_zero)
{
    OpJacobi<COMP, RY, T>::run({g, nbx, nby, ax2, ay2, az2, xo, rhs, mu, x, xo_zero}, blk_here());
}

template <int COMP, int RY, typename T>
void OpJacobi<COMP, RY, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_jacobi<COMP, RY, T>), grid, block(), 0, s, a.g, a.nbx, a.nby, a.ax2, a.ay2, a.az2, a.xo, a.rhs, a.mu, a.x, a.xo_zero);
}

// Even iteration_time: 3-D colouring (i+j+k) = 1+flg (mod 2); every cell of a
// colour reads only the other colour, so the sweep is fully parallel.  One
// thread per cell of the colour.
template <int COMP, int RY, typename T>
struct OpSor3d {
    struct Args {
        Grid g;
        int nbx;
        int nby;
        int flg;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> x;
        FieldView<T> rhs;
        FieldView<T> mu;
    };
    static constexpr int THREADS = BX * RY;
    static dim3 block() { return dim3(BX, RY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int RY, typename T>
__device__ __forceinline__ void OpSor3d<COMP, RY, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    int nby = args_.nby;
    int flg = args_.flg;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> x = args_.x;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    using C = T;
    int t, j, k;
    if (!cell_of<COMP, RY>(g, nbx, nby, bk_, t, j, k)) return;
    const int i = ((k + j + flg) & 1) + 1 + 2 * t;
    if (i > imax_of<COMP>(g.maxl)) return;
    MuFaces<C> m = mu_cook<COMP, C>(mu_raw<COMP, C>(mu, i, j, k, g.P));
    C zp, zm;
    z_nbrs<COMP, C>(x, i, j, k, g.maxn, zp, zm);
    C c = x(i, j, k);
    C res = helmholtz<C>(m, ax2, ay2, az2, C(x(i + 1, j, k)), C(x(i - 1, j, k)), C(x(i, wrapj(j + 1, g.P), k)),
                         C(x(i, wrapj(j - 1, g.P), k)), zp, zm, C(rhs(i, j, k)));
    x(i, j, k) = c + C(1.2) * (res - c);
}

template <int COMP, int RY, typename T>
__global__ void __launch_bounds__(BX* RY) k_sor3d(Grid g, int nbx, int nby, int flg, T ax2, T ay2, T az2,
                                                  FieldView<T> x, FieldView<T> rhs, FieldView<T> mu)
{
    OpSor3d<COMP, RY, T>::run({g, nbx, nby, flg, ax2, ay2, az2, x, rhs, mu}, blk_here());
}

template <int COMP, int RY, typename T>
void OpSor3d<COMP, RY, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_sor3d<COMP, RY, T>), grid, block(), 0, s, a.g, a.nbx, a.nby, a.flg, a.ax2, a.ay2, a.az2, a.x, a.rhs, a.mu);
}

// Odd iteration_time: 2-D colouring (i+j) = 1+flg (mod 2) with k ascending,
// i.e. Gauss-Seidel up each column (sor_iteration_allflux.f90:100-146 run on
// one thread).  One thread per column of the colour.  The only dependence
// between levels is the updated x(k-1); every other input of level k (the
// other colour's neighbours, the old x(k+1), RHS, viscosity) is loaded SOR_D
// levels ahead into a register ring, so the march is not serialised on
// memory latency.
constexpr int SOR_D = 4;

template <typename C>
struct SorLevel {
    MuRaw<C> mu;
    C xp, xm, yp, ym, rhs, self, up;
};

template <int COMP, typename C, typename T>
__device__ __forceinline__ void sor_load(SorLevel<C>& L, const FieldView<T>& x, const FieldView<T>& rhs,
                                         const FieldView<T>& mu, int i, int j, int jp, int jm, int k, int k1, int P)
{
    L.mu = mu_raw<COMP, C>(mu, i, j, k, P);
    L.xp = x(i + 1, j, k);
    L.xm = x(i - 1, j, k);
    L.yp = x(i, jp, k);
    L.ym = x(i, jm, k);
    L.rhs = rhs(i, j, k);
    L.self = x(i, j, k);
    // w reads its wall value x(maxn-1) above the top level; u and v have no
    // level above k1 = maxn-2
    L.up = (k < k1 || COMP == 2) ? C(x(i, j, k + 1)) : C(0);
}

template <int COMP, int RY, int D, typename T>
struct OpSor2d {
    struct Args {
        Grid g;
        int flg;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> x;
        FieldView<T> rhs;
        FieldView<T> mu;
    };
    static constexpr int THREADS = BX * RY;
    static dim3 block() { return dim3(BX, RY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int RY, int D, typename T>
__device__ __forceinline__ void OpSor2d<COMP, RY, D, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int flg = args_.flg;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> x = args_.x;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    using C = T;
    constexpr int SOR_D = D;
    // 1-D grid of nbx x nby tiles, XCD-aware: each XCD marches a contiguous
    // band of rows, so the neighbour rows and viscosity rows a tile shares
    // with the next stay in one L2
    const int nbx = (g.maxl / 2 + BX - 1) / BX;
    const int t = xcd_tile(bk_.x, bk_.n);
    const int j = (t / nbx) * RY + threadIdx.y + 1;
    const int i = ((j + flg) & 1) + 1 + 2 * ((t % nbx) * BX + threadIdx.x);
    if (i > imax_of<COMP>(g.maxl) || j > g.P) return;
    const int P = g.P, maxn = g.maxn;
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2;
    SorLevel<C> ring[SOR_D];
#pragma unroll
    for (int d = 0; d < SOR_D; ++d)
        if (k0 + d <= k1) sor_load<COMP, C>(ring[d], x, rhs, mu, i, j, jp, jm, k0 + d, k1, P);
    C prev = COMP == 2 ? C(x(i, j, k0 - 1)) : C(0);  // x(i,j,k-1) after its update
    for (int kb = k0; kb <= k1; kb += SOR_D) {
#pragma unroll
        for (int d = 0; d < SOR_D; ++d) {
            const int k = kb + d;
            if (k > k1) break;
            const SorLevel<C>& L = ring[d];
            const C cur = L.self, nxt = L.up;
            C zp, zm;
            if (COMP == 2) {
                zp = nxt;
                zm = prev;
            } else {
                zm = (k == 0) ? C(-2) * cur + nxt / C(3) : prev;
                zp = (k == maxn - 2) ? C(-2) * cur + prev / C(3) : nxt;
            }
            const C res = helmholtz<C>(mu_cook<COMP, C>(L.mu), ax2, ay2, az2, L.xp, L.xm, L.yp, L.ym, zp, zm, L.rhs);
            const C upd = cur + C(1.2) * (res - cur);
            x(i, j, k) = upd;
            prev = upd;
            if (k + SOR_D <= k1) sor_load<COMP, C>(ring[d], x, rhs, mu, i, j, jp, jm, k + SOR_D, k1, P);
        }
    }
}

template <int COMP, int RY, int D, typename T>
__global__ void __launch_bounds__(BX* RY) k_sor2d(Grid g, int flg, T ax2, T ay2, T az2, FieldView<T> x,
                                                  FieldView<T> rhs, FieldView<T> mu)
{
    OpSor2d<COMP, RY, D, T>::run({g, flg, ax2, ay2, az2, x, rhs, mu}, blk_here());
}

template <int COMP, int RY, int D, typename T>
void OpSor2d<COMP, RY, D, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_sor2d<COMP, RY, D, T>), grid, block(), 0, s, a.g, a.flg, a.ax2, a.ay2, a.az2, a.x, a.rhs, a.mu);
}

// ---------------------------------------------------------------------------
// Both colour sweeps of one SOR iteration in one pass: the sweep flg = 0
// ("red", i+j(+k) odd) and then flg = 1 ("black"), as main_ns_hybrid.F90
// calls do_sor_* twice per iteration (:782-785, 818-821, 854-857).  Out of
// place: xin holds x after the Jacobi step and its boundary conditions and is
// only read; xout receives every updated cell.  Each cell is updated from the
// same inputs, with the same arithmetic, as the two in-place sweeps give it:
//  - a red cell reads its black neighbours before the black sweep: xin;
//  - a black cell reads its red neighbours after the red sweep.  A block
//    computes the red cells of its tile and of a one-cell ring around it
//    (the ring redundantly, as the neighbouring tiles do) and hands them to
//    the black cells through LDS, one level at a time;
//  - the 2-D colouring (odd iteration_time) is Gauss-Seidel up each column,
//    so both colours march k, black one level behind red;
//  - with the 3-D colouring (even iteration_time) a black cell also reads
//    the red cells above and below it, so black trails red by two levels;
//    nothing is sequential in k, and the levels are split into chunks of KS
//    (each chunk recomputing the red level on either side of it).
// A tile is RB_TX = 62 interior columns by TY rows; with the ring it spans
// 64 columns, 32 cells of each colour per row.  The red cells of the TY + 2
// rows take the first (TY+2)*32 lanes and the black cells of the TY interior
// rows the next TY*32, so every wavefront is of one colour.  Cells outside
// the updated range (boundary columns, wall levels) pass their xin value.
constexpr int RB_TX = 62;
constexpr int RB_W = RB_TX + 2;

template <typename C>
struct RbLevel {
    MuRaw<C> mu;
    C self, rhs, xp, xm, yp, ym, zl, zh;
};

// Loads of one cell, all unconditional, so that the level loop keeps a
// uniform control flow and the compiler waits for exactly the loads a level
// uses (a conditional load there makes it drain every load in flight).
// (is, ms) is the stencil position clamped into the updated range and
// (it, mt) the cell itself clamped into the array, so every load stays
// inside the arrays whether or not the lane has a cell there; a lane without
// one discards what it loaded.  zl/zh: x one level below / above ms, within
// the array's levels 0..mtop (ZL: load zl at all).
template <int COMP, bool ZL, typename C, typename T>
__device__ __forceinline__ void rb_load(RbLevel<C>& L, const FieldView<T>& x, const FieldView<T>& rhs,
                                        const FieldView<T>& mu, int is, int it, int j, int jp, int jm, int ms,
                                        int mt, int mtop, int P)
{
    L.mu = mu_raw<COMP, C>(mu, is, j, ms, P);
    L.self = x(it, j, mt);
    L.rhs = rhs(is, j, ms);
    L.xp = x(is + 1, j, ms);
    L.xm = x(is - 1, j, ms);
    L.yp = x(is, jp, ms);
    L.ym = x(is, jm, ms);
    L.zl = ZL ? C(x(is, j, max(ms - 1, 0))) : C(0);
    L.zh = x(is, j, min(ms + 1, mtop));
}

// z neighbours of the cell at level k with current value cur, given the
// values below (lo) and above (hi) at the time of its update; u and v use
// the wall ghost -2 x0 + x1/3 at both ends (jacobi_iteration_allflux.f90:63-73).
// One division, always done, for whichever end the level is at, so that no
// branch surrounds it (the two ends never coincide: maxn >= 6).
template <int COMP, typename C>
__device__ __forceinline__ void rb_z(int k, int maxn, C cur, C lo, C hi, C& zp, C& zm)
{
    if (COMP == 2) {
        zp = hi;
        zm = lo;
        return;
    }
    // the ghost (and its division) only on the wall levels: k is uniform
    // across a wavefront in every caller, so this is a scalar branch
    zm = lo;
    zp = hi;
    if (k == 0 || k == maxn - 2) {
        const C ghost = C(-2) * cur + (k == 0 ? hi : lo) / C(3);
        if (k == 0) zm = ghost;
        if (k == maxn - 2) zp = ghost;
    }
}

template <int COMP, typename C>
__device__ __forceinline__ C rb_update(const RbLevel<C>& L, C ax2, C ay2, C az2, C xp, C xm, C yp, C ym, C zp,
                                       C zm)
{
    const C cur = L.self;
    const C res = helmholtz<C>(mu_cook<COMP, C>(L.mu), ax2, ay2, az2, xp, xm, yp, ym, zp, zm, L.rhs);
    return cur + C(1.2) * (res - cur);
}

// The fused passes stage each level of x and of the viscosity into LDS for
// the whole tile and its halo (rows j0-2 .. j0+TY+1 of x, j0-1 .. j0+TY+2 of
// mu; columns i0-2 .. i0+63): every value is loaded from HBM once per block
// with coalesced loads, a few per lane, instead of the ~20 scattered loads
// per cell a stencil needs, so a register ring of a few values per lane
// keeps many levels in flight.  Columns and rows outside the arrays are
// clamped to the nearest existing ones: only cells that are not updated read
// them, and those discard what they compute.
template <int TY>
struct RbPlane {
    static constexpr int R = TY + 4, W = RB_W + 2, N = R * W;
};

// Viscosity values of one cell from staged planes, in mu_raw's order:
// m(di, dj, dk) returns mu at the cell's column + di, its row wrapj(j + dj)
// and level k + dk.
template <int COMP, typename C, typename M>
__device__ __forceinline__ MuRaw<C> mu_raw_at(const M& m)
{
    MuRaw<C> q;
    if (COMP == 0) {
        q.a = m(0, 1, 1);
        q.b = m(1, 1, 1);
        q.q[0][0] = m(0, 2, 1), q.q[0][1] = m(1, 2, 1);
        q.q[1][0] = m(0, 0, 1), q.q[1][1] = m(1, 0, 1);
        q.q[2][0] = m(0, 1, 2), q.q[2][1] = m(1, 1, 2);
        q.q[3][0] = m(0, 1, 0), q.q[3][1] = m(1, 1, 0);
    } else if (COMP == 1) {
        q.a = m(0, 1, 1);
        q.b = m(0, 0, 1);
        q.q[0][0] = m(1, 1, 1), q.q[0][1] = m(1, 0, 1);
        q.q[1][0] = m(-1, 1, 1), q.q[1][1] = m(-1, 0, 1);
        q.q[2][0] = m(0, 1, 2), q.q[2][1] = m(0, 0, 2);
        q.q[3][0] = m(0, 1, 0), q.q[3][1] = m(0, 0, 0);
    } else {
        q.a = m(0, 1, 0);
        q.b = m(0, 1, 1);
        q.q[0][0] = m(1, 1, 0), q.q[0][1] = m(1, 1, 1);
        q.q[1][0] = m(-1, 1, 0), q.q[1][1] = m(-1, 1, 1);
        q.q[2][0] = m(0, 2, 0), q.q[2][1] = m(0, 2, 1);
        q.q[3][0] = m(0, 0, 0), q.q[3][1] = m(0, 0, 1);
    }
    return q;
}

// Viscosity values of one column marching up its levels, one level per call
// of next(): the planes consecutive levels share are kept in registers, so a
// level reads 6 (u, v) or 5 (w) of its 14 values.  m(di, dj, dk) reads mu at
// the column + di, row wrapj(j + dj), the new level + dk.  raw() gives
// mu_raw's values of the current level.
template <int COMP, typename C>
struct MuColumn {
    C z[3][2];  // u: (i, i+1) on row j1; v: (j1, j0) on column i; planes k, k+1, k+2
    C e[4];     // u: rows j2, j0 at plane k+1 (i, i+1); v: columns i+1, i-1 at plane k+1 (j1, j0)
    C w[2][5];  // w: (i,j1), (i+1,j1), (i-1,j1), (i,j2), (i,j0) on planes k, k+1

    template <typename M>
    __device__ __forceinline__ void next(const M& m)
    {
        if (COMP == 2) {
#pragma unroll
            for (int q = 0; q < 5; ++q) w[0][q] = w[1][q];
            w[1][0] = m(0, 1, 1), w[1][1] = m(1, 1, 1), w[1][2] = m(-1, 1, 1), w[1][3] = m(0, 2, 1),
            w[1][4] = m(0, 0, 1);
            return;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) z[0][q] = z[1][q], z[1][q] = z[2][q];
        if (COMP == 0) {
            z[2][0] = m(0, 1, 2), z[2][1] = m(1, 1, 2);
            e[0] = m(0, 2, 1), e[1] = m(1, 2, 1), e[2] = m(0, 0, 1), e[3] = m(1, 0, 1);
        } else {
            z[2][0] = m(0, 1, 2), z[2][1] = m(0, 0, 2);
            e[0] = m(1, 1, 1), e[1] = m(1, 0, 1), e[2] = m(-1, 1, 1), e[3] = m(-1, 0, 1);
        }
    }
    // the planes below the first level's (they are shifted in by two calls
    // of next(), so the first raw() needs no more)
    template <typename M>
    __device__ __forceinline__ void prime(const M& m)
    {
        if (COMP == 2) {
            w[1][0] = m(0, 1, 1), w[1][1] = m(1, 1, 1), w[1][2] = m(-1, 1, 1), w[1][3] = m(0, 2, 1),
            w[1][4] = m(0, 0, 1);
            return;
        }
        if (COMP == 0) {
            z[1][0] = m(0, 1, 1), z[1][1] = m(1, 1, 1);
            z[2][0] = m(0, 1, 2), z[2][1] = m(1, 1, 2);
        } else {
            z[1][0] = m(0, 1, 1), z[1][1] = m(0, 0, 1);
            z[2][0] = m(0, 1, 2), z[2][1] = m(0, 0, 2);
        }
    }
    __device__ __forceinline__ MuRaw<C> raw() const
    {
        MuRaw<C> r;
        if (COMP == 2) {
            r.a = w[0][0], r.b = w[1][0];
#pragma unroll
            for (int f = 0; f < 4; ++f) r.q[f][0] = w[0][f + 1], r.q[f][1] = w[1][f + 1];
        } else {
            r.a = z[1][0], r.b = z[1][1];
            r.q[0][0] = e[0], r.q[0][1] = e[1];
            r.q[1][0] = e[2], r.q[1][1] = e[3];
            r.q[2][0] = z[2][0], r.q[2][1] = z[2][1];
            r.q[3][0] = z[0][0], r.q[3][1] = z[0][1];
        }
        return r;
    }
};

// Odd iteration_time.  Step s: red lanes update level s, black lanes level
// s-1 from the red plane of level s-1.  Red lanes march their column's red
// cell and black lanes their column's black cell, so `prev` is the updated
// value one level down in the lane's own column, as the in-place sweep has
// it.  Stage t (x plane t, mu plane t+1) is loaded D steps before step t-2
// writes it to LDS; at step s the planes x s-1..s+1 and mu s-1..s+2 are read
// (rings of four and five).  The steps are padded to whole rounds of the
// register ring, with nothing stored outside k0..k1.
// LDS of the fused odd pass: staged x and mu planes and the red planes.  One
// variable per tile shape and type, whatever the component, so that a launch
// running the three components' tiles (k_rb2d3) holds one set per block.
template <int TX, int TY, typename C>
struct Rb2dShared {
    C xs[4][TY + 4][TX + 4];
    C ms[5][TY + 4][TX + 4];
    C plane[3][TY + 2][TX + 2];  // red values of two levels and a spare for lanes with nothing to publish
};
template <int TX, int TY, typename C>
__device__ __forceinline__ Rb2dShared<TX, TY, C>& rb2d_shared()
{
    __shared__ Rb2dShared<TX, TY, C> sh;
    return sh;
}

template <int COMP, int TX, int TY, int D, typename T>
struct OpRb2d {
    struct Args {
        Grid g;
        int nbx;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> xin;
        FieldView<T> xout;
        FieldView<T> rhs;
        FieldView<T> mu;
    };
    static constexpr int THREADS = RbLanes<TX, TY>::NT;
    static dim3 block() { return dim3(RbLanes<TX, TY>::NT); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int TX, int TY, int D, typename T>
__device__ __forceinline__ void OpRb2d<COMP, TX, TY, D, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> xin = args_.xin;
    FieldView<T> xout = args_.xout;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    using C = T;
    using LN = RbLanes<TX, TY>;
    constexpr int SR = TY + 4, SW = TX + 4, SN = SR * SW, WR = TX + 2;  // staged planes; red plane width
    constexpr int NT = LN::NT, NS = (SN + NT - 1) / NT;
    Rb2dShared<TX, TY, C>& sh = rb2d_shared<TX, TY, C>();
    C(&xs)[4][SR][SW] = sh.xs;
    C(&ms)[5][SR][SW] = sh.ms;
    C(&plane)[3][TY + 2][WR] = sh.plane;
    const int t = xcd_tile(bk_.x, bk_.n);
    const int i0 = 1 + (t % nbx) * TX, j0 = 1 + (t / nbx) * TY;
    const int P = g.P, maxn = g.maxn, imax = imax_of<COMP>(g.maxl);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2, mtop = COMP == 2 ? maxn - 1 : maxn - 2;
    // one colour per wavefront (RL is a multiple of 64): a uniform value
    const bool red = __builtin_amdgcn_readfirstlane(threadIdx.x) < LN::RL;
    const int off = red ? 0 : 1;  // the level a lane updates at step s is s - off
    // red lanes: the (TX+2)/2 red cells of each of the TY+2 rows of the tile
    // and its ring; black lanes: the TX/2 black cells of each interior row
    const int q = red ? threadIdx.x : threadIdx.x - LN::RL;
    const int per = red ? LN::RPR : LN::BPR;
    const bool lane = q < (red ? LN::RC : LN::BC);  // lanes past the last cell hold none
    const int r = lane ? q / per + off : 1, h = lane ? q % per : 0;
    const int jj = j0 - 1 + r;
    const int j = wrapj(jj, P);
    const int c = 2 * h + off + ((i0 + jj) & 1);
    const int i = i0 - 1 + c;
    const bool upd = lane && i >= 1 && i <= imax;
    const bool inner = r >= 1 && r <= TY && c >= 1 && c <= TX;
    const int is = min(max(i, 1), imax);
    // neighbour slots in the red plane, clamped (only black lanes inside the
    // tile use them, and theirs are never clamped)
    const int rp = min(r + 1, TY + 1), rm = max(r - 1, 0), cp = min(c + 1, WR - 1), cm = max(c - 1, 0);
    // this thread's staged elements: offsets within a plane of x and of mu
    long ox[NS], om[NS];
    int sl[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int e = min(int(threadIdx.x) + n * NT, SN - 1);
        const int er = e / SW, ec = e % SW;
        sl[n] = e;
        ox[n] = min(max(i0 - 2 + ec, 0), g.maxl) + long(wrapj(j0 - 2 + er, P)) * xin.sy;
        om[n] = min(i0 - 2 + ec, g.maxl + 1) + long(wrapj(j0 - 1 + er, P)) * mu.sy;
    }
    auto stage_load = [&](C* vx, C* vm, int st) {
        const long kx = min(max(st, 0), mtop) * xin.sz, km = min(max(st + 1, 0), maxn) * mu.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            vx[n] = xin.p[ox[n] + kx];
            vm[n] = mu.p[om[n] + km];
        }
    };
    auto stage_store = [&](const C* vx, const C* vm, int st) {
        C* px = &xs[st & 3][0][0];
        C* pm = &ms[(st + 1) % 5][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            px[sl[n]] = vx[n];
            pm[sl[n]] = vm[n];
        }
    };
    auto lev = [&](int s) { return min(max(s - off, k0), k1); };
    // stages k0-1 .. k0+1 into LDS; stages k0+2 .. k0+1+D and the RHS of the
    // first D steps into the register ring
    {
        C vx[NS], vm[NS];
#pragma unroll
        for (int st = -1; st <= 1; ++st) {
            stage_load(vx, vm, k0 + st);
            stage_store(vx, vm, k0 + st);
        }
    }
    C rx[D][NS], rm_[D][NS], rh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        stage_load(rx[d], rm_[d], k0 + 2 + d);
        rh[d] = rhs(is, j, lev(k0 + d));
    }
    // start the loop with nothing in flight: loads pending from before it,
    // in another order than the loop issues them, would make the compiler
    // wait for every load at the top of each round of the ring
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    C prev = xs[(k0 - 1) & 3][r + 1][c + 1];  // w: the wall below level 1 (u, v start with the wall ghost)
    // the lane's viscosity column, primed for the level below its first step's
    MuColumn<COMP, C> mc;
    auto mu_at = [&](int k) {
        // k >= -3 (a black lane three levels below its first), so k + dk + 5 > 0
        return [&, k](int di, int dj, int dk) { return ms[(k + dk + 5) % 5][r + dj][c + 1 + di]; };
    };
    mc.prime(mu_at(k0 - off - 2));
    mc.next(mu_at(k0 - off - 1));
    const int send = k0 + (k1 + 2 - k0 + D - 1) / D * D;
    for (int sb = k0; sb < send; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = sb + d, k = s - off, b = k & 1;
            const bool on = k >= k0 && k <= k1;
            // stage s+2 to LDS (read from step s+1 on), then reload the slot
            stage_store(rx[d], rm_[d], s + 2);
            stage_load(rx[d], rm_[d], s + 2 + D);
            RbLevel<C> L;
            mc.next(mu_at(k));
            L.mu = mc.raw();
            L.self = xs[k & 3][r + 1][c + 1];
            L.zh = xs[(k + 1) & 3][r + 1][c + 1];
            L.rhs = rh[d];
            rh[d] = rhs(is, j, lev(s + D));
            C xp, xm, yp, ym;
            if (red) {  // a wavefront-uniform branch: the other colour's old values
                xp = xs[k & 3][r + 1][c + 2], xm = xs[k & 3][r + 1][c];
                yp = xs[k & 3][r + 2][c + 1], ym = xs[k & 3][r][c + 1];
            } else {  // the red plane of this level
                xp = plane[b][r][cp], xm = plane[b][r][cm], yp = plane[b][rp][c], ym = plane[b][rm][c];
            }
            C zp, zm;
            rb_z<COMP, C>(k, maxn, L.self, prev, L.zh, zp, zm);
            const C nv = rb_update<COMP, C>(L, ax2, ay2, az2, xp, xm, yp, ym, zp, zm);
            const C v = upd ? nv : L.self;
            prev = on ? v : prev;
            plane[red && on && lane ? b : 2][r][c] = v;
            if (on && upd && inner) xout(i, j, k) = v;
            __syncthreads();
        }
    }
}

template <int COMP, int TX, int TY, int D, typename T>
__global__ void __launch_bounds__((RbLanes<TX, TY>::NT)) k_rb2d(Grid g, int nbx, T ax2, T ay2, T az2,
                                                              FieldView<T> xin, FieldView<T> xout, FieldView<T> rhs,
                                                              FieldView<T> mu)
{
    OpRb2d<COMP, TX, TY, D, T>::run({g, nbx, ax2, ay2, az2, xin, xout, rhs, mu}, blk_here());
}

template <int COMP, int TX, int TY, int D, typename T>
void OpRb2d<COMP, TX, TY, D, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_rb2d<COMP, TX, TY, D, T>), grid, block(), 0, s, a.g, a.nbx, a.ax2, a.ay2, a.az2, a.xin, a.xout, a.rhs, a.mu);
}

// The odd pass of u, v and w in one launch: blocks [0, b1) run u's tiles,
// [b1, b2) v's and [b2, ..) w's, with b1 and b2 multiples of 8 so that each
// component's tile b runs on XCD b % 8 as in its own launch.  One pass of
// one component is a single round of blocks (its k march is sequential, so
// it cannot be split into chunks of levels); the three together, at W waves
// per SIMD, keep three blocks on each CU and leave no CU idle at the end of
// one component's pass.
template <int TX, int TY, int D, int W, typename T>
__global__ void __launch_bounds__((RbLanes<TX, TY>::NT)) __attribute__((amdgpu_waves_per_eu(W)))
k_rb2d3(typename OpRb2d<0, TX, TY, D, T>::Args a0, typename OpRb2d<1, TX, TY, D, T>::Args a1,
        typename OpRb2d<2, TX, TY, D, T>::Args a2, int b1, int b2, int n0, int n1, int n2)
{
    const int b = int(blockIdx.x);
    if (b < b1) {
        if (b < n0) OpRb2d<0, TX, TY, D, T>::run(a0, Blk{b, 0, 0, n0});
    } else if (b < b2) {
        if (b - b1 < n1) OpRb2d<1, TX, TY, D, T>::run(a1, Blk{b - b1, 0, 0, n1});
    } else if (b - b2 < n2) {
        OpRb2d<2, TX, TY, D, T>::run(a2, Blk{b - b2, 0, 0, n2});
    }
}

// Even iteration_time, levels kb..ke of one chunk.  Step s: red lanes update
// level s+1, black lanes level s-1 (reading the red planes of s-2, s-1, s).
// The cells of a colour zigzag between the two columns of a lane's pair from
// one level to the next.  Red levels run from kb-1 to ke+1 (where they exist:
// the wall levels of w pass their xin value).  Planes are staged as in
// k_rb2d: at step s the x planes s-1..s+2 and mu planes s-1..s+3 are read
// (rings of five and six) and stage s+3 is written; red planes live in a ring
// of four, so one barrier per step separates every write from the reads of
// that plane.
template <int COMP, int TX, int TY, int D, typename T>
struct OpRb3d {
    struct Args {
        Grid g;
        int nbx;
        int nby;
        int ks;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> xin;
        FieldView<T> xout;
        FieldView<T> rhs;
        FieldView<T> mu;
    };
    static constexpr int THREADS = RbLanes<TX, TY>::NT;
    static dim3 block() { return dim3(RbLanes<TX, TY>::NT); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int TX, int TY, int D, typename T>
__device__ __forceinline__ void OpRb3d<COMP, TX, TY, D, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    int nby = args_.nby;
    int ks = args_.ks;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> xin = args_.xin;
    FieldView<T> xout = args_.xout;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    using C = T;
    using LN = RbLanes<TX, TY>;
    constexpr int SR = TY + 4, SW = TX + 4, SN = SR * SW, WR = TX + 2;
    constexpr int NT = LN::NT, NS = (SN + NT - 1) / NT;
    __shared__ C xs[5][SR][SW];
    __shared__ C ms[6][SR][SW];
    __shared__ C plane[5][TY + 2][WR];  // red values of four levels and a spare
    const int t = xcd_tile(bk_.x, bk_.n);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int i0 = 1 + bx * TX, j0 = 1 + by * TY;
    const int P = g.P, maxn = g.maxn, imax = imax_of<COMP>(g.maxl);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2, mtop = COMP == 2 ? maxn - 1 : maxn - 2;
    const int kb = k0 + bz * ks, ke = min(kb + ks - 1, k1);
    const int rlo = max(kb - 1, 0), rhi = min(ke + 1, mtop);
    const bool red = __builtin_amdgcn_readfirstlane(threadIdx.x) < LN::RL;
    const int off = red ? 0 : 1, dl = red ? 1 : -1;  // the level a lane updates at step s is s + dl
    const int q = red ? threadIdx.x : threadIdx.x - LN::RL;
    const int per = red ? LN::RPR : LN::BPR;
    const bool lane = q < (red ? LN::RC : LN::BC);
    const int r = lane ? q / per + off : 1, h = lane ? q % per : 0;
    const int jj = j0 - 1 + r;
    const int j = wrapj(jj, P);
    // the lane's column slot at level m: red where i+j+m is odd
    auto col = [&](int m) { return 2 * h + off + ((i0 + jj + m) & 1); };
    const int rp = min(r + 1, TY + 1), rm = max(r - 1, 0);
    long ox[NS], om[NS];
    int sl[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int e = min(int(threadIdx.x) + n * NT, SN - 1);
        const int er = e / SW, ec = e % SW;
        sl[n] = e;
        ox[n] = min(max(i0 - 2 + ec, 0), g.maxl) + long(wrapj(j0 - 2 + er, P)) * xin.sy;
        om[n] = min(i0 - 2 + ec, g.maxl + 1) + long(wrapj(j0 - 1 + er, P)) * mu.sy;
    }
    auto stage_load = [&](C* vx, C* vm, int st) {
        const long kx = min(max(st, 0), mtop) * xin.sz, km = min(max(st + 1, 0), maxn) * mu.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            vx[n] = xin.p[ox[n] + kx];
            vm[n] = mu.p[om[n] + km];
        }
    };
    // plane p sits in slot (p + 5) % 5 of xs and (p + 6) % 6 of ms; every
    // plane here is >= -3, so those are positive
    auto stage_store = [&](const C* vx, const C* vm, int st) {
        C* px = &xs[(st + 5) % 5][0][0];
        C* pm = &ms[(st + 1 + 6) % 6][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            px[sl[n]] = vx[n];
            pm[sl[n]] = vm[n];
        }
    };
    // the lane's RHS at step s (its cell's level and column), clamped into the arrays
    auto rhs_at = [&](int s) {
        const int m = min(max(s + dl, k0), k1);
        return C(rhs(min(max(i0 - 1 + col(m), 1), imax), j, m));
    };
    const int s0 = rlo - 1, send = s0 + (ke + 2 - s0 + D - 1) / D * D;
    {
        C vx[NS], vm[NS];
#pragma unroll
        for (int u = -1; u <= 2; ++u) {
            stage_load(vx, vm, s0 + u);
            stage_store(vx, vm, s0 + u);
        }
    }
    C rx[D][NS], rm_[D][NS], rh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        stage_load(rx[d], rm_[d], s0 + 3 + d);
        rh[d] = rhs_at(s0 + d);
    }
    __builtin_amdgcn_s_waitcnt(0);  // as in k_rb2d
    __syncthreads();
    for (int sb = s0; sb < send; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = sb + d, m = s + dl, b = m & 3;
            stage_store(rx[d], rm_[d], s + 3);
            stage_load(rx[d], rm_[d], s + 3 + D);
            const int c = col(m), ii = i0 - 1 + c;
            const int cp = min(c + 1, WR - 1), cm = max(c - 1, 0);
            const bool upd = lane && ii >= 1 && ii <= imax && m >= k0 && m <= k1;
            const bool mine = r >= 1 && r <= TY && c >= 1 && c <= TX && m >= kb && m <= ke;
            const bool on = red ? (m >= rlo && m <= rhi) : (m >= kb && m <= ke);
            RbLevel<C> L;
            L.mu = mu_raw_at<COMP, C>(
                [&](int di, int dj, int dk) { return ms[(m + dk + 6) % 6][r + dj][c + 1 + di]; });
            L.self = xs[(m + 5) % 5][r + 1][c + 1];
            L.rhs = rh[d];
            rh[d] = rhs_at(s + D);
            C xp, xm, yp, ym, lo, hi;
            if (red) {  // the other colour's old values
                const int pm = (m + 5) % 5;
                xp = xs[pm][r + 1][c + 2], xm = xs[pm][r + 1][c];
                yp = xs[pm][r + 2][c + 1], ym = xs[pm][r][c + 1];
                lo = xs[(m + 4) % 5][r + 1][c + 1], hi = xs[(m + 6) % 5][r + 1][c + 1];
            } else {  // the red planes of this level and the two around it
                xp = plane[b][r][cp], xm = plane[b][r][cm], yp = plane[b][rp][c], ym = plane[b][rm][c];
                lo = plane[(m - 1) & 3][r][c], hi = plane[(m + 1) & 3][r][c];
            }
            C zp, zm;
            rb_z<COMP, C>(m, maxn, L.self, lo, hi, zp, zm);
            const C nv = rb_update<COMP, C>(L, ax2, ay2, az2, xp, xm, yp, ym, zp, zm);
            const C v = upd ? nv : L.self;
            plane[red && on && lane ? b : 4][r][c] = v;
            if (on && upd && mine) xout(ii, j, m) = v;
            __syncthreads();
        }
    }
}

template <int COMP, int TX, int TY, int D, typename T>
__global__ void __launch_bounds__((RbLanes<TX, TY>::NT)) k_rb3d(Grid g, int nbx, int nby, int ks, T ax2, T ay2,
                                                                T az2, FieldView<T> xin, FieldView<T> xout,
                                                                FieldView<T> rhs, FieldView<T> mu)
{
    OpRb3d<COMP, TX, TY, D, T>::run({g, nbx, nby, ks, ax2, ay2, az2, xin, xout, rhs, mu}, blk_here());
}

template <int COMP, int TX, int TY, int D, typename T>
void OpRb3d<COMP, TX, TY, D, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_rb3d<COMP, TX, TY, D, T>), grid, block(), 0, s, a.g, a.nbx, a.nby, a.ks, a.ax2, a.ay2, a.az2, a.xin, a.xout, a.rhs, a.mu);
}

// Staged planes of a tile of 64 x TY columns: x on rows j0-1 .. j0+TY and mu
// on rows j0 .. j0+TY+1, both on columns i0-1 .. i0+64, clamped into the
// arrays (a clamped element is read by no updated cell).  Stage t is x plane
// t and mu plane t+1.  NT threads share the loads, NS elements each; LDS
// holds rings of four planes, indexed by plane & 3.
template <typename C, int TY, int NT>
struct Staged {
    static constexpr int R = TY + 2, W = BX + 2, N = R * W, NS = (N + NT - 1) / NT;
    long ox[NS], om[NS];
    int sl[NS];

    __device__ void init(const Grid& g, const FieldView<C>& xo, const FieldView<C>& mu, int i0, int j0, int tid)
    {
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            const int e = min(tid + n * NT, N - 1);
            const int a = e / W, b = e % W;
            sl[n] = e;
            ox[n] = min(max(i0 - 1 + b, 0), g.maxl) + long(wrapj(j0 - 1 + a, g.P)) * xo.sy;
            om[n] = min(i0 - 1 + b, g.maxl + 1) + long(wrapj(j0 + a, g.P)) * mu.sy;
        }
    }
    // x planes clamped to 0..mtop, mu planes to 0..maxn; no x when WX is false
    template <bool WX>
    __device__ void load(C* vx, C* vm, int st, const FieldView<C>& xo, const FieldView<C>& mu, int mtop,
                         int maxn) const
    {
        const long kx = long(min(max(st, 0), mtop)) * xo.sz, km = long(min(max(st + 1, 0), maxn)) * mu.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            if (WX) vx[n] = xo.p[ox[n] + kx];
            vm[n] = mu.p[om[n] + km];
        }
    }
    template <bool WX>
    __device__ void store(const C* vx, const C* vm, int st, C (*xs)[R][W], C (*ms)[R][W]) const
    {
        C* px = &xs[WX ? (st & 3) : 0][0][0];
        C* pm = &ms[(st + 1) & 3][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            if (WX) px[sl[n]] = vx[n];
            pm[sl[n]] = vm[n];
        }
    }
};

// The Jacobi sweep with staged planes.  A block of 64 x TY threads owns
// 64 x TY columns and a chunk of ks levels.  Level k reads x_old on the planes
// k-1..k+1 and mu on k..k+2 over the tile and a one-cell halo: the block
// loads each plane once, with coalesced loads, D stages ahead into
// registers, and writes stage k+2 to LDS at step k.  A few loads per lane and
// level instead of the ~20 scattered ones of k_jacobi, with its arithmetic.
// ZERO: x_old is zero (the first iteration), nothing of it is read.
template <int COMP, int TY, int D, bool ZERO, typename T>
struct OpJacobiSt {
    struct Args {
        Grid g;
        int nbx;
        int nby;
        int ks;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> xo;
        FieldView<T> rhs;
        FieldView<T> mu;
        FieldView<T> x;
    };
    static constexpr int THREADS = BX * TY;
    static dim3 block() { return dim3(BX, TY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int TY, int D, bool ZERO, typename T>
__device__ __forceinline__ void OpJacobiSt<COMP, TY, D, ZERO, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    int nby = args_.nby;
    int ks = args_.ks;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> xo = args_.xo;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    FieldView<T> x = args_.x;
    using C = T;
    using S = Staged<C, TY, BX * TY>;
    constexpr int NS = S::NS;
    __shared__ C xs[ZERO ? 1 : 4][S::R][S::W];
    __shared__ C ms[4][S::R][S::W];
    const int t = xcd_tile(bk_.x, bk_.n);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int i0 = 1 + bx * BX, j0 = 1 + by * TY;
    const int maxn = g.maxn, imax = imax_of<COMP>(g.maxl);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2, mtop = COMP == 2 ? maxn - 1 : maxn - 2;
    const int kb = k0 + bz * ks, ke = min(kb + ks - 1, k1);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx, j = j0 + ty;  // j <= P: the tile rows divide the period
    const bool mine = i <= imax;
    const int ic = min(i, imax);
    S st;
    st.init(g, xo, mu, i0, j0, ty * BX + tx);
    {
        C vx[NS], vm[NS];
#pragma unroll
        for (int u = -1; u <= 1; ++u) {
            st.template load<!ZERO>(vx, vm, kb + u, xo, mu, mtop, maxn);
            st.template store<!ZERO>(vx, vm, kb + u, xs, ms);
        }
    }
    C rx[D][NS], rm_[D][NS], rh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        st.template load<!ZERO>(rx[d], rm_[d], kb + 2 + d, xo, mu, mtop, maxn);
        rh[d] = rhs_load(&rhs(ic, j, min(kb + d, ke)));
    }
    // start the loop with nothing in flight (see k_rb2d)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int kend = kb + (ke - kb + D) / D * D;
    for (int sb = kb; sb < kend; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int k = sb + d;
            st.template store<!ZERO>(rx[d], rm_[d], k + 2, xs, ms);
            st.template load<!ZERO>(rx[d], rm_[d], k + 2 + D, xo, mu, mtop, maxn);
            const C rv = rh[d];
            rh[d] = rhs_load(&rhs(ic, j, min(k + D, ke)));
            const MuRaw<C> mr =
                mu_raw_at<COMP, C>([&](int di, int dj, int dk) { return ms[(k + dk) & 3][ty + dj][tx + 1 + di]; });
            C xp = 0, xm = 0, yp = 0, ym = 0, zp = 0, zm = 0;
            if (!ZERO) {
                const int b = k & 3;
                xp = xs[b][ty + 1][tx + 2];
                xm = xs[b][ty + 1][tx];
                yp = xs[b][ty + 2][tx + 1];
                ym = xs[b][ty][tx + 1];
                rb_z<COMP, C>(k, maxn, C(xs[b][ty + 1][tx + 1]), C(xs[(k - 1) & 3][ty + 1][tx + 1]),
                              C(xs[(k + 1) & 3][ty + 1][tx + 1]), zp, zm);
            }
            const C v = helmholtz<C>(mu_cook<COMP, C>(mr), ax2, ay2, az2, xp, xm, yp, ym, zp, zm, rv);
            if (mine && k <= ke) {
                // FP64: a streaming store (no L2 allocation for a line no later
                // level of this launch reads) measured 3 % faster; FP32 slower
                if constexpr (sizeof(C) == 8)
                    __builtin_nontemporal_store(v, &x(i, j, k));
                else
                    x(i, j, k) = v;
            }
            __syncthreads();
        }
    }
}

template <int COMP, int TY, int D, bool ZERO, typename T>
__global__ void __launch_bounds__(BX* TY) k_jacobi_st(Grid g, int nbx, int nby, int ks, T ax2, T ay2, T az2,
                                                     FieldView<T> xo, FieldView<T> rhs, FieldView<T> mu,
                                                     FieldView<T> x)
{
    OpJacobiSt<COMP, TY, D, ZERO, T>::run({g, nbx, nby, ks, ax2, ay2, az2, xo, rhs, mu, x}, blk_here());
}

template <int COMP, int TY, int D, bool ZERO, typename T>
void OpJacobiSt<COMP, TY, D, ZERO, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_jacobi_st<COMP, TY, D, ZERO, T>), grid, block(), 0, s, a.g, a.nbx, a.nby, a.ks, a.ax2, a.ay2, a.az2, a.xo, a.rhs, a.mu, a.x);
}

// The even-parity SOR colour sweep (3-D colouring, in place) with staged
// planes: a block of 32 x TY threads owns 64 x TY columns, a thread the two
// columns (i0+2tx, i0+2tx+1), of which it updates the one of the sweep's
// colour at each level.  Only that colour is written, and it reads only the
// other colour and its own old value, so loading the planes ahead is safe in
// place.  k_sor3d's arithmetic.
template <int COMP, int TY, int D, typename T>
struct OpSor3dSt {
    struct Args {
        Grid g;
        int nbx;
        int nby;
        int ks;
        int flg;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> x;
        FieldView<T> rhs;
        FieldView<T> mu;
    };
    static constexpr int THREADS = 32 * TY;
    static dim3 block() { return dim3(32 * TY); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int TY, int D, typename T>
__device__ __forceinline__ void OpSor3dSt<COMP, TY, D, T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int nbx = args_.nbx;
    int nby = args_.nby;
    int ks = args_.ks;
    int flg = args_.flg;
    T ax2 = args_.ax2;
    T ay2 = args_.ay2;
    T az2 = args_.az2;
    FieldView<T> x = args_.x;
    FieldView<T> rhs = args_.rhs;
    FieldView<T> mu = args_.mu;
    using C = T;
    using S = Staged<C, TY, 32 * TY>;
    constexpr int NS = S::NS;
    __shared__ C xs[4][S::R][S::W];
    __shared__ C ms[4][S::R][S::W];
    const int t = xcd_tile(bk_.x, bk_.n);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int i0 = 1 + bx * BX, j0 = 1 + by * TY;
    const int maxn = g.maxn, imax = imax_of<COMP>(g.maxl);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2, mtop = COMP == 2 ? maxn - 1 : maxn - 2;
    const int kb = k0 + bz * ks, ke = min(kb + ks - 1, k1);
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    const int j = j0 + ty;
    // the column of this thread's pair that holds the sweep's colour at level k
    auto col = [&](int k) { return 2 * tx + ((k + j + flg) & 1); };
    S st;
    st.init(g, x, mu, i0, j0, threadIdx.x);
    {
        C vx[NS], vm[NS];
#pragma unroll
        for (int u = -1; u <= 1; ++u) {
            st.template load<true>(vx, vm, kb + u, x, mu, mtop, maxn);
            st.template store<true>(vx, vm, kb + u, xs, ms);
        }
    }
    C rx[D][NS], rm_[D][NS], rh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        st.template load<true>(rx[d], rm_[d], kb + 2 + d, x, mu, mtop, maxn);
        const int kd = min(kb + d, ke);
        rh[d] = rhs_load(&rhs(min(i0 + col(kd), imax), j, kd));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int kend = kb + (ke - kb + D) / D * D;
    for (int sb = kb; sb < kend; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int k = sb + d;
            st.template store<true>(rx[d], rm_[d], k + 2, xs, ms);
            st.template load<true>(rx[d], rm_[d], k + 2 + D, x, mu, mtop, maxn);
            const C rv = rh[d];
            const int kn = min(k + D, ke);
            rh[d] = rhs_load(&rhs(min(i0 + col(kn), imax), j, kn));
            const int c = col(k), i = i0 + c;
            const MuRaw<C> mr =
                mu_raw_at<COMP, C>([&](int di, int dj, int dk) { return ms[(k + dk) & 3][ty + dj][c + 1 + di]; });
            const int b = k & 3;
            const C cur = xs[b][ty + 1][c + 1];
            C zp, zm;
            rb_z<COMP, C>(k, maxn, cur, C(xs[(k - 1) & 3][ty + 1][c + 1]), C(xs[(k + 1) & 3][ty + 1][c + 1]), zp,
                          zm);
            const C res = helmholtz<C>(mu_cook<COMP, C>(mr), ax2, ay2, az2, C(xs[b][ty + 1][c + 2]),
                                       C(xs[b][ty + 1][c]), C(xs[b][ty + 2][c + 1]), C(xs[b][ty][c + 1]), zp, zm, rv);
            if (i <= imax && k <= ke) x(i, j, k) = cur + C(1.2) * (res - cur);
            __syncthreads();
        }
    }
}

template <int COMP, int TY, int D, typename T>
__global__ void __launch_bounds__(32 * TY) k_sor3d_st(Grid g, int nbx, int nby, int ks, int flg, T ax2, T ay2,
                                                     T az2, FieldView<T> x, FieldView<T> rhs, FieldView<T> mu)
{
    OpSor3dSt<COMP, TY, D, T>::run({g, nbx, nby, ks, flg, ax2, ay2, az2, x, rhs, mu}, blk_here());
}

template <int COMP, int TY, int D, typename T>
void OpSor3dSt<COMP, TY, D, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_sor3d_st<COMP, TY, D, T>), grid, block(), 0, s, a.g, a.nbx, a.nby, a.ks, a.flg, a.ax2, a.ay2, a.az2, a.x, a.rhs, a.mu);
}

// An even iteration_time's Jacobi sweep, the boundary conditions after it
// and the first colour sweep (flg 0), in one k-march from xs into xd: the
// values main:758-887 leave in xd between its first SOR colour sweep and the
// second, bit for bit.  A block owns a tile of 64 x 8 columns and a chunk of
// levels.  At step m it forms the Jacobi value J(m) of every cell of the tile
// and its one-cell ring (the ring rows wrapped in y), from x and mu planes
// staged as in k_rb2d, and publishes the plane of J values in LDS; then each
// cell of the tile at level m-1 takes its J value or, for the first colour,
// the SOR update from the J values around it: the in-plane ones from LDS, its
// own column's J(m-2) and J(m) from registers, with the viscosity faces,
// 1/diag and RHS it used for J(m-1).  The boundary cells read are those the
// boundary kernel would write: x(0) = J(1); x(imax+1) = J(imax) (v, w); the
// walls of w are 0.  The block also writes those boundary cells.  u's outlet
// column i = imax keeps J here: its boundary cells need the pressure (the
// outlet extrapolation), and k_jsor0_outlet then writes them and updates the
// column's first-colour cells, which read only second-colour neighbours.  xd
// then holds exactly what the Jacobi launch, the boundary kernel and the
// colour sweep leave.  J is formed one level below and above the chunk, so
// chunks overlap only in reads.
template <int COMP, int D, typename T>
struct OpJSor0 {
    static constexpr int TX = 64, TY = 8, RW = TX + 2, RR = TY + 2, NC = RW * RR;
    static constexpr int NT = (NC + 63) / 64 * 64;
    struct Args {
        Grid g;
        int nbx;
        int nby;
        int ks;
        T ax2;
        T ay2;
        T az2;
        FieldView<T> xs;
        FieldView<T> rhs;
        FieldView<T> mu;
        FieldView<T> xd;
    };
    static constexpr int THREADS = NT;
    static dim3 block() { return dim3(NT); }
    static __device__ __forceinline__ void run(const Args& a, const Blk& bk);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <int COMP, int D, typename T>
__device__ __forceinline__ void OpJSor0<COMP, D, T>::run(const Args& a, const Blk& bk)
{
    using C = T;
    const Grid& g = a.g;
    const FieldView<T>& xs = a.xs;
    const FieldView<T>& mu = a.mu;
    const FieldView<T>& rhs = a.rhs;
    const FieldView<T>& xd = a.xd;
    const C ax2 = a.ax2, ay2 = a.ay2, az2 = a.az2;
    constexpr int SR = TY + 4, SW = TX + 4, SN = SR * SW, NS = (SN + NT - 1) / NT;
    __shared__ C xsh[4][SR][SW];
    __shared__ C msh[4][SR][SW];
    __shared__ C jr[2][RR][RW];
    const int t = xcd_tile(bk.x, bk.n);
    const int bx = t % a.nbx, rest = t / a.nbx, by = rest % a.nby, bz = rest / a.nby;
    const int i0 = 1 + bx * TX, j0 = 1 + by * TY;
    const int P = g.P, maxn = g.maxn, maxl = g.maxl, imax = imax_of<COMP>(maxl);
    const int k0 = kmin_of<COMP>(), k1 = maxn - 2, mtop = COMP == 2 ? maxn - 1 : maxn - 2;
    const int kb = k0 + bz * a.ks, ke = min(kb + a.ks - 1, k1);
    // this lane's cell of the tile and ring: (i0-1+c, j0-1+r)
    const int e = int(threadIdx.x);
    const bool lane = e < NC;
    const int r = lane ? e / RW : 1, c = lane ? e % RW : 1;
    const int i = i0 - 1 + c, jj = j0 - 1 + r, j = wrapj(jj, P);
    const bool tile = lane && r >= 1 && r <= TY && c >= 1 && c <= TX && i <= imax;
    const int ic = min(max(i, 1), imax);
    // staged planes, as k_rb2d: x on rows j0-2.., mu on rows j0-1.., both on
    // columns i0-2.., clamped into the arrays
    long ox[NS], om[NS];
    int sl[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int q = min(e + n * NT, SN - 1);
        const int er = q / SW, ec = q % SW;
        sl[n] = q;
        ox[n] = min(max(i0 - 2 + ec, 0), maxl) + long(wrapj(j0 - 2 + er, P)) * xs.sy;
        om[n] = min(i0 - 2 + ec, maxl + 1) + long(wrapj(j0 - 1 + er, P)) * mu.sy;
    }
    auto stage_load = [&](C* vx, C* vm, int st) {
        const long kx = min(max(st, 0), mtop) * xs.sz, km = min(max(st + 1, 0), maxn) * mu.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            vx[n] = xs.p[ox[n] + kx];
            vm[n] = mu.p[om[n] + km];
        }
    };
    // x plane p in slot p & 3, mu plane p in slot p & 3 (stage st: x plane
    // st, mu plane st + 1; planes here are >= -2, so & 3 is their residue)
    auto stage_store = [&](const C* vx, const C* vm, int st) {
        C* px = &xsh[st & 3][0][0];
        C* pm = &msh[(st + 1) & 3][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            px[sl[n]] = vx[n];
            pm[sl[n]] = vm[n];
        }
    };
    auto lev = [&](int m) { return min(max(m, k0), k1); };
    const int m0 = kb - 1;
    {
        C vx[NS], vm[NS];
#pragma unroll
        for (int st = m0 - 1; st <= m0 + 1; ++st) {
            stage_load(vx, vm, st);
            stage_store(vx, vm, st);
        }
    }
    C rx[D][NS], rm_[D][NS], rh[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        stage_load(rx[d], rm_[d], m0 + 2 + d);
        rh[d] = rhs_load(&rhs(ic, j, lev(m0 + d)));
    }
    __builtin_amdgcn_s_waitcnt(0);  // as in k_rb2d
    __syncthreads();
    // this column's J values at m-2 and m-1, and the faces and RHS of m-1
    C jm2 = C(0), jm1 = C(0), rprev = C(0), rdprev = C(0);
    MuFaces<C> fprev{};
    const int send = m0 + (ke + 2 - m0 + D - 1) / D * D;
    for (int sb = m0; sb < send; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int m = sb + d, b = m & 3;
            stage_store(rx[d], rm_[d], m + 2);
            stage_load(rx[d], rm_[d], m + 2 + D);
            const C rv = rh[d];
            rh[d] = rhs_load(&rhs(ic, j, lev(m + D)));
            // J(m) of this cell (the walls of w are 0)
            const MuFaces<C> f = mu_cook<COMP, C>(
                mu_raw_at<COMP, C>([&](int di, int dj, int dk) { return msh[(m + dk) & 3][r + dj][c + 1 + di]; }));
            C zp, zm;
            rb_z<COMP, C>(m, maxn, C(xsh[b][r + 1][c + 1]), C(xsh[(m - 1) & 3][r + 1][c + 1]),
                          C(xsh[(m + 1) & 3][r + 1][c + 1]), zp, zm);
            const C rd = helmholtz_rdiag<C>(f, ax2, ay2, az2);
            C jv = helmholtz_rd<C>(rd, f, ax2, ay2, az2, C(xsh[b][r + 1][c + 2]), C(xsh[b][r + 1][c]),
                                   C(xsh[b][r + 2][c + 1]), C(xsh[b][r][c + 1]), zp, zm, rv);
            if (COMP == 2 && (m <= 0 || m >= maxn - 1)) jv = C(0);
            if (lane) jr[m & 1][r][c] = jv;
            // level m-1: J, or the first colour's update from the J values
            const int k = m - 1;
            if (tile && k >= kb && k <= ke) {
                const C cur = jm1;
                // u's outlet column waits for k_jsor0_outlet, which needs the
                // pressure; v's and w's x(imax+1) is J(imax)
                const bool defer = COMP == 0 && i == imax;
                C v = cur;
                if (!defer && ((i + jj + k) & 1) != 0) {
                    const int q = k & 1;
                    const C xm = i == 1 ? cur : C(jr[q][r][c - 1]);
                    const C xp = i == imax ? cur : C(jr[q][r][c + 1]);
                    C zp2, zm2;
                    rb_z<COMP, C>(k, maxn, cur, jm2, jv, zp2, zm2);
                    const C res = helmholtz_rd<C>(rdprev, fprev, ax2, ay2, az2, xp, xm, C(jr[q][r + 1][c]),
                                                  C(jr[q][r - 1][c]), zp2, zm2, rprev);
                    v = cur + C(1.2) * (res - cur);
                }
                xd(i, jj, k) = v;
                if (i == 1) xd(0, jj, k) = cur;
                if (COMP != 0 && i == imax) xd(imax + 1, jj, k) = cur;
            }
            jm2 = jm1;
            jm1 = jv;
            fprev = f;
            rprev = rv;
            rdprev = rd;
            __syncthreads();
        }
    }
    // the walls of w, for the chunks that hold the first and last levels
    if (COMP == 2 && tile) {
        if (kb == k0) {
            xd(i, jj, 0) = C(0);
            if (i == 1) xd(0, jj, 0) = C(0);
            if (i == imax) xd(imax + 1, jj, 0) = C(0);
        }
        if (ke == k1) {
            xd(i, jj, maxn - 1) = C(0);
            if (i == 1) xd(0, jj, maxn - 1) = C(0);
            if (i == imax) xd(imax + 1, jj, maxn - 1) = C(0);
        }
    }
}

template <int COMP, int D, typename T>
__global__ void __launch_bounds__((OpJSor0<COMP, D, T>::NT)) k_jsor0(typename OpJSor0<COMP, D, T>::Args a)
{
    OpJSor0<COMP, D, T>::run(a, blk_here());
}

template <int COMP, int D, typename T>
void OpJSor0<COMP, D, T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_jsor0<COMP, D, T>), grid, block(), 0, s, a);
}

// u's outlet column after k_jsor0 (which left J there): the boundary cells
// x(L-1) (the pressure extrapolation, as k_bc_uv) and x(L) = 0, then the
// first-colour update of the column's cells with k_sor3d's arithmetic.  One
// thread per (j, k); each writes only its own cells and reads only
// second-colour cells (J values) and its own x(L-1).
template <typename T>
__global__ void k_jsor0_outlet(Grid g, T ax2, T ay2, T az2, T dtdx, FieldView<T> x, FieldView<T> rhs,
                               FieldView<T> mu, FieldView<T> pres)
{
    using C = T;
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = int(blockIdx.y) + 1;
    if (k > g.maxn - 2) return;
    const int L = g.maxl, i = L - 2;
    const C cur = x(i, j, k);
    const C bcr = cur + dtdx * (pres(L, j, k + 1) + pres(L - 2, j, k + 1) - T(2) * pres(L - 1, j, k + 1));
    x(L - 1, j, k) = bcr;
    x(L, j, k) = C(0);
    if (((i + j + k) & 1) == 0) return;
    const MuFaces<C> m = mu_cook<0, C>(mu_raw<0, C>(mu, i, j, k, g.P));
    C zp, zm;
    z_nbrs<0, C>(x, i, j, k, g.maxn, zp, zm);
    const C res = helmholtz<C>(m, ax2, ay2, az2, bcr, C(x(i - 1, j, k)), C(x(i, wrapj(j + 1, g.P), k)),
                               C(x(i, wrapj(j - 1, g.P), k)), zp, zm, C(rhs(i, j, k)));
    x(i, j, k) = cur + C(1.2) * (res - cur);
}

// Outer-loop boundary conditions.
// u: main_ns_hybrid.F90:764-776 (inlet copy, outlet pressure extrapolation
//    with the unshifted j of pres, fictitious face maxl = 0);
// v: main_ns_hybrid.F90:826-836 / jacobi_iteration_allflux.f90:178-189;
// w: main_ns_hybrid.F90:864-879 / jacobi_iteration_allflux.f90:257-279.
template <typename T>
struct OpBcUv {
    struct Args {
        Grid g;
        int comp;
        T dtdx;
        FieldView<T> x;
        FieldView<T> pres;
    };
    static constexpr int THREADS = 64;
    static dim3 block() { return dim3(64); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <typename T>
__device__ __forceinline__ void OpBcUv<T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    int comp = args_.comp;
    T dtdx = args_.dtdx;
    FieldView<T> x = args_.x;
    FieldView<T> pres = args_.pres;
    const int k = bk_.x * blockDim.x + threadIdx.x;
    const int j = bk_.y + 1;
    if (k > g.maxn - 2) return;
    const int L = g.maxl;
    x(0, j, k) = x(1, j, k);
    if (comp == 0) {
        x(L - 1, j, k) = x(L - 2, j, k) + dtdx * (pres(L, j, k + 1) + pres(L - 2, j, k + 1) - T(2) * pres(L - 1, j, k + 1));
        x(L, j, k) = T(0);
    } else {
        x(L, j, k) = x(L - 1, j, k);
    }
}

template <typename T>
__global__ void k_bc_uv(Grid g, int comp, T dtdx, FieldView<T> x, FieldView<T> pres)
{
    OpBcUv<T>::run({g, comp, dtdx, x, pres}, blk_here());
}

template <typename T>
void OpBcUv<T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_bc_uv<T>), grid, block(), 0, s, a.g, a.comp, a.dtdx, a.x, a.pres);
}

// Two independent segments: threads [0, P*(maxn-2)) copy the inlet/outlet
// columns for k in 1..maxn-2; threads [P*(maxn-2), P*(maxn-2) + (maxl+1)*P)
// zero both walls of column (i, j).
template <typename T>
struct OpBcW {
    struct Args {
        Grid g;
        FieldView<T> w;
    };
    static constexpr int THREADS = 256;
    static dim3 block() { return dim3(256); }
    static __device__ __forceinline__ void run(const Args& args_, const Blk& bk_);
    static void launch(dim3 grid, hipStream_t s, const Args& a);
};

template <typename T>
__device__ __forceinline__ void OpBcW<T>::run(const Args& args_, const Blk& bk_)
{
    Grid g = args_.g;
    FieldView<T> w = args_.w;
    const long q = bk_.x * (long)blockDim.x + threadIdx.x;
    const int L = g.maxl, n = g.maxn, P = g.P;
    const long n1 = long(P) * (n - 2), n2 = long(L + 1) * P;
    if (q < n1) {
        const int k = int(q % (n - 2)) + 1, j = int(q / (n - 2)) + 1;
        w(0, j, k) = w(1, j, k);
        w(L, j, k) = w(L - 1, j, k);
    } else if (q < n1 + n2) {
        const long r = q - n1;
        const int i = int(r % (L + 1)), j = int(r / (L + 1)) + 1;
        w(i, j, 0) = T(0);
        w(i, j, n - 1) = T(0);
    }
}

template <typename T>
__global__ void k_bc_w(Grid g, FieldView<T> w)
{
    OpBcW<T>::run({g, w}, blk_here());
}

template <typename T>
void OpBcW<T>::launch(dim3 grid, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL((k_bc_w<T>), grid, block(), 0, s, a.g, a.w);
}

static inline int nz_of(int comp, const Grid& g) { return comp == 2 ? g.maxn - 2 : g.maxn - 1; }

// Launch shapes selectable at run time (tile rows 1, 2, 4, 8; odd-sweep
// prefetch depth 4 or 8): the autotuner picks among them per grid.
template <int COMP, typename T, typename F>
static hipError_t with_rows(int rows, F&& f)
{
    switch (rows) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 8: return f(std::integral_constant<int, 8>{});
    default: return f(std::integral_constant<int, 4>{});
    }
}

// The launchers choose each sweep's kernel and launch shape from the case's
// tiling and hand it to an emitter: DirectEmit launches it on a stream; the
// grouped step's emitter collects it with the other cases' (group.hpp).
// `slots` scales the resident-block count the march chunks are sized for
// (a grouped launch shares the device among its cases).
struct DirectEmit {
    hipStream_t s;
    int slots(int dev_slots) const { return dev_slots; }
    template <class Op>
    hipError_t go(const typename Op::Args& a, dim3 grid)
    {
        Op::launch(grid, s, a);
        return hipGetLastError();
    }
};

template <typename T, class E>
hipError_t jacobi_emit(E& e, const Grid& g, const Phys& ph, int comp, FieldView<T> xold, FieldView<T> rhs,
                       FieldView<T> mu, FieldView<T> x, bool xold_is_zero, const Tiling& tl)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0);
    const int z = xold_is_zero ? 1 : 0;
    if (tl.jst) {
        // staged sweep: tile rows that divide the period, chunks of levels that fill
        // the last round of resident blocks
        const int ty = (tl.jst == 4 || tl.jst == 8 || tl.jst == 16) && g.P % tl.jst == 0 ? tl.jst : 2;
        auto go_st = [&](auto comp_c, auto ty_c, auto zero_c) {
            constexpr int CP = decltype(comp_c)::value, TY = decltype(ty_c)::value;
            constexpr bool ZERO = decltype(zero_c)::value;
            // 16-row tiles (16 waves) load 4 levels ahead, within 128 registers
            constexpr int D = TY == 16 ? 4 : 8;
            const int imax = CP == 0 ? g.maxl - 2 : g.maxl - 1, nz = nz_of(CP, g);
            const int nbx = (imax + BX - 1) / BX, nby = g.P / TY;
            static const int slots = resident_blocks(k_jacobi_st<CP, TY, D, ZERO, T>, BX * TY);
            const int ks = tl.ks > 0 ? std::min(tl.ks, nz) : march_chunk(nbx * nby, nz, e.slots(slots), 32);
            const int nbz = (nz + ks - 1) / ks;
            return e.template go<OpJacobiSt<CP, TY, D, ZERO, T>>({g, nbx, nby, ks, a2, b2, c2, xold, rhs, mu, x},
                                                               dim3(nbx * nby * nbz));
        };
        auto by_ty = [&](auto comp_c, auto zero_c) {
            if (ty == 16) return go_st(comp_c, std::integral_constant<int, 16>{}, zero_c);
            if (ty == 8) return go_st(comp_c, std::integral_constant<int, 8>{}, zero_c);
            if (ty == 4) return go_st(comp_c, std::integral_constant<int, 4>{}, zero_c);
            return go_st(comp_c, std::integral_constant<int, 2>{}, zero_c);
        };
        auto by_zero = [&](auto comp_c) {
            return z ? by_ty(comp_c, std::true_type{}) : by_ty(comp_c, std::false_type{});
        };
        if (comp == 0) return by_zero(std::integral_constant<int, 0>{});
        if (comp == 1) return by_zero(std::integral_constant<int, 1>{});
        return by_zero(std::integral_constant<int, 2>{});
    }
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        return with_rows<CP, T>(tl.rows, [&](auto ry) {
            constexpr int RY = decltype(ry)::value;
            const int nbx = (g.maxl - 1 + BX - 1) / BX, nby = (g.P + RY - 1) / RY;
            return e.template go<OpJacobi<CP, RY, T>>({g, nbx, nby, a2, b2, c2, xold, rhs, mu, x, z},
                                                      dim3(nbx * nby * nz_of(CP, g)));
        });
    };
    if (comp == 0) return go(std::integral_constant<int, 0>{});
    if (comp == 1) return go(std::integral_constant<int, 1>{});
    return go(std::integral_constant<int, 2>{});
}

template <typename T>
hipError_t launch_jacobi(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xold, FieldView<T> rhs,
                         FieldView<T> mu, FieldView<T> x, bool xold_is_zero, const Tiling& tl)
{
    DirectEmit e{s};
    return jacobi_emit<T>(e, g, ph, comp, xold, rhs, mu, x, xold_is_zero, tl);
}

template <typename T, class E>
hipError_t sor_emit(E& e, const Grid& g, const Phys& ph, int comp, FieldView<T> x, FieldView<T> rhs, FieldView<T> mu,
                    int flg, int iteration_time, const Tiling& tl)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0);
    const int half = (g.maxl - 1 + 1) / 2;
    const int nbx = (half + BX - 1) / BX;
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        if (iteration_time % 2 == 0 && tl.jst) {
            // staged planes, as the Jacobi sweep: rows that divide the period, chunks of
            // levels sized to the resident blocks
            const int ty = (tl.jst == 4 || tl.jst == 8 || tl.jst == 16) && g.P % tl.jst == 0 ? tl.jst : 2;
            const int imax = CP == 0 ? g.maxl - 2 : g.maxl - 1, nz = nz_of(CP, g);
            const int nbx3 = (imax + BX - 1) / BX;
            auto go_st = [&](auto ty_c) {
                constexpr int TY = decltype(ty_c)::value;
                const int nby = g.P / TY;
                static const int slots = resident_blocks(k_sor3d_st<CP, TY, 8, T>, 32 * TY);
                const int ks = tl.ks > 0 ? std::min(tl.ks, nz) : march_chunk(nbx3 * nby, nz, e.slots(slots), 32);
                const int nbz = (nz + ks - 1) / ks;
                return e.template go<OpSor3dSt<CP, TY, 8, T>>({g, nbx3, nby, ks, flg, a2, b2, c2, x, rhs, mu},
                                                             dim3(nbx3 * nby * nbz));
            };
            if (ty == 16) return go_st(std::integral_constant<int, 16>{});
            if (ty == 8) return go_st(std::integral_constant<int, 8>{});
            if (ty == 4) return go_st(std::integral_constant<int, 4>{});
            return go_st(std::integral_constant<int, 2>{});
        }
        if (iteration_time % 2 == 0) {
            return with_rows<CP, T>(tl.rows, [&](auto ry) {
                constexpr int RY = decltype(ry)::value;
                const int nby = (g.P + RY - 1) / RY;
                return e.template go<OpSor3d<CP, RY, T>>({g, nbx, nby, flg, a2, b2, c2, x, rhs, mu},
                                                         dim3(nbx * nby * nz_of(CP, g)));
            });
        }
        return with_rows<CP, T>(tl.rows_odd, [&](auto ry) {
            constexpr int RY = decltype(ry)::value;
            const dim3 grd(nbx * ((g.P + RY - 1) / RY));
            if (tl.depth_odd == 8)
                return e.template go<OpSor2d<CP, RY, 8, T>>({g, flg, a2, b2, c2, x, rhs, mu}, grd);
            return e.template go<OpSor2d<CP, RY, 4, T>>({g, flg, a2, b2, c2, x, rhs, mu}, grd);
        });
    };
    if (comp == 0) return go(std::integral_constant<int, 0>{});
    if (comp == 1) return go(std::integral_constant<int, 1>{});
    return go(std::integral_constant<int, 2>{});
}

template <typename T>
hipError_t launch_sor(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> x, FieldView<T> rhs,
                      FieldView<T> mu, int flg, int iteration_time, const Tiling& tl)
{
    DirectEmit e{s};
    return sor_emit<T>(e, g, ph, comp, x, rhs, mu, flg, iteration_time, tl);
}

// Tile rows of a fused launch: the requested count where it divides the
// period, else 2 (the period is even).
static inline int rb_rows(int want, int P) { return (want == 4 || want == 8) && P % want == 0 ? want : 2; }
constexpr int RB_KS = 32;  // levels per chunk of the 3-D colouring pass

template <typename T, class E>
hipError_t sor_pair_emit(E& e, const Grid& g, const Phys& ph, int comp, FieldView<T> xin, FieldView<T> xout,
                         FieldView<T> rhs, FieldView<T> mu, int iteration_time, const Tiling& tl)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0);
    const int imax = comp == 0 ? g.maxl - 2 : g.maxl - 1;
    const bool odd = iteration_time % 2 != 0;
    const int ty = rb_rows(odd ? tl.rb_rows_odd : tl.rb_rows, g.P);
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        auto launch = [&](auto ty_c) {
            constexpr int TY = decltype(ty_c)::value;
            const int nby = g.P / TY;
            if (odd) {
                // tiles of 8 rows are 32 columns wide (16 x 16 tiles fill the
                // chip at 512 x 128), the others 62 (a 64-column row with its
                // ring)
                constexpr int TX = TY == 8 ? 32 : 62;
                const int nbx2 = (imax + TX - 1) / TX;
                return e.template go<OpRb2d<CP, TX, TY, 8, T>>({g, nbx2, a2, b2, c2, xin, xout, rhs, mu},
                                                              dim3(nbx2 * nby));
            }
            constexpr int TX = TY == 8 ? 32 : 62;
            const int nbx3 = (imax + TX - 1) / TX, nk = (nz_of(CP, g) + RB_KS - 1) / RB_KS;
            return e.template go<OpRb3d<CP, TX, TY, 8, T>>({g, nbx3, nby, RB_KS, a2, b2, c2, xin, xout, rhs, mu},
                                                          dim3(nbx3 * nby * nk));
        };
        if (ty == 8) return launch(std::integral_constant<int, 8>{});
        if (ty == 4) return launch(std::integral_constant<int, 4>{});
        return launch(std::integral_constant<int, 2>{});
    };
    if (comp == 0) return go(std::integral_constant<int, 0>{});
    if (comp == 1) return go(std::integral_constant<int, 1>{});
    return go(std::integral_constant<int, 2>{});
}

template <typename T>
hipError_t launch_sor_pair(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xin, FieldView<T> xout,
                           FieldView<T> rhs, FieldView<T> mu, int iteration_time, const Tiling& tl)
{
    DirectEmit e{s};
    return sor_pair_emit<T>(e, g, ph, comp, xin, xout, rhs, mu, iteration_time, tl);
}

// sor_pair_emit of u, v and w at one odd iteration_time as one launch
// (k_rb2d3).  8-row tiles only: hipErrorInvalidValue where sor_pair_emit
// would not take them (sor_pair3_fits).
template <typename T>
hipError_t launch_sor_pair3(hipStream_t s, const Grid& g, const Phys& ph, const FieldView<T>* xin,
                            const FieldView<T>* xout, const FieldView<T>* rhs, FieldView<T> mu, const Tiling& tl)
{
    if (!sor_pair3_fits(g, tl)) return hipErrorInvalidValue;
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), bb = T(ay / 2.0), c2 = T(az / 2.0);
    constexpr int TX = 32, TY = 8;
    const int nby = g.P / TY;
    const int n0 = (g.maxl - 2 + TX - 1) / TX * nby, n12 = (g.maxl - 1 + TX - 1) / TX * nby;
    const int b1 = (n0 + GROUP_ALIGN - 1) / GROUP_ALIGN * GROUP_ALIGN;
    const int b2 = b1 + (n12 + GROUP_ALIGN - 1) / GROUP_ALIGN * GROUP_ALIGN;
    const int nb0 = n0 / nby, nb12 = n12 / nby;
    auto go = [&](auto d_c, auto w_c) {
        constexpr int D = decltype(d_c)::value, W = decltype(w_c)::value;
        hipLaunchKernelGGL((k_rb2d3<TX, TY, D, W, T>), dim3(b2 + n12), dim3(RbLanes<TX, TY>::NT), 0, s,
                           typename OpRb2d<0, TX, TY, D, T>::Args{g, nb0, a2, bb, c2, xin[0], xout[0], rhs[0], mu},
                           typename OpRb2d<1, TX, TY, D, T>::Args{g, nb12, a2, bb, c2, xin[1], xout[1], rhs[1], mu},
                           typename OpRb2d<2, TX, TY, D, T>::Args{g, nb12, a2, bb, c2, xin[2], xout[2], rhs[2], mu},
                           b1, b2, n0, n12, n12);
        return hipGetLastError();
    };
    // 4 waves per SIMD (three 5-wave blocks per CU) with the deepest
    // prefetch that fits 128 registers unspilled: 2 levels in FP64, 4 in
    // FP32.  Measured at 512 x 128 x 256 (one launch against the three
    // one-component passes): FP64 500 against 520 us, FP32 step 17.3 against
    // 17.9 ms; 4 levels at 3 waves per SIMD, or 8 (FP32) / 4 (FP64) at 4
    // waves with spills, were slower.
    if (sizeof(T) == 8) return go(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{});
    return go(std::integral_constant<int, 4>{}, std::integral_constant<int, 4>{});
}

template <typename T, class E>
hipError_t jsor0_emit(E& e, const Grid& g, const Phys& ph, int comp, FieldView<T> xs, FieldView<T> xd,
                      FieldView<T> rhs, FieldView<T> mu, FieldView<T> pres, const Tiling& tl)
{
    if (!jsor0_fits(g)) return hipErrorInvalidValue;
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0), dtdx = T(ph.dt / ph.dx);
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        // FP64 loads 4 levels ahead: 8 would not fit the 168 registers of
        // an 11-wave block
        constexpr int D = sizeof(T) == 8 ? 4 : 8;
        using Op = OpJSor0<CP, D, T>;
        const int imax = CP == 0 ? g.maxl - 2 : g.maxl - 1, nz = nz_of(CP, g);
        const int nbx = (imax + Op::TX - 1) / Op::TX, nby = g.P / Op::TY;
        static const int slots = resident_blocks(k_jsor0<CP, D, T>, Op::NT);
        const int ks = tl.ks > 0 ? std::min(tl.ks, nz) : march_chunk(nbx * nby, nz, e.slots(slots), 32);
        const int nbz = (nz + ks - 1) / ks;
        hipError_t rc = e.template go<Op>({g, nbx, nby, ks, a2, b2, c2, xs, rhs, mu, xd},
                                          dim3(nbx * nby * nbz));
        if (rc == hipSuccess && CP == 0) {
            hipLaunchKernelGGL(k_jsor0_outlet<T>, dim3((g.maxn - 1 + 63) / 64, g.P), dim3(64), 0, e.s, g, a2, b2, c2,
                               dtdx, xd, rhs, mu, pres);
            rc = hipGetLastError();
        }
        return rc;
    };
    if (comp == 0) return go(std::integral_constant<int, 0>{});
    if (comp == 1) return go(std::integral_constant<int, 1>{});
    return go(std::integral_constant<int, 2>{});
}

template <typename T>
hipError_t launch_jsor0(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xs, FieldView<T> xd,
                        FieldView<T> rhs, FieldView<T> mu, FieldView<T> pres, const Tiling& tl)
{
    DirectEmit e{s};
    return jsor0_emit<T>(e, g, ph, comp, xs, xd, rhs, mu, pres, tl);
}

template <typename T, class E>
hipError_t helmholtz_bc_emit(E& e, const Grid& g, const Phys& ph, int comp, FieldView<T> x, FieldView<T> pres)
{
    if (comp == 2) {
        const long n = long(g.P) * (g.maxn - 2) + long(g.maxl + 1) * g.P;
        return e.template go<OpBcW<T>>({g, x}, dim3(unsigned((n + 255) / 256)));
    }
    return e.template go<OpBcUv<T>>({g, comp, T(ph.dt / ph.dx), x, pres}, dim3((g.maxn - 1 + 63) / 64, g.P));
}

template <typename T>
hipError_t launch_helmholtz_bc(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> x,
                               FieldView<T> pres)
{
    DirectEmit e{s};
    return helmholtz_bc_emit<T>(e, g, ph, comp, x, pres);
}

// ---- the Helmholtz outer loops of many cases in lock-step (group.hpp) ----
// Each case's launches are the ones its own enqueue_pre makes (the same
// launchers, with the launch shapes below: every shape computes the same
// values); the emitter collects one sub-step of every case and appends it as
// one grouped launch per kernel variant.
namespace {

struct CollectEmit {
    struct Base {
        virtual ~Base() = default;
        virtual void flush(GroupPlan& P) = 0;
    };
    template <class Op>
    struct Table : Base {
        GroupItems<typename Op::Args> it;
        void flush(GroupPlan& P) override { P.add<Op>(it.v); }
    };
    int share = 1;
    std::vector<std::pair<const void*, std::unique_ptr<Base>>> tabs;
    int slots(int dev_slots) const { return std::max(1, dev_slots / share); }
    template <class Op>
    hipError_t go(const typename Op::Args& a, dim3 grid)
    {
        static const char key = 0;  // one address per kernel variant
        Table<Op>* t = nullptr;
        for (auto& e : tabs)
            if (e.first == &key) t = static_cast<Table<Op>*>(e.second.get());
        if (!t) {
            tabs.emplace_back(&key, std::make_unique<Table<Op>>());
            t = static_cast<Table<Op>*>(tabs.back().second.get());
        }
        t->it.add(a, grid);
        return hipSuccess;
    }
    void flush(GroupPlan& P)
    {
        for (auto& e : tabs) e.second->flush(P);
        tabs.clear();
    }
};

template <typename T>
int group_component(GroupPlan& P, const std::vector<HhGroupCase<T>>& cs, int q, bool diag)
{
    const size_t n = cs.size();
    std::vector<int> src(n), dst(n);
    int its = 0;
    for (size_t k = 0; k < n; ++k) {
        int others[2], m = 0;
        for (int b = 0; b < 3; ++b)
            if (b != cs[k].i2[q]) others[m++] = b;
        src[k] = others[0];
        dst[k] = others[1];
        its = std::max(its, cs[k].mom[q] / 2);
    }
    CollectEmit E;
    E.share = int(n);
    // The grouped launches use staged sweeps with long marches (chunks sized
    // for a share of the device) and the fused odd pass in 8-row tiles
    // wherever the period allows: measured on the 64-case FP32 batch, 94.5 ms
    // per step against 105.6 ms with each case's own (small-grid) shapes,
    // whose one-level blocks leave a grouped launch bound by block dispatch.
    std::vector<Tiling> tls(n);
    for (size_t k = 0; k < n; ++k) {
        Tiling t = cs[k].tl;
        const int P = cs[k].g.P;
        t.jst = P % 8 == 0 ? 8 : (P % 4 == 0 ? 4 : 2);
        t.fused = 1;  // odd iteration_time fused, even in two colour sweeps
        t.rb_rows_odd = 8;
        t.ks = 0;
        tls[k] = t;
    }
    for (int it = 1; it <= its; ++it) {
        std::vector<size_t> on;
        for (size_t k = 0; k < n; ++k)
            if (it <= cs[k].mom[q] / 2) on.push_back(k);
        auto fused = [&](size_t k) { return ((tls[k].fused >> (it % 2 ? 0 : 1)) & 1) != 0; };
        auto X = [&](size_t k, int b) { return cs[k].x[q][b]; };
        auto XJ = [&](size_t k) { return fused(k) ? X(k, cs[k].i2[q]) : X(k, dst[k]); };
        for (size_t k : on) {
            const HhGroupCase<T>& c = cs[k];
            if (int rc = jacobi_emit<T>(E, c.g, c.ph, q, X(k, src[k]), c.rhs[q], c.mu, XJ(k), it == 1, tls[k])) return rc;
        }
        E.flush(P);
        for (size_t k : on) helmholtz_bc_emit<T>(E, cs[k].g, cs[k].ph, q, XJ(k), cs[k].pres);
        E.flush(P);
        // the fused pass, or the first of the two colour sweeps
        for (size_t k : on) {
            const HhGroupCase<T>& c = cs[k];
            if (fused(k)) sor_pair_emit<T>(E, c.g, c.ph, q, XJ(k), X(k, dst[k]), c.rhs[q], c.mu, it, tls[k]);
            else sor_emit<T>(E, c.g, c.ph, q, X(k, dst[k]), c.rhs[q], c.mu, 0, it, tls[k]);
        }
        E.flush(P);
        for (size_t k : on)
            if (!fused(k)) sor_emit<T>(E, cs[k].g, cs[k].ph, q, X(k, dst[k]), cs[k].rhs[q], cs[k].mu, 1, it, tls[k]);
        E.flush(P);
        for (size_t k : on) helmholtz_bc_emit<T>(E, cs[k].g, cs[k].ph, q, X(k, dst[k]), cs[k].pres);
        E.flush(P);
        for (size_t k : on) std::swap(src[k], dst[k]);
    }
    if (diag) {
        // get_difference(x3, x3_old), as enqueue_pre
        std::vector<const HhGroupCase<T>*> on;
        std::vector<FieldView<T>> a, b;
        for (size_t k = 0; k < n; ++k)
            if (cs[k].mom[q] / 2 >= 2 && cs[k].part) {
                on.push_back(&cs[k]);
                a.push_back(cs[k].x[q][src[k]]);
                b.push_back(cs[k].x[q][dst[k]]);
            }
        group_difference<T>(P, on, a, b, q);
    }
    return 0;
}

}  // namespace

template <typename T>
int hh_group_loops(GroupPlan& P, const std::vector<HhGroupCase<T>>& cs, bool diag)
{
    if (cs.empty()) return 0;
    for (int q = 0; q < 3; ++q)
        if (int rc = group_component<T>(P, cs, q, diag)) return rc;
    return 0;
}
template int hh_group_loops<double>(GroupPlan&, const std::vector<HhGroupCase<double>>&, bool);
template int hh_group_loops<float>(GroupPlan&, const std::vector<HhGroupCase<float>>&, bool);

#define TPLS_INST_HH(T)                                                                                              \
    template hipError_t launch_jacobi<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,     \
                                         FieldView<T>, FieldView<T>, bool, const Tiling&);                           \
    template hipError_t launch_sor<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,        \
                                      FieldView<T>, int, int, const Tiling&);                                        \
    template hipError_t launch_sor_pair<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,   \
                                           FieldView<T>, FieldView<T>, int, const Tiling&);                          \
    template hipError_t launch_helmholtz_bc<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>); \
    template hipError_t launch_sor_pair3<T>(hipStream_t, const Grid&, const Phys&, const FieldView<T>*,              \
                                            const FieldView<T>*, const FieldView<T>*, FieldView<T>, const Tiling&); \
    template hipError_t launch_jsor0<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,      \
                                        FieldView<T>, FieldView<T>, FieldView<T>, const Tiling&);
TPLS_INST_HH(double)
TPLS_INST_HH(float)

}  // namespace tpls

