// Crank-Nicolson viscous Helmholtz sweeps: jacobi_iteration_allflux.f90:21-282,
// sor_iteration_allflux.f90:19-377 and the outer-loop boundary conditions of
// main_ns_hybrid.F90:758-880.
//
// Component ranges (comp 0/1/2 = u/v/w): u on i in 1..maxl-2 (the outlet face
// maxl-1 is a boundary condition), k in 0..maxn-2 with the wall ghost
// -2 x0 + x1/3; v on i in 1..maxl-1, same k; w on i in 1..maxl-1,
// k in 1..maxn-2 between the walls w = 0.
#include <type_traits>

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
    auto avg = [&](int f) { return (r.a + r.b + r.q[f][0] + r.q[f][1]) / C(4); };
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
__device__ __forceinline__ bool cell_of(const Grid& g, int nbx, int nby, int& i, int& j, int& k)
{
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
    i = bx * BX + threadIdx.x;
    j = by * RY + threadIdx.y + 1;
    k = bz + kmin_of<COMP>();
    return j <= g.P;
}

template <int COMP, int RY, typename T>
__global__ void __launch_bounds__(BX* RY) k_jacobi(Grid g, int nbx, int nby, T ax2, T ay2, T az2, FieldView<T> xo,
                                                   FieldView<T> rhs, FieldView<T> mu, FieldView<T> x, int xo_zero)
{
    using C = T;
    int i, j, k;
    if (!cell_of<COMP, RY>(g, nbx, nby, i, j, k)) return;
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

// Even iteration_time: 3-D colouring (i+j+k) = 1+flg (mod 2); every cell of a
// colour reads only the other colour, so the sweep is fully parallel.  One
// thread per cell of the colour.
template <int COMP, int RY, typename T>
__global__ void __launch_bounds__(BX* RY) k_sor3d(Grid g, int nbx, int nby, int flg, T ax2, T ay2, T az2,
                                                  FieldView<T> x, FieldView<T> rhs, FieldView<T> mu)
{
    using C = T;
    int t, j, k;
    if (!cell_of<COMP, RY>(g, nbx, nby, t, j, k)) return;
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
__global__ void __launch_bounds__(BX* RY) k_sor2d(Grid g, int flg, T ax2, T ay2, T az2, FieldView<T> x,
                                                  FieldView<T> rhs, FieldView<T> mu)
{
    using C = T;
    constexpr int SOR_D = D;
    // 1-D grid of nbx x nby tiles, XCD-aware: each XCD marches a contiguous
    // band of rows, so the neighbour rows and viscosity rows a tile shares
    // with the next stay in one L2
    const int nbx = (g.maxl / 2 + BX - 1) / BX;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
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

// Outer-loop boundary conditions.
// u: main_ns_hybrid.F90:764-776 (inlet copy, outlet pressure extrapolation
//    with the unshifted j of pres, fictitious face maxl = 0);
// v: main_ns_hybrid.F90:826-836 / jacobi_iteration_allflux.f90:178-189;
// w: main_ns_hybrid.F90:864-879 / jacobi_iteration_allflux.f90:257-279.
template <typename T>
__global__ void k_bc_uv(Grid g, int comp, T dtdx, FieldView<T> x, FieldView<T> pres)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
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

// Two independent segments: threads [0, P*(maxn-2)) copy the inlet/outlet
// columns for k in 1..maxn-2; threads [P*(maxn-2), P*(maxn-2) + (maxl+1)*P)
// zero both walls of column (i, j).
template <typename T>
__global__ void k_bc_w(Grid g, FieldView<T> w)
{
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
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

template <typename T>
hipError_t launch_jacobi(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xold, FieldView<T> rhs,
                         FieldView<T> mu, FieldView<T> x, bool xold_is_zero, const Tiling& tl)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0);
    const int z = xold_is_zero ? 1 : 0;
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        return with_rows<CP, T>(tl.rows, [&](auto ry) {
            constexpr int RY = decltype(ry)::value;
            const int nbx = (g.maxl - 1 + BX - 1) / BX, nby = (g.P + RY - 1) / RY;
            const dim3 grd(nbx * nby * nz_of(CP, g));
            hipLaunchKernelGGL((k_jacobi<CP, RY, T>), grd, dim3(BX, RY), 0, s, g, nbx, nby, a2, b2, c2, xold, rhs,
                               mu, x, z);
            return hipGetLastError();
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
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const T a2 = T(ax / 2.0), b2 = T(ay / 2.0), c2 = T(az / 2.0);
    const int half = (g.maxl - 1 + 1) / 2;
    const int nbx = (half + BX - 1) / BX;
    auto go = [&](auto comp_c) {
        constexpr int CP = decltype(comp_c)::value;
        if (iteration_time % 2 == 0) {
            return with_rows<CP, T>(tl.rows, [&](auto ry) {
                constexpr int RY = decltype(ry)::value;
                const int nby = (g.P + RY - 1) / RY;
                const dim3 grd(nbx * nby * nz_of(CP, g));
                hipLaunchKernelGGL((k_sor3d<CP, RY, T>), grd, dim3(BX, RY), 0, s, g, nbx, nby, flg, a2, b2, c2, x,
                                   rhs, mu);
                return hipGetLastError();
            });
        }
        return with_rows<CP, T>(tl.rows_odd, [&](auto ry) {
            constexpr int RY = decltype(ry)::value;
            const dim3 grd(nbx * ((g.P + RY - 1) / RY));
            if (tl.depth_odd == 8)
                hipLaunchKernelGGL((k_sor2d<CP, RY, 8, T>), grd, dim3(BX, RY), 0, s, g, flg, a2, b2, c2, x, rhs, mu);
            else
                hipLaunchKernelGGL((k_sor2d<CP, RY, 4, T>), grd, dim3(BX, RY), 0, s, g, flg, a2, b2, c2, x, rhs, mu);
            return hipGetLastError();
        });
    };
    if (comp == 0) return go(std::integral_constant<int, 0>{});
    if (comp == 1) return go(std::integral_constant<int, 1>{});
    return go(std::integral_constant<int, 2>{});
}

template <typename T>
hipError_t launch_helmholtz_bc(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> x,
                               FieldView<T> pres)
{
    if (comp == 2) {
        const long n = long(g.P) * (g.maxn - 2) + long(g.maxl + 1) * g.P;
        hipLaunchKernelGGL(k_bc_w<T>, dim3((n + 255) / 256), dim3(256), 0, s, g, x);
    } else {
        hipLaunchKernelGGL(k_bc_uv<T>, dim3((g.maxn - 1 + 63) / 64, g.P), dim3(64), 0, s, g, comp, T(ph.dt / ph.dx), x,
                           pres);
    }
    return hipGetLastError();
}

#define TPLS_INST_HH(T)                                                                                              \
    template hipError_t launch_jacobi<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,     \
                                         FieldView<T>, FieldView<T>, bool, const Tiling&);                           \
    template hipError_t launch_sor<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,        \
                                      FieldView<T>, int, int, const Tiling&);                                        \
    template hipError_t launch_helmholtz_bc<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>);
TPLS_INST_HH(double)
TPLS_INST_HH(float)

}  // namespace tpls
