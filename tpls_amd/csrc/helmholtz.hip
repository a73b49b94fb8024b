// Crank-Nicolson viscous Helmholtz sweeps: jacobi_iteration_allflux.f90:21-282,
// sor_iteration_allflux.f90:19-377 and the outer-loop boundary conditions of
// main_ns_hybrid.F90:758-880.
//
// Component ranges (comp 0/1/2 = u/v/w): u on i in 1..maxl-2 (the outlet face
// maxl-1 is a boundary condition), k in 0..maxn-2 with the wall ghost
// -2 x0 + x1/3; v on i in 1..maxl-1, same k; w on i in 1..maxl-1,
// k in 1..maxn-2 between the walls w = 0.
#include "tpls_kernels.hpp"

namespace tpls {

template <typename C, typename T>
__device__ __forceinline__ MuFaces<C> mu_faces(int comp, const FieldView<T>& mu, int i, int j, int k, int P)
{
    if (comp == 0) return mu_faces_u<C>(mu, i, j, k, P);
    if (comp == 1) return mu_faces_v<C>(mu, i, j, k, P);
    return mu_faces_w<C>(mu, i, j, k, P);
}

__device__ __forceinline__ int imax_of(int comp, int maxl) { return comp == 0 ? maxl - 2 : maxl - 1; }

// Wall-ghost neighbours in z for u and v (jacobi_iteration_allflux.f90:63-73).
template <typename C, typename T>
__device__ __forceinline__ void z_nbrs(int comp, const FieldView<T>& x, int i, int j, int k, int maxn, C& zp, C& zm)
{
    if (comp == 2) {
        zp = x(i, j, k + 1);
        zm = x(i, j, k - 1);
        return;
    }
    zm = (k == 0) ? C(-2) * C(x(i, j, 0)) + C(x(i, j, 1)) / C(3) : C(x(i, j, k - 1));
    zp = (k == maxn - 2) ? C(-2) * C(x(i, j, maxn - 2)) + C(x(i, j, maxn - 3)) / C(3) : C(x(i, j, k + 1));
}

template <typename T>
__global__ void __launch_bounds__(BX* BY) k_jacobi(Grid g, int comp, T ax2, T ay2, T az2, FieldView<T> xo,
                                                   FieldView<T> rhs, FieldView<T> mu, FieldView<T> x, int xo_zero)
{
    using C = T;
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + (comp == 2 ? 1 : 0);
    if (i > imax_of(comp, g.maxl) || j > g.P) return;
    MuFaces<C> m = mu_faces<C>(comp, mu, i, j, k, g.P);
    C xp = 0, xm = 0, yp = 0, ym = 0, zp = 0, zm = 0;
    if (!xo_zero) {
        xp = xo(i + 1, j, k);
        xm = xo(i - 1, j, k);
        yp = xo(i, wrapj(j + 1, g.P), k);
        ym = xo(i, wrapj(j - 1, g.P), k);
        z_nbrs<C>(comp, xo, i, j, k, g.maxn, zp, zm);
    }
    x(i, j, k) = helmholtz<C>(m, ax2, ay2, az2, xp, xm, yp, ym, zp, zm, C(rhs(i, j, k)));
}

// Even iteration_time: 3-D colouring (i+j+k) = 1+flg (mod 2); every cell of a
// colour reads only the other colour, so the sweep is fully parallel.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_sor3d(Grid g, int comp, int flg, T ax2, T ay2, T az2, FieldView<T> x,
                                                  FieldView<T> rhs, FieldView<T> mu)
{
    using C = T;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + (comp == 2 ? 1 : 0);
    const int i = ((k + j + flg) & 1) + 1 + 2 * (blockIdx.x * BX + threadIdx.x);
    if (i > imax_of(comp, g.maxl) || j > g.P) return;
    MuFaces<C> m = mu_faces<C>(comp, mu, i, j, k, g.P);
    C zp, zm;
    z_nbrs<C>(comp, x, i, j, k, g.maxn, zp, zm);
    C c = x(i, j, k);
    C res = helmholtz<C>(m, ax2, ay2, az2, C(x(i + 1, j, k)), C(x(i - 1, j, k)), C(x(i, wrapj(j + 1, g.P), k)),
                         C(x(i, wrapj(j - 1, g.P), k)), zp, zm, C(rhs(i, j, k)));
    x(i, j, k) = c + C(1.2) * (res - c);
}

// Odd iteration_time: 2-D colouring (i+j) = 1+flg (mod 2) with k ascending,
// i.e. Gauss-Seidel up each column (sor_iteration_allflux.f90:100-146 run on
// one thread).  One thread per column of the colour.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_sor2d(Grid g, int comp, int flg, T ax2, T ay2, T az2, FieldView<T> x,
                                                  FieldView<T> rhs, FieldView<T> mu)
{
    using C = T;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int i = ((j + flg) & 1) + 1 + 2 * (blockIdx.x * BX + threadIdx.x);
    if (i > imax_of(comp, g.maxl) || j > g.P) return;
    const int P = g.P, maxn = g.maxn;
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    const int k0 = comp == 2 ? 1 : 0, k1 = maxn - 2;
    C prev = 0;        // x(i,j,k-1) after its update
    C cur = x(i, j, k0);
    for (int k = k0; k <= k1; ++k) {
        MuFaces<C> m = mu_faces<C>(comp, mu, i, j, k, P);
        C nxt = (k < k1 || comp == 2) ? C(x(i, j, k + 1)) : C(0);
        C zp, zm;
        if (comp == 2) {
            zp = nxt;
            zm = (k == k0) ? C(x(i, j, k - 1)) : prev;
        } else {
            zm = (k == 0) ? C(-2) * cur + nxt / C(3) : prev;
            zp = (k == maxn - 2) ? C(-2) * cur + prev / C(3) : nxt;
        }
        C res = helmholtz<C>(m, ax2, ay2, az2, C(x(i + 1, j, k)), C(x(i - 1, j, k)), C(x(i, jp, k)), C(x(i, jm, k)),
                             zp, zm, C(rhs(i, j, k)));
        C upd = cur + C(1.2) * (res - cur);
        x(i, j, k) = upd;
        prev = upd;
        cur = nxt;
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

template <typename T>
hipError_t launch_jacobi(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> xold, FieldView<T> rhs,
                         FieldView<T> mu, FieldView<T> x, bool xold_is_zero)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    dim3 grd((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, nz_of(comp, g));
    hipLaunchKernelGGL(k_jacobi<T>, grd, dim3(BX, BY), 0, s, g, comp, T(ax / 2.0), T(ay / 2.0), T(az / 2.0), xold,
                       rhs, mu, x, xold_is_zero ? 1 : 0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sor(hipStream_t s, const Grid& g, const Phys& ph, int comp, FieldView<T> x, FieldView<T> rhs,
                      FieldView<T> mu, int flg, int iteration_time)
{
    const double ax = ph.dt / (ph.dx * ph.dx), ay = ph.dt / (ph.dy * ph.dy), az = ph.dt / (ph.dz * ph.dz);
    const int half = (g.maxl - 1 + 1) / 2;
    if (iteration_time % 2 == 0) {
        dim3 grd((half + BX - 1) / BX, (g.P + BY - 1) / BY, nz_of(comp, g));
        hipLaunchKernelGGL(k_sor3d<T>, grd, dim3(BX, BY), 0, s, g, comp, flg, T(ax / 2.0), T(ay / 2.0), T(az / 2.0), x,
                           rhs, mu);
    } else {
        dim3 grd((half + BX - 1) / BX, (g.P + BY - 1) / BY, 1);
        hipLaunchKernelGGL(k_sor2d<T>, grd, dim3(BX, BY), 0, s, g, comp, flg, T(ax / 2.0), T(ay / 2.0), T(az / 2.0), x,
                           rhs, mu);
    }
    return hipGetLastError();
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
                                         FieldView<T>, FieldView<T>, bool);                                          \
    template hipError_t launch_sor<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>,        \
                                      FieldView<T>, int, int);                                                       \
    template hipError_t launch_helmholtz_bc<T>(hipStream_t, const Grid&, const Phys&, int, FieldView<T>, FieldView<T>);
TPLS_INST_HH(double)
TPLS_INST_HH(float)

}  // namespace tpls
