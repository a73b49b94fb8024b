// Diffuse-interface method: the Cahn-Hilliard update of
// cahn_hilliard_solver.f90:41-384 (dim), as main_ns_hybrid.F90:606-611
// calls it, on gfx950.
//
// The reference forms phi = phi2 + 0.5 and phiOld = phi_old + 0.5 as
// temporaries; the kernels read phi2 and add the shift 0.5 per read (shift 0
// when the caller passes phi itself).  phi_old equals phi2 whenever the time
// loop calls dim (main_ns_hybrid.F90:444, 2507-2508), so the case passes
// phi2 for both.  Its work arrays bulkE0 and
// d2phi get their boundary planes by copies (z planes, inlet and outlet
// columns) and their y halos by the periodic exchange before any stencil
// reads them; every such plane is a copy of an interior value, so the
// stencils here read the interior point the copy came from (clampx / clampz
// and wrapj) and the work arrays are stored over the interior only.
//
// The relaxation iterate phiNew is updated point by point from d2phi and
// grad^4 of phiNew, which depend only on the previous iterate, so each sweep
// is two launches: d2phi of the iterate, then grad^4, the update, the clip to
// [0, 1] and the iterate's boundary planes in one pass (in place: every
// thread reads and writes only its own point and the planes it owns).
//
// Scalar factors (the diagonal, the relaxation coefficient, the inlet tanh
// profile) are computed on the host in FP64 with the C library, in the
// reference's operation order.  Every expression rounds as written (no FMA
// contraction), which makes the FP64 path bit-identical to a sequential
// evaluation of the reference loops.
#include <algorithm>
#include <cmath>

#include "tiles.hpp"
#include "tpls_kernels.hpp"

#pragma clang fp contract(off)

namespace tpls {

namespace {

// x + h, or x itself when the shift is zero
template <typename T>
__device__ __forceinline__ T sh(T x, T h)
{
    return h != T(0) ? x + h : x;
}

__device__ __forceinline__ int clampx(int i, int L) { return i < 1 ? 1 : (i > L - 1 ? L - 1 : i); }
__device__ __forceinline__ int clampz(int k, int n) { return k < 1 ? 1 : (k > n - 1 ? n - 1 : k); }

// 7-point Laplacian, cahn_hilliard_solver.f90:108-110 / 256-258 / 305-307,
// of a stored field plus h whose neighbours are read as stored (phi2 or the
// iterate, boundary planes included)
template <typename T>
__device__ __forceinline__ T lap_stored(const FieldView<T>& a, T h, int i, int j, int k, int P, T dx, T dy, T dz)
{
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    auto v = [&](int ii, int jj, int kk) { return sh(a(ii, jj, kk), h); };
    const T c = v(i, j, k);
    return (v(i + 1, j, k) - T(2) * c + v(i - 1, j, k)) / dx / dx +
           (v(i, jp, k) - T(2) * c + v(i, jm, k)) / dy / dy +
           (v(i, j, k + 1) - T(2) * c + v(i, j, k - 1)) / dz / dz;
}

// the same of a work array stored over the interior only (boundary planes
// read as the interior value they copy)
template <typename T>
__device__ __forceinline__ T lap_work(const FieldView<T>& a, int i, int j, int k, const Grid& g, T dx, T dy, T dz)
{
    const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
    const int ip = clampx(i + 1, g.maxl), im = clampx(i - 1, g.maxl);
    const int kp = clampz(k + 1, g.maxn), km = clampz(k - 1, g.maxn);
    const T c = a(i, j, k);
    return (a(ip, j, k) - T(2) * c + a(im, j, k)) / dx / dx + (a(i, jp, k) - T(2) * c + a(i, jm, k)) / dy / dy +
           (a(i, j, kp) - T(2) * c + a(i, j, km)) / dz / dz;
}

// bulk chemical potential of phi (cahn_hilliard_solver.f90:104)
template <typename T>
__device__ __forceinline__ T bulk(T p)
{
    return (T(1) - T(3) * p + T(2) * p * p) * p / T(2);
}

// face mobility sqrt(c^2 (1-|c|)^2) (cahn_hilliard_solver.f90:176-191)
template <typename T>
__device__ __forceinline__ T mob(T c1)
{
    return sqrt(c1 * c1 * (T(1) - fabs(c1)) * (T(1) - fabs(c1)));
}

}  // namespace

// phiNew's first iterate: the entry phi3, which is the last step's phi3 (=
// phi2) or zero before the first step (main_ns_hybrid.F90:443); all of
// i = 0..maxl, k = 0..maxn.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_dim_init(Grid g, int zero, FieldView<T> phi2, FieldView<T> pn)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i > g.maxl || j > g.P) return;
    pn(i, j, k) = zero ? T(0) : phi2(i, j, k);
}

// d2phi of phi2 + h or of the iterate (h = 0), over the interior
// (cahn_hilliard_solver.f90:101-115, 251-261)
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_dim_lap(Grid g, T dx, T dy, T dz, T h, FieldView<T> a, FieldView<T> d2)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    d2(i, j, k) = lap_stored<T>(a, h, i, j, k, g.P, dx, dy, dz);
}

// Explicit part, cahn_hilliard_solver.f90:165-237: the mobility-weighted
// divergences of grad(bulkE0) and grad(d2phi), the source term source0 and
// the new ConvPrevPhi.
template <typename T>
__global__ void __launch_bounds__(BX* BY)
    k_dim_source(Grid g, T dx, T dy, T dz, T epn, T epn2, T Pe, T c_tao, T c_epn, T h, FieldView<T> phi2,
                 FieldView<T> d2, FieldView<T> conv2, FieldView<T> cprev, FieldView<T> src)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const int L = g.maxl, n = g.maxn, P = g.P;
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    // phi = phi2 + h at the point and its six neighbours, boundary planes as stored
    const T p = sh(phi2(i, j, k), h);
    const T pxm = sh(phi2(i - 1, j, k), h), pxp = sh(phi2(i + 1, j, k), h);
    const T pym = sh(phi2(i, jm, k), h), pyp = sh(phi2(i, jp, k), h);
    const T pzm = sh(phi2(i, j, k - 1), h), pzp = sh(phi2(i, j, k + 1), h);
    const T a1 = mob((pxm + p) / T(2));
    const T a2 = mob((pxp + p) / T(2));
    const T a3 = mob((pym + p) / T(2));
    const T a4 = mob((pyp + p) / T(2));
    const T a5 = mob(k == 1 ? p : (pzm + p) / T(2));
    const T a6 = mob((pzp + p) / T(2));
    // bulkE0 at the neighbours: its boundary planes copy the adjacent interior
    // value (cahn_hilliard_solver.f90:121-163), i.e. bulk(phi) there
    const T b = bulk(p);
    const T bxm = i == 1 ? b : bulk(pxm), bxp = i == L - 1 ? b : bulk(pxp);
    const T bym = bulk(pym), byp = bulk(pyp);
    const T bzm = k == 1 ? b : bulk(pzm), bzp = k == n - 1 ? b : bulk(pzp);
    const T temp1 = (a2 * (bxp - b) - a1 * (b - bxm)) / dx / dx + (a4 * (byp - b) - a3 * (b - bym)) / dy / dy +
                    (a6 * (bzp - b) - a5 * (b - bzm)) / dz / dz;
    const T d = d2(i, j, k);
    const T dxp = d2(clampx(i + 1, L), j, k), dxm = d2(clampx(i - 1, L), j, k);
    const T dyp = d2(i, jp, k), dym = d2(i, jm, k);
    const T dzp = d2(i, j, clampz(k + 1, n)), dzm = d2(i, j, clampz(k - 1, n));
    const T temp2 = (dxp - T(2) * d + dxm) / dx / dx + (dyp - T(2) * d + dym) / dy / dy +
                    (dzp - T(2) * d + dzm) / dz / dz;
    const T temp3 = (a2 * (dxp - d) - a1 * (d - dxm)) / dx / dx + (a4 * (dyp - d) - a3 * (d - dym)) / dy / dy +
                    (a6 * (dzp - d) - a5 * (d - dzm)) / dz / dz;
    const T temp4 = (temp1 - epn2 * temp3) / epn - c_tao * d + c_epn * temp2;
    const T temp5 = temp4 / Pe - conv2(i, j, k);
    src(i, j, k) = temp5 * T(2) - cprev(i, j, k);
    cprev(i, j, k) = temp5;
}

// One relaxation sweep's update, cahn_hilliard_solver.f90:281-354: grad^4 of
// the iterate from d2phi, the relaxed point update clipped to [0, 1], and the
// iterate's boundary planes (z copies, inlet profile, outlet copy).  In place.
// resmax (non-null on the last sweep when diagnostics are on): the largest
// |change| of the sweep (the reference's res, :325-344), kept as the bit
// pattern of a non-negative double, whose integer order is its value order.
template <typename T>
__global__ void __launch_bounds__(BX* BY)
    k_dim_update(Grid g, T dx, T dy, T dz, T dt, T c_rel, T tao, T epn2, T alfa1, T diagonal, T h,
                 FieldView<T> phi2, FieldView<T> phio, FieldView<T> src, FieldView<T> d2,
                 const T* __restrict__ inlet, FieldView<T> pn, unsigned long long* resmax)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double change = 0.0;
    if (i <= g.maxl - 1 && j <= g.P) {
        const T B = lap_work(d2, i, j, k, g, dx, dy, dz);
        const T p = sh(phi2(i, j, k), h), po = sh(phio(i, j, k), h);
        const T old = pn(i, j, k);
        const T some =
            T(2) * p - T(0.5) * po - T(1.5) * old + dt * src(i, j, k) + c_rel * (tao * d2(i, j, k) - epn2 * B);
        T nw = old + alfa1 * some / diagonal;
        if (nw < T(0)) nw = T(0);
        if (nw > T(1)) nw = T(1);
        pn(i, j, k) = nw;
        if (k == 1) pn(i, j, 0) = nw;
        if (k == g.maxn - 1) pn(i, j, g.maxn) = nw;
        if (i == g.maxl - 1) pn(g.maxl, j, k) = nw;
        if (i == 1) pn(0, j, k) = inlet[k];
        change = double(fabs(nw - old));
    }
    if (resmax) {
        // every lane of the 64-wide row takes part (out-of-range lanes with 0)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) change = fmax(change, __shfl_xor(change, off, 64));
        if (threadIdx.x == 0) atomicMax(resmax, (unsigned long long)__double_as_longlong(change));
    }
}

// One relaxation sweep in one pass (the two launches above fused), out of
// place: pi is the previous iterate, po receives the new one.  A block owns a
// tile of DIM_TI x TJ columns and marches up a chunk of ks levels.  Per level
// c it stages plane c+1 of the iterate with a two-cell ring (rows wrapped,
// columns clamped to the stored boundary planes 0..maxl; loaded one level
// ahead into registers), forms d2phi of
// plane c on the tile with a one-cell ring (ring columns outside the
// interior take the clamped interior value, as the copies of
// cahn_hilliard_solver.f90:264-292 make them), then updates plane c-1.  Every
// output reads the same values k_dim_lap + k_dim_update read, so the results
// are identical.  The boundary planes of po that no sweep writes (the inlet
// and outlet columns at k = 0 and maxn) must already hold the first iterate:
// k_dim_init fills both buffers.
constexpr int DIM_TI = 64;

template <typename T, int TJ>
__global__ void __launch_bounds__(DIM_TI* TJ)
    k_dim_sweep(Grid g, int nbx, int nby, int ks, T dx, T dy, T dz, T dt, T c_rel, T tao, T epn2, T alfa1,
                T diagonal, T h, FieldView<T> phi2, FieldView<T> phio, FieldView<T> src,
                const T* __restrict__ inlet, FieldView<T> pi, FieldView<T> po, unsigned long long* resmax)
{
    constexpr int PW = DIM_TI + 4, PH = TJ + 4;  // iterate window: two-cell ring
    constexpr int DW = DIM_TI + 2, DH = TJ + 2;  // d2phi window: one-cell ring
    constexpr int NT = DIM_TI * TJ;
    __shared__ T ps[4][PH][PW];
    __shared__ T ds[4][DH][DW];
    const int t = blockIdx.x;
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int L = g.maxl, n = g.maxn, P = g.P;
    const int i0 = 1 + bx * DIM_TI, j0 = 1 + by * TJ;
    const int kb = 1 + bz * ks, ke = min(kb + ks - 1, n - 1);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * DIM_TI + tx;
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i <= L - 1 && j <= P;
    auto wrapP = [&](int jj) { return (jj - 1 + 2 * P) % P + 1; };
    // plane c of the iterate: loaded into registers (this thread's share of
    // the window), then stored into its ring slot
    constexpr int NQ = (PW * PH + NT - 1) / NT;
    auto fetch = [&](int c, T (&r)[NQ]) {
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const int q = min(tid + m * NT, PW * PH - 1);
            const int li = q % PW, lj = q / PW;
            r[m] = pi(min(max(i0 - 2 + li, 0), L), wrapP(j0 - 2 + lj), c);
        }
    };
    auto put = [&](int c, const T (&r)[NQ]) {
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const int q = tid + m * NT;
            if (q < PW * PH) ps[c & 3][q / PW][q % PW] = r[m];
        }
    };
    auto stage = [&](int c) {
        T r[NQ];
        fetch(c, r);
        put(c, r);
    };
    // d2phi of plane c (1 <= c <= n-1) over the one-cell ring window
    auto d2plane = [&](int c) {
        for (int q = tid; q < DW * DH; q += NT) {
            const int a = q % DW, b = q / DW;
            const int ca = min(max(i0 - 1 + a, 1), L - 1);
            const int li = ca - (i0 - 2), lj = b + 1;
            const T cc = ps[c & 3][lj][li];
            ds[c & 3][b][a] = (ps[c & 3][lj][li + 1] - T(2) * cc + ps[c & 3][lj][li - 1]) / dx / dx +
                              (ps[c & 3][lj + 1][li] - T(2) * cc + ps[c & 3][lj - 1][li]) / dy / dy +
                              (ps[(c + 1) & 3][lj][li] - T(2) * cc + ps[(c - 1) & 3][lj][li]) / dz / dz;
        }
    };
    double change = 0.0;
    // update of plane k from d2phi planes clampz(k-1), k, clampz(k+1)
    auto update = [&](int k) {
        if (!mine) return;
        const int a = tx + 1, b = ty + 1;
        const int km = k > 1 ? k - 1 : 1, kp = k < n - 1 ? k + 1 : n - 1;
        const T d = ds[k & 3][b][a];
        const T B = (ds[k & 3][b][a + 1] - T(2) * d + ds[k & 3][b][a - 1]) / dx / dx +
                    (ds[k & 3][b + 1][a] - T(2) * d + ds[k & 3][b - 1][a]) / dy / dy +
                    (ds[kp & 3][b][a] - T(2) * d + ds[km & 3][b][a]) / dz / dz;
        const T p = sh(phi2(i, j, k), h), pold = sh(phio(i, j, k), h);
        const T old = ps[k & 3][ty + 2][tx + 2];
        const T some = T(2) * p - T(0.5) * pold - T(1.5) * old + dt * src(i, j, k) + c_rel * (tao * d - epn2 * B);
        T nw = old + alfa1 * some / diagonal;
        if (nw < T(0)) nw = T(0);
        if (nw > T(1)) nw = T(1);
        po(i, j, k) = nw;
        if (k == 1) po(i, j, 0) = nw;
        if (k == n - 1) po(i, j, n) = nw;
        if (i == L - 1) po(L, j, k) = nw;
        if (i == 1) po(0, j, k) = inlet[k];
        change = fmax(change, double(fabs(nw - old)));
    };
    const int clo = max(kb - 1, 1), chi = min(ke + 1, n - 1);
    stage(clo - 1);
    stage(clo);
    // plane c+2 is in flight while plane c's d2phi and plane c-1's update run
    T nxt[NQ];
    fetch(clo + 1, nxt);
    for (int c = clo; c <= chi; ++c) {
        put(c + 1, nxt);
        if (c + 2 <= n) fetch(c + 2, nxt);
        __syncthreads();
        d2plane(c);
        __syncthreads();
        if (c - 1 >= kb) update(c - 1);
    }
    if (ke == n - 1) update(ke);
    if (resmax) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) change = fmax(change, __shfl_xor(change, off, 64));
        if (tx == 0) atomicMax(resmax, (unsigned long long)__double_as_longlong(change));
    }
}

// phi3 = phiNew - 0.5 over i = 0..maxl, k = 0..maxn (main_ns_hybrid.F90:611)
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_dim_final(Grid g, FieldView<T> pn, FieldView<T> phi3)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i > g.maxl || j > g.P) return;
    phi3(i, j, k) = pn(i, j, k) - T(0.5);
}

void dim_coefficients(double dx, double dy, double dz, double dt, double epn, double Pe, DimCoef* c)
{
    // cahn_hilliard_solver.f90:73-98, 240-245 (dx**2.d0 and dx**4.d0 are
    // real-exponent powers)
    const double nmda = 0.25, tao = 0.25;
    const double epn2 = epn * epn;
    c->epn = epn;
    c->epn2 = epn2;
    c->Pe = Pe;
    c->dt = dt;
    c->c_tao = nmda * tao / 2.0;
    c->c_epn = epn2 * nmda / 2.0;
    c->c_rel = dt * nmda / 2.0 / Pe;
    c->tao = tao;
    c->alfa1 = 1.1;
    const double a1 = 1.0 / std::pow(dx, 2.0) + 1.0 / std::pow(dy, 2.0) + 1.0 / std::pow(dz, 2.0);
    const double a2 = 2.0 / std::pow(dx, 4.0) + 2.0 / std::pow(dy, 4.0) + 2.0 / std::pow(dz, 4.0);
    c->diagonal = 1.5 + nmda / 2.0 / Pe * dt * (2.0 * tao * a1 + epn2 * (4.0 * a1 * a1 + a2));
}

void dim_inlet(int maxn, double dz, double height, double epn, double* prof)
{
    // cahn_hilliard_solver.f90:346-353; k = 0 and maxn are not used
    prof[0] = prof[maxn] = 0.0;
    for (int k = 1; k <= maxn - 1; ++k) {
        const double z_val = k * dz - 0.5 * dz;
        prof[k] = 0.5 + 0.5 * std::tanh((z_val - height) / (2.0 * std::pow(2.0, 0.5) * epn));
    }
}

static inline dim3 grid_int(const Grid& g) { return dim3((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn - 1); }
static inline dim3 grid_all(const Grid& g) { return dim3((g.maxl + 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn + 1); }

// levels per block of the fused sweep
constexpr int DIM_KS = 32;
constexpr int DIM_TJ = 8;

template <typename T>
hipError_t launch_dim_core(hipStream_t s, const Grid& g, const Phys& ph, const DimCoef& c, int its, double shift,
                           FieldView<T> phi, FieldView<T> phio, FieldView<T> conv2, FieldView<T> cprev,
                           FieldView<T> pn, FieldView<T> pn2, FieldView<T> d2, FieldView<T> src, const T* inlet,
                           double* res, int fused, int ks_fixed)
{
    if (res) {
        hipError_t e = launch_fill<double>(s, res, 1, 0.0);
        if (e != hipSuccess) return e;
    }
    const dim3 blk(BX, BY);
    const T dx = T(ph.dx), dy = T(ph.dy), dz = T(ph.dz), h = T(shift);
    hipLaunchKernelGGL(k_dim_lap<T>, grid_int(g), blk, 0, s, g, dx, dy, dz, h, phi, d2);
    hipLaunchKernelGGL(k_dim_source<T>, grid_int(g), blk, 0, s, g, dx, dy, dz, T(c.epn), T(c.epn2), T(c.Pe),
                       T(c.c_tao), T(c.c_epn), h, phi, d2, conv2, cprev, src);
    const int nbx = (g.maxl - 1 + DIM_TI - 1) / DIM_TI, nby = (g.P + DIM_TJ - 1) / DIM_TJ;
    static const int slots = resident_blocks(k_dim_sweep<T, DIM_TJ>, DIM_TI * DIM_TJ);
    const int ks = ks_fixed > 0 ? std::min(ks_fixed, g.maxn - 1) : march_chunk(nbx * nby, g.maxn - 1, slots, DIM_KS);
    const int nbz = (g.maxn - 1 + ks - 1) / ks;
    for (int it = 1; it <= its; ++it) {
        unsigned long long* rm = res && it == its ? reinterpret_cast<unsigned long long*>(res) : nullptr;
        if (fused) {
            // odd sweeps read pn and write pn2, even sweeps the reverse
            FieldView<T> a = (it & 1) ? pn : pn2, b = (it & 1) ? pn2 : pn;
            hipLaunchKernelGGL((k_dim_sweep<T, DIM_TJ>), dim3(nbx * nby * nbz), dim3(DIM_TI, DIM_TJ), 0, s, g, nbx,
                               nby, ks, dx, dy, dz, T(c.dt), T(c.c_rel), T(c.tao), T(c.epn2), T(c.alfa1),
                               T(c.diagonal), h, phi, phio, src, inlet, a, b, rm);
        } else {
            hipLaunchKernelGGL(k_dim_lap<T>, grid_int(g), blk, 0, s, g, dx, dy, dz, T(0), pn, d2);
            hipLaunchKernelGGL(k_dim_update<T>, grid_int(g), blk, 0, s, g, dx, dy, dz, T(c.dt), T(c.c_rel),
                               T(c.tao), T(c.epn2), T(c.alfa1), T(c.diagonal), h, phi, phio, src, d2, inlet, pn,
                               rm);
        }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dim(hipStream_t s, const Grid& g, const Phys& ph, const DimCoef& c, int its, int zero,
                      FieldView<T> phi2, FieldView<T> conv2, FieldView<T> cprev, FieldView<T> phi3, FieldView<T> pn,
                      FieldView<T> pn2, FieldView<T> d2, FieldView<T> src, const T* inlet, double* res, int fused,
                      int ks)
{
    const dim3 blk(BX, BY);
    hipLaunchKernelGGL(k_dim_init<T>, grid_all(g), blk, 0, s, g, zero, phi2, pn);
    if (fused) hipLaunchKernelGGL(k_dim_init<T>, grid_all(g), blk, 0, s, g, zero, phi2, pn2);
    hipError_t e =
        launch_dim_core<T>(s, g, ph, c, its, 0.5, phi2, phi2, conv2, cprev, pn, pn2, d2, src, inlet, res, fused, ks);
    if (e != hipSuccess) return e;
    // the last iterate: in pn2 after an odd number of fused sweeps
    hipLaunchKernelGGL(k_dim_final<T>, grid_all(g), blk, 0, s, g, fused && (its & 1) ? pn2 : pn, phi3);
    return hipGetLastError();
}

#define TPLS_INST_DIM(T)                                                                                         \
    template hipError_t launch_dim<T>(hipStream_t, const Grid&, const Phys&, const DimCoef&, int, int,          \
                                      FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>,     \
                                      FieldView<T>, FieldView<T>, FieldView<T>, const T*, double*, int, int);   \
    template hipError_t launch_dim_core<T>(hipStream_t, const Grid&, const Phys&, const DimCoef&, int, double,  \
                                           FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>,              \
                                           FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, const T*,    \
                                           double*, int, int);
TPLS_INST_DIM(double)
TPLS_INST_DIM(float)

}  // namespace tpls
