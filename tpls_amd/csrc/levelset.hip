// Level-set update, Russo-Smereka reinitialisation and CSF force:
// two_phase_levelset.f90:20-118 and levelset.f90:17-414.
#include "tiles.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

// phi3 = phi2 - dt * AB3(conv), two_phase_levelset.f90:56-70, p interior.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_tpls_update(Grid g, T dt, FieldView<T> phi2, FieldView<T> c0,
                                                        FieldView<T> c1, FieldView<T> c2, FieldView<T> phi3)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    T rhs = T(23.0 / 12.0) * c2(i, j, k) - T(4.0 / 3.0) * c1(i, j, k) + T(5.0 / 12.0) * c0(i, j, k);
    phi3(i, j, k) = phi2(i, j, k) - dt * rhs;
}

// Boundary planes of a level-set field.  Pass 1, one thread per (j, k),
// k in 1..n-1: inlet phi = z - height and outlet copy (two_phase_levelset.f90:
// 72-82, levelset.f90:170-191).  Pass 2, one thread per (i, j), i in 0..maxl:
// linear extrapolation to k = 0 and k = n (two_phase_levelset.f90:94-95,
// levelset.f90:193-216).
template <typename T>
__global__ void k_phi_inout(Grid g, T dz, T height, FieldView<T> phi)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    if (k > g.maxn - 1) return;
    T z = T(k) * dz - T(0.5) * dz;
    phi(0, j, k) = z - height;
    phi(g.maxl, j, k) = phi(g.maxl - 1, j, k);
}

template <typename T>
__global__ void k_phi_zext(Grid g, FieldView<T> phi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    const int n = g.maxn;
    if (i > g.maxl) return;
    phi(i, j, 0) = T(2) * phi(i, j, 1) - phi(i, j, 2);
    phi(i, j, n) = T(2) * phi(i, j, n - 1) - phi(i, j, n - 2);
}

// levelset.f90:44-91: sign, interface indicator and subcell distance.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_levelset_status(Grid g, T h, FieldView<T> phi, FieldView<T> D,
                                                            FieldView<int> ist, FieldView<T> sg)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
    T c = phi(i, j, k);
    T xp = phi(i + 1, j, k), xm = phi(i - 1, j, k), yp = phi(i, jp, k), ym = phi(i, jm, k);
    T zp = phi(i, j, k + 1), zm = phi(i, j, k - 1);
    sg(i, j, k) = c < T(0) ? T(-1) : (c > T(0) ? T(1) : T(0));
    T mn = fmin(fmin(fmin(fmin(fmin(c * xp, c * xm), c * yp), c * ym), c * zp), c * zm);
    int s = mn <= T(0) ? 1 : 0;
    ist(i, j, k) = s;
    if (s) {
        T gx = xp - xm, gy = yp - ym, gz = zp - zm;
        T a1 = sqrt(gx * gx + gy * gy + gz * gz) * T(0.5);
        T mx = fmax(fmax(fmax(fmax(fmax(fmax(fmax(a1, fabs(xp - c)), fabs(xm - c)), fabs(yp - c)), fabs(ym - c)),
                             fabs(zp - c)),
                        fabs(zm - c)),
                   h);
        D(i, j, k) = h * c / mx;
    }
}

// levelset.f90:125-166: one pseudo-time step, dtau = 0.3 h.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_levelset_iter(Grid g, T h, FieldView<T> phi, FieldView<T> D,
                                                          FieldView<int> ist, FieldView<T> sg,
                                                          FieldView<T> pr, FieldView<T> pn)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const T dtau = T(0.3) * h;
    T c = pr(i, j, k);
    T dc;
    if (ist(i, j, k) == 1) {
        dc = sg(i, j, k) * fabs(c) - D(i, j, k);
    } else {
        dc = T(0);
        const T p0 = phi(i, j, k);
        if (p0 != T(0)) {
            const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
            T dxm = c - pr(i - 1, j, k), dxp = pr(i + 1, j, k) - c;
            T dym = c - pr(i, jm, k), dyp = pr(i, jp, k) - c;
            T dzm = c - pr(i, j, k - 1), dzp = pr(i, j, k + 1) - c;
            T gx, gy, gz;
            if (p0 > T(0)) {
                T aP = fmax(dxm, T(0)), bM = fmin(dxp, T(0));
                T cP = fmax(dym, T(0)), dM = fmin(dyp, T(0));
                T eP = fmax(dzm, T(0)), fM = fmin(dzp, T(0));
                gx = fmax(aP * aP, bM * bM);
                gy = fmax(cP * cP, dM * dM);
                gz = fmax(eP * eP, fM * fM);
                dc = sg(i, j, k) * (sqrt(gx + gy + gz) - h);
            } else {
                T bP = fmax(dxp, T(0)), aM = fmin(dxm, T(0));
                T dP = fmax(dyp, T(0)), cM = fmin(dym, T(0));
                T fP = fmax(dzp, T(0)), eM = fmin(dzm, T(0));
                gx = fmax(bP * bP, aM * aM);
                gz = fmax(fP * fP, eM * eM);
                gy = fmax(dP * dP, cM * cM);
                dc = sg(i, j, k) * (sqrt(gx + gz + gy) - h);  // levelset.f90:154 sums x, z, y
            }
        }
    }
    pn(i, j, k) = c - dtau * dc / h;
}

// The same two kernels for the case's step, with the per-cell state packed
// into one int: code = ist | s << 1.  For an interface cell (ist = 1) s is
// sign + 1; otherwise s selects k_levelset_iter's branch on phi: 0 phi == 0
// (no change), 1 phi > 0 (sign +1), 2 phi < 0 (sign -1), 3 phi NaN (the
// negative branch with sign 0).  The iteration then reads the iterate, the
// code and, on interface cells only, D: 20 B per cell instead of 44, the same
// values.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_levelset_status_code(Grid g, T h, FieldView<T> phi, FieldView<T> D,
                                                                 FieldView<int> code)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
    T c = phi(i, j, k);
    T xp = phi(i + 1, j, k), xm = phi(i - 1, j, k), yp = phi(i, jp, k), ym = phi(i, jm, k);
    T zp = phi(i, j, k + 1), zm = phi(i, j, k - 1);
    const int sgn = c < T(0) ? -1 : (c > T(0) ? 1 : 0);
    T mn = fmin(fmin(fmin(fmin(fmin(c * xp, c * xm), c * yp), c * ym), c * zp), c * zm);
    int st = mn <= T(0) ? 1 : 0;
    if (st) {
        T gx = xp - xm, gy = yp - ym, gz = zp - zm;
        T a1 = sqrt(gx * gx + gy * gy + gz * gz) * T(0.5);
        T mx = fmax(fmax(fmax(fmax(fmax(fmax(fmax(a1, fabs(xp - c)), fabs(xm - c)), fabs(yp - c)), fabs(ym - c)),
                             fabs(zp - c)),
                        fabs(zm - c)),
                   h);
        D(i, j, k) = h * c / mx;
        code(i, j, k) = 1 | ((sgn + 1) << 1);
    } else {
        const int br = c != T(0) ? (c > T(0) ? 1 : (c < T(0) ? 2 : 3)) : 0;
        code(i, j, k) = br << 1;
    }
}

// One iteration's new value of a cell from its code, iterate c, D and the
// six neighbours (read only for the upwind branches).
template <typename T>
__device__ __forceinline__ T iter_code_value(int cd, T c, T dv, T xm, T xp, T ym, T yp, T zm, T zp, T h)
{
    const T dtau = T(0.3) * h;
    const int sv = cd >> 1;
    T dc;
    if (cd & 1) {
        dc = T(sv - 1) * fabs(c) - dv;
    } else {
        dc = T(0);
        if (sv != 0) {
            const T sg = sv == 1 ? T(1) : (sv == 2 ? T(-1) : T(0));
            T dxm = c - xm, dxp = xp - c;
            T dym = c - ym, dyp = yp - c;
            T dzm = c - zm, dzp = zp - c;
            T gx, gy, gz;
            if (sv == 1) {
                T aP = fmax(dxm, T(0)), bM = fmin(dxp, T(0));
                T cP = fmax(dym, T(0)), dM = fmin(dyp, T(0));
                T eP = fmax(dzm, T(0)), fM = fmin(dzp, T(0));
                gx = fmax(aP * aP, bM * bM);
                gy = fmax(cP * cP, dM * dM);
                gz = fmax(eP * eP, fM * fM);
                dc = sg * (sqrt(gx + gy + gz) - h);
            } else {
                T bP = fmax(dxp, T(0)), aM = fmin(dxm, T(0));
                T dP = fmax(dyp, T(0)), cM = fmin(dym, T(0));
                T fP = fmax(dzp, T(0)), eM = fmin(dzm, T(0));
                gx = fmax(bP * bP, aM * aM);
                gz = fmax(fP * fP, eM * eM);
                gy = fmax(dP * dP, cM * cM);
                dc = sg * (sqrt(gx + gz + gy) - h);  // levelset.f90:154 sums x, z, y
            }
        }
    }
    return c - dtau * dc / h;
}

template <typename T>
__global__ void __launch_bounds__(BX* BY) k_levelset_iter_code(Grid g, T h, FieldView<T> D, FieldView<int> code,
                                                               FieldView<T> pr, FieldView<T> pn)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const int cd = code(i, j, k);
    const T c = pr(i, j, k);
    const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
    const T dv = (cd & 1) ? T(D(i, j, k)) : T(0);
    pn(i, j, k) = iter_code_value<T>(cd, c, dv, pr(i - 1, j, k), pr(i + 1, j, k), pr(i, jm, k), pr(i, jp, k),
                                     pr(i, j, k - 1), pr(i, j, k + 1), h);
}

// The same iteration marching a 64 x 8 tile up a chunk of levels: the
// iterate's planes (tile and a one-cell ring, rows wrapped) are loaded once
// per block into an LDS ring of four, D prefetched levels ahead of use like
// the code (on every cell: a load under the cell's condition would drain the
// memory counters at every level).  Same values as k_levelset_iter_code.
constexpr int LS_TY = 8;
template <typename T, int D>
__global__ void __launch_bounds__(BX* LS_TY) k_levelset_iter_st(Grid g, int nbx, int nby, int ks, T h, FieldView<T> Dd,
                                                                FieldView<int> code, FieldView<T> pr,
                                                                FieldView<T> pn)
{
    constexpr int R = LS_TY + 2, W = BX + 2, N = R * W, NT = BX * LS_TY, NS = (N + NT - 1) / NT;
    __shared__ T ps[4][R][W];
    const int t = xcd_tile(int(blockIdx.x), int(gridDim.x));
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int i0 = 1 + bx * BX, j0 = 1 + by * LS_TY;
    const int kb = 1 + bz * ks, ke = min(kb + ks - 1, g.maxn - 1);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int i = i0 + tx, j = j0 + ty;  // j <= P: the tile rows divide the period
    const bool mine = i <= g.maxl - 1;
    const int ic = min(i, g.maxl - 1);
    long off[NS];
    int sl[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int e = min(tid + n * NT, N - 1);
        const int r = e / W, c = e % W;
        sl[n] = e;
        off[n] = min(max(i0 - 1 + c, 0), g.maxl) + long(wrapj(j0 - 1 + r, g.P)) * pr.sy;
    }
    auto load = [&](T* v, int k) {
        const long kz = long(min(max(k, 0), g.maxn)) * pr.sz;
#pragma unroll
        for (int n = 0; n < NS; ++n) v[n] = pr.p[off[n] + kz];
    };
    auto store = [&](const T* v, int k) {
        T* q = &ps[k & 3][0][0];
#pragma unroll
        for (int n = 0; n < NS; ++n) q[sl[n]] = v[n];
    };
    {
        T v[NS];
#pragma unroll
        for (int k = kb - 1; k <= kb + 1; ++k) {
            load(v, k);
            store(v, k);
        }
    }
    T rx[D][NS], dv[D];
    int cd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        load(rx[d], kb + 2 + d);
        const int kk = min(kb + d, ke);
        cd[d] = code(ic, j, kk);
        dv[d] = Dd(ic, j, kk);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int kend = kb + (ke - kb + D) / D * D;
    for (int sb = kb; sb < kend; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int k = sb + d, b = k & 3;
            store(rx[d], k + 2);
            load(rx[d], k + 2 + D);
            const int cv = cd[d];
            const T dd = dv[d];
            const int kn = min(k + D, ke);
            cd[d] = code(ic, j, kn);
            dv[d] = Dd(ic, j, kn);
            const T v = iter_code_value<T>(cv, T(ps[b][ty + 1][tx + 1]), (cv & 1) ? dd : T(0), T(ps[b][ty + 1][tx]),
                                           T(ps[b][ty + 1][tx + 2]), T(ps[b][ty][tx + 1]), T(ps[b][ty + 2][tx + 1]),
                                           T(ps[(k - 1) & 3][ty + 1][tx + 1]), T(ps[(k + 1) & 3][ty + 1][tx + 1]), h);
            if (mine && k <= ke) pn(i, j, k) = v;
            __syncthreads();
        }
    }
}

// levelset.f90:270-363: curvature (clipped above at 1/min(dx,dy,dz)) and the
// CSF force delta(phi) * scap * kappa * grad(phi), p interior.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_csf(Grid g, T dx, T dy, T dz, T scap, T sw, FieldView<T> fi,
                                                FieldView<T> fx, FieldView<T> fy, FieldView<T> fz)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.maxl - 1 || j > g.P) return;
    const int jp = wrapj(j + 1, g.P), jm = wrapj(j - 1, g.P);
    T c = fi(i, j, k);
    T xp = fi(i + 1, j, k), xm = fi(i - 1, j, k), yp = fi(i, jp, k), ym = fi(i, jm, k);
    T zp = fi(i, j, k + 1), zm = fi(i, j, k - 1);
    T d2x = (xp - T(2) * c + xm) / (dx * dx);
    T d2y = (yp - T(2) * c + ym) / (dy * dy);
    T d2z = (zp - T(2) * c + zm) / (dz * dz);
    T gx = (xp - xm) / (T(2) * dx);
    T gy = (yp - ym) / (T(2) * dy);
    T gz = (zp - zm) / (T(2) * dz);
    T dxz = ((fi(i + 1, j, k + 1) - fi(i - 1, j, k + 1)) / (T(2) * dx) -
             (fi(i + 1, j, k - 1) - fi(i - 1, j, k - 1)) / (T(2) * dx)) /
            (T(2) * dz);
    T dyz = ((fi(i, jp, k + 1) - fi(i, jm, k + 1)) / (T(2) * dy) - (fi(i, jp, k - 1) - fi(i, jm, k - 1)) / (T(2) * dy)) /
            (T(2) * dz);
    T dxy = ((fi(i + 1, jp, k) - fi(i - 1, jp, k)) / (T(2) * dx) - (fi(i + 1, jm, k) - fi(i - 1, jm, k)) / (T(2) * dx)) /
            (T(2) * dy);
    T nrm = gx * gx + gy * gy + gz * gz;
    T kap;
    if (nrm < T(1e-10)) {
        kap = T(0);
    } else {
        T rn = sqrt(nrm);
        T t1 = (d2x + d2y + d2z) / rn;
        T t2 = gx * gx * d2x + gy * gy * d2y + gz * gz * d2z;
        T t3 = gx * gy * dxy + gx * gz * dxz + gy * gz * dyz;
        T n15 = nrm * rn;
        t2 = t2 / n15;
        t3 = T(2) * t3 / n15;
        kap = -(t1 - t2 - t3);
    }
    const T max_kappa = T(1) / fmin(fmin(dx, dy), dz);
    if (kap > max_kappa) kap = max_kappa;
    T dirac;
    if (c < -sw || c > sw)
        dirac = T(0);
    else
        dirac = (T(1) / (T(2) * sw)) + (T(1) / (T(2) * sw)) * cos(T(TPLS_PI_SP) * c / sw);
    T base = dirac * scap * kap;
    fx(i, j, k) = base * gx;
    fy(i, j, k) = base * gy;
    fz(i, j, k) = base * gz;
}

// levelset.f90:365-412: z extrapolation on the interior columns, then x
// extrapolation of whole inlet/outlet columns (which the column threads at
// i = 0 and i = maxl rebuild from their neighbours' extrapolated values).
template <typename T>
__device__ __forceinline__ T zext(const FieldView<T>& f, int i, int j, int k, int n)
{
    if (k == 0) return T(2) * f(i, j, 1) - f(i, j, 2);
    if (k == n) return T(2) * f(i, j, n - 1) - f(i, j, n - 2);
    return f(i, j, k);
}

// One kernel, two independent index segments: threads [0, (maxl-1)*P) do
// the z extrapolation of interior column (i, j); threads
// [(maxl-1)*P, (maxl-1)*P + (maxn+1)*P) do the x extrapolation of (j, k) at
// both ends, rebuilding the extrapolated k = 0, n values of columns 1, 2,
// maxl-2, maxl-1 locally.
template <typename T>
__global__ void k_csf_bc(Grid g, FieldView<T> fx, FieldView<T> fy, FieldView<T> fz)
{
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int L = g.maxl, n = g.maxn, P = g.P;
    const long n1 = long(L - 1) * P, n2 = long(n + 1) * P;
    FieldView<T> fs[3] = {fx, fy, fz};
    if (q < n1) {
        const int i = int(q % (L - 1)) + 1, j = int(q / (L - 1)) + 1;
        for (int c = 0; c < 3; ++c) {
            const FieldView<T>& f = fs[c];
            T a = T(2) * f(i, j, 1) - f(i, j, 2);
            T b = T(2) * f(i, j, n - 1) - f(i, j, n - 2);
            f(i, j, 0) = a;
            f(i, j, n) = b;
        }
    } else if (q < n1 + n2) {
        const long r = q - n1;
        const int k = int(r % (n + 1)), j = int(r / (n + 1)) + 1;
        for (int c = 0; c < 3; ++c) {
            const FieldView<T>& f = fs[c];
            f(0, j, k) = T(2) * zext(f, 1, j, k, n) - zext(f, 2, j, k, n);
            f(L, j, k) = T(2) * zext(f, L - 1, j, k, n) - zext(f, L - 2, j, k, n);
        }
    }
}

static inline dim3 grid3(const Grid& g, int nzb) { return dim3((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, nzb); }

template <typename T>
hipError_t launch_tpls_update(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi2, FieldView<T> c0,
                              FieldView<T> c1, FieldView<T> c2, FieldView<T> phi3)
{
    hipLaunchKernelGGL(k_tpls_update<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(ph.dt), phi2, c0, c1, c2,
                       phi3);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_phi_bc(hipStream_t s, const Grid& g, double dz, double height, FieldView<T> phi, int inlet_outlet)
{
    if (inlet_outlet)
        hipLaunchKernelGGL(k_phi_inout<T>, dim3((g.maxn - 1 + 63) / 64, g.P), dim3(64), 0, s, g, T(dz), T(height), phi);
    hipLaunchKernelGGL(k_phi_zext<T>, dim3((g.maxl + 1 + 63) / 64, g.P), dim3(64), 0, s, g, phi);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_levelset_status(hipStream_t s, const Grid& g, double h, FieldView<T> phi, FieldView<T> D,
                                  FieldView<int> ist, FieldView<T> sign)
{
    hipLaunchKernelGGL(k_levelset_status<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(h), phi, D, ist, sign);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_levelset_iter(hipStream_t s, const Grid& g, double h, FieldView<T> phi, FieldView<T> D,
                                FieldView<int> ist, FieldView<T> sign, FieldView<T> pr, FieldView<T> pr_new)
{
    hipLaunchKernelGGL(k_levelset_iter<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(h), phi, D, ist, sign,
                       pr, pr_new);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_levelset_status_code(hipStream_t s, const Grid& g, double h, FieldView<T> phi, FieldView<T> D,
                                       FieldView<int> code)
{
    hipLaunchKernelGGL(k_levelset_status_code<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(h), phi, D, code);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_levelset_iter_code(hipStream_t s, const Grid& g, double h, FieldView<T> D, FieldView<int> code,
                                     FieldView<T> pr, FieldView<T> pr_new)
{
    if (g.P % LS_TY == 0) {
        // the staged march: tiles of 64 x 8 columns, chunks of levels that
        // fill the last round of resident blocks
        constexpr int DP = 4;
        static const int slots = resident_blocks(k_levelset_iter_st<T, DP>, BX * LS_TY);
        const int nbx = (g.maxl - 1 + BX - 1) / BX, nby = g.P / LS_TY, nz = g.maxn - 1;
        const int ks = march_chunk(nbx * nby, nz, slots, 32);
        const int nbz = (nz + ks - 1) / ks;
        hipLaunchKernelGGL((k_levelset_iter_st<T, DP>), dim3(nbx * nby * nbz), dim3(BX, LS_TY), 0, s, g, nbx, nby, ks,
                           T(h), D, code, pr, pr_new);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_levelset_iter_code<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(h), D, code, pr,
                       pr_new);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_csf(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi, FieldView<T> fx,
                      FieldView<T> fy, FieldView<T> fz)
{
    hipLaunchKernelGGL(k_csf<T>, grid3(g, g.maxn - 1), dim3(BX, BY), 0, s, g, T(ph.dx), T(ph.dy), T(ph.dz),
                       T(ph.scap), T(ph.sw), phi, fx, fy, fz);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_csf_bc(hipStream_t s, const Grid& g, FieldView<T> fx, FieldView<T> fy, FieldView<T> fz)
{
    const long n = long(g.maxl - 1) * g.P + long(g.maxn + 1) * g.P;
    hipLaunchKernelGGL(k_csf_bc<T>, dim3((n + 255) / 256), dim3(256), 0, s, g, fx, fy, fz);
    return hipGetLastError();
}

#define TPLS_INST_LS(T)                                                                                              \
    template hipError_t launch_tpls_update<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, FieldView<T>,     \
                                              FieldView<T>, FieldView<T>, FieldView<T>);                             \
    template hipError_t launch_phi_bc<T>(hipStream_t, const Grid&, double, double, FieldView<T>, int);               \
    template hipError_t launch_levelset_status<T>(hipStream_t, const Grid&, double, FieldView<T>, FieldView<T>,      \
                                                  FieldView<int>, FieldView<T>);                                     \
    template hipError_t launch_levelset_iter<T>(hipStream_t, const Grid&, double, FieldView<T>, FieldView<T>,        \
                                                FieldView<int>, FieldView<T>, FieldView<T>, FieldView<T>);           \
    template hipError_t launch_levelset_status_code<T>(hipStream_t, const Grid&, double, FieldView<T>,            \
                                                       FieldView<T>, FieldView<int>);                              \
    template hipError_t launch_levelset_iter_code<T>(hipStream_t, const Grid&, double, FieldView<T>, FieldView<int>, \
                                                     FieldView<T>, FieldView<T>);                                  \
    template hipError_t launch_csf<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, FieldView<T>,             \
                                      FieldView<T>, FieldView<T>);                                                   \
    template hipError_t launch_csf_bc<T>(hipStream_t, const Grid&, FieldView<T>, FieldView<T>, FieldView<T>);

TPLS_INST_LS(double)
TPLS_INST_LS(float)

}  // namespace tpls
