// Momentum flux assembly: momentum_allflux.f90:22-521 and the right-hand
// sides of main_ns_hybrid.F90:629-757.
//
// The fused kernel (mode bit 4) produces, per face cell, the convective term
// conv2, the diffusive term (kept in registers), the CSF face average csf2 and
// the AB3 + Crank-Nicolson right-hand side in one pass.
#include "tiles.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

template <typename T>
__global__ void __launch_bounds__(BX* BY) k_viscosity(Grid g, T sw, T Re, T mum, T mup, FieldView<T> phi,
                                                      FieldView<T> mu, int ilo, int ihi, int shift)
{
    const int i = blockIdx.x * BX + threadIdx.x + ilo;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i > ihi || j > g.P) return;
    T pv = phi(i, j, k);
    if (shift) pv = pv - T(0.5);
    mu(i, j, k) = visc_of_phi<T>(pv, sw, Re, mum, mup);
}

// A grid spacing as the stencils divide by it: a / h is the correctly rounded
// division in FP64 (the reference's operation) and a times the hardware
// reciprocal of h (within an ulp) in FP32; 2 * h is a spacing too.
template <typename C>
struct Sp {
    C h, r;
};
template <typename C>
__device__ __forceinline__ Sp<C> spacing(C h)
{
    return Sp<C>{h, recip(h)};
}
template <typename C>
__device__ __forceinline__ C operator/(C a, Sp<C> s)
{
    if constexpr (sizeof(C) == 4)
        return a * s.r;
    else
        return a / s.h;
}
template <typename C>
__device__ __forceinline__ Sp<C> operator*(C f, Sp<C> s)
{
    return spacing(f * s.h);
}

// ---- convective terms, momentum_allflux.f90:281-337 / 350-402 / 415-463 ----
template <typename C, typename T>
__device__ __forceinline__ C conv_u(const T& u2, const T& v2, const T& w2, const T& mu, int i, int j, int k, int maxn,
                                    int P, Sp<C> dx, Sp<C> dy, Sp<C> dz)
{
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    const int j1 = jp, j2 = wrapj(j + 2, P), j0 = j;
    C tv = (C(v2(i, j, k)) + C(v2(i + 1, j, k)) + C(v2(i, jp, k)) + C(v2(i + 1, jp, k))) * C(0.25);
    C tw = (C(w2(i, j, k)) + C(w2(i + 1, j, k)) + C(w2(i, j, k + 1)) + C(w2(i + 1, j, k + 1))) * C(0.25);
    C uc = u2(i, j, k), uxp = u2(i + 1, j, k), uxm = u2(i - 1, j, k);
    C dudx = C(0.5) * (uxp - uxm) / dx;
    C Xc = dudx * uc;
    C dudy = C(0.5) * (C(u2(i, jp, k)) - C(u2(i, jm, k))) / dy;
    C Yc = dudy * tv;
    C dudz;
    if (k == 0)
        dudz = (uc + C(u2(i, j, k + 1))) / (C(2) * dz);
    else if (k == maxn - 2)
        dudz = -(uc + C(u2(i, j, k - 1))) / (C(2) * dz);
    else
        dudz = C(0.5) * (C(u2(i, j, k + 1)) - C(u2(i, j, k - 1))) / dz;
    C Zc = dudz * tw;
    C c1 = Xc + Yc + Zc;
    C mxp = mu(i + 1, j1, k + 1), mxm = mu(i, j1, k + 1);
    C myp = (C(mu(i + 1, j2, k + 1)) + C(mu(i, j2, k + 1)) + mxp + mxm) * C(0.25);
    C mym = (C(mu(i + 1, j0, k + 1)) + C(mu(i, j0, k + 1)) + mxp + mxm) * C(0.25);
    C mzp = (C(mu(i + 1, j1, k + 2)) + C(mu(i, j1, k + 2)) + mxp + mxm) * C(0.25);
    C mzm = (C(mu(i + 1, j1, k)) + C(mu(i, j1, k)) + mxp + mxm) * C(0.25);
    C aE = mxp * (uxp - uc) / dx;
    C aW = mxm * (uc - uxm) / dx;
    C bN = myp * (C(v2(i + 1, jp, k)) - C(v2(i, jp, k))) / dx;
    C bS = mym * (C(v2(i + 1, j, k)) - C(v2(i, j, k))) / dx;
    C cO = mzp * (C(w2(i + 1, j, k + 1)) - C(w2(i, j, k + 1))) / dx;
    C cI = mzm * (C(w2(i + 1, j, k)) - C(w2(i, j, k))) / dx;
    C c2 = ((aE - aW) / dx) + ((bN - bS) / dy) + ((cO - cI) / dz);
    return c1 - c2;
}

template <typename C, typename T>
__device__ __forceinline__ C conv_v(const T& u2, const T& v2, const T& w2, const T& mu, int i, int j, int k, int maxn,
                                    int P, Sp<C> dx, Sp<C> dy, Sp<C> dz)
{
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P);
    C tu = (C(u2(i - 1, jm, k)) + C(u2(i - 1, j, k)) + C(u2(i, jm, k)) + C(u2(i, j, k))) * C(0.25);
    C tw = (C(w2(i, jm, k)) + C(w2(i, j, k)) + C(w2(i, jm, k + 1)) + C(w2(i, j, k + 1))) * C(0.25);
    C vc = v2(i, j, k), vyp = v2(i, jp, k), vym = v2(i, jm, k);
    C dvdy = C(0.5) * (vyp - vym) / dy;
    C Yc = dvdy * vc;
    C dvdx = C(0.5) * (C(v2(i + 1, j, k)) - C(v2(i - 1, j, k))) / dx;
    C Xc = dvdx * tu;
    C dvdz;
    if (k == 0)
        dvdz = (vc + C(v2(i, j, k + 1))) / (C(2) * dz);
    else if (k == maxn - 2)
        dvdz = -(vc + C(v2(i, j, k - 1))) / (C(2) * dz);
    else
        dvdz = C(0.5) * (C(v2(i, j, k + 1)) - C(v2(i, j, k - 1))) / dz;
    C Zc = dvdz * tw;
    C c1 = Xc + Yc + Zc;
    C a = mu(i, jp, k + 1), b = mu(i, j, k + 1);
    C mxp = (C(mu(i + 1, jp, k + 1)) + C(mu(i + 1, j, k + 1)) + a + b) * C(0.25);
    C mxm = (C(mu(i - 1, jp, k + 1)) + C(mu(i - 1, j, k + 1)) + a + b) * C(0.25);
    C myp = a, mym = b;
    C mzp = (C(mu(i, jp, k + 2)) + C(mu(i, j, k + 2)) + a + b) * C(0.25);
    C mzm = (C(mu(i, jp, k)) + C(mu(i, j, k)) + a + b) * C(0.25);
    C aE = mxp * (C(u2(i, j, k)) - C(u2(i, jm, k))) / dy;
    C aW = mxm * (C(u2(i - 1, j, k)) - C(u2(i - 1, jm, k))) / dy;
    C bN = myp * (vyp - vc) / dy;
    C bS = mym * (vc - vym) / dy;
    C cO = mzp * (C(w2(i, j, k + 1)) - C(w2(i, jm, k + 1))) / dy;
    C cI = mzm * (C(w2(i, j, k)) - C(w2(i, jm, k))) / dy;
    C c2 = ((aE - aW) / dx) + ((bN - bS) / dy) + ((cO - cI) / dz);
    return c1 - c2;
}

template <typename C, typename T>
__device__ __forceinline__ C conv_w(const T& u2, const T& v2, const T& w2, const T& mu, int i, int j, int k, int P,
                                    Sp<C> dx, Sp<C> dy, Sp<C> dz)
{
    const int jp = wrapj(j + 1, P), jm = wrapj(j - 1, P), j2 = wrapj(j + 2, P);
    C tu = (C(u2(i - 1, j, k - 1)) + C(u2(i - 1, j, k)) + C(u2(i, j, k - 1)) + C(u2(i, j, k))) * C(0.25);
    C tv = (C(v2(i, j, k - 1)) + C(v2(i, j, k)) + C(v2(i, jp, k - 1)) + C(v2(i, jp, k))) * C(0.25);
    C wc = w2(i, j, k), wzp = w2(i, j, k + 1), wzm = w2(i, j, k - 1);
    C dwdz = C(0.5) * (wzp - wzm) / dz;
    C Zc = dwdz * wc;
    C dwdx = C(0.5) * (C(w2(i + 1, j, k)) - C(w2(i - 1, j, k))) / dx;
    C Xc = dwdx * tu;
    C dwdy = C(0.5) * (C(w2(i, jp, k)) - C(w2(i, jm, k))) / dy;
    C Yc = dwdy * tv;
    C c1 = Xc + Yc + Zc;
    C a = mu(i, jp, k + 1), b = mu(i, jp, k);
    C mxp = (C(mu(i + 1, jp, k + 1)) + C(mu(i + 1, jp, k)) + a + b) * C(0.25);
    C mxm = (C(mu(i - 1, jp, k + 1)) + C(mu(i - 1, jp, k)) + a + b) * C(0.25);
    C myp = (C(mu(i, j2, k + 1)) + C(mu(i, j2, k)) + a + b) * C(0.25);
    C mym = (C(mu(i, j, k + 1)) + C(mu(i, j, k)) + a + b) * C(0.25);
    C mzp = a, mzm = b;
    C aE = mxp * (C(u2(i, j, k)) - C(u2(i, j, k - 1))) / dz;
    C aW = mxm * (C(u2(i - 1, j, k)) - C(u2(i - 1, j, k - 1))) / dz;
    C bN = myp * (C(v2(i, jp, k)) - C(v2(i, jp, k - 1))) / dz;
    C bS = mym * (C(v2(i, j, k)) - C(v2(i, j, k - 1))) / dz;
    C cO = mzp * (wzp - wc) / dz;
    C cI = mzm * (wc - wzm) / dz;
    C c2 = ((aE - aW) / dx) + ((bN - bS) / dy) + ((cO - cI) / dz);
    return c1 - c2;
}

// ---- diffusive terms, momentum_allflux.f90:70-110 / 121-164 / 175-206 ----
template <typename C, typename T>
__device__ __forceinline__ C diff_uv(const T& x2, const MuFaces<C>& m, int i, int j, int k, int maxn,
                                     int P, Sp<C> dx, Sp<C> dy, Sp<C> dz)
{
    C c = x2(i, j, k);
    C xm = (k == 0) ? C(-2) * c + C(x2(i, j, 1)) / C(3) : C(x2(i, j, k - 1));
    C xp = (k == maxn - 2) ? C(-2) * c + C(x2(i, j, maxn - 3)) / C(3) : C(x2(i, j, k + 1));
    C xE = m.xp * ((C(x2(i + 1, j, k)) - c) / dx);
    C xW = m.xm * ((c - C(x2(i - 1, j, k))) / dx);
    C yN = m.yp * ((C(x2(i, wrapj(j + 1, P), k)) - c) / dy);
    C yS = m.ym * ((c - C(x2(i, wrapj(j - 1, P), k))) / dy);
    C zO = m.zp * ((xp - c) / dz);
    C zI = m.zm * ((c - xm) / dz);
    return ((xE - xW) / dx) + ((yN - yS) / dy) + ((zO - zI) / dz);
}

template <typename C, typename T>
__device__ __forceinline__ C diff_w(const T& w2, const MuFaces<C>& m, int i, int j, int k, int P, Sp<C> dx,
                                    Sp<C> dy, Sp<C> dz)
{
    C c = w2(i, j, k);
    C xE = m.xp * ((C(w2(i + 1, j, k)) - c) / dx);
    C xW = m.xm * ((c - C(w2(i - 1, j, k))) / dx);
    C yN = m.yp * ((C(w2(i, wrapj(j + 1, P), k)) - c) / dy);
    C yS = m.ym * ((c - C(w2(i, wrapj(j - 1, P), k))) / dy);
    C zO = m.zp * ((C(w2(i, j, k + 1)) - c) / dz);
    C zI = m.zm * ((c - C(w2(i, j, k - 1))) / dz);
    return ((xE - xW) / dx) + ((yN - yS) / dy) + ((zO - zI) / dz);
}

// One thread per face cell (i in 1..maxl-1, j in 1..P, k over the component's
// full z range).  comp 0/1/2 = u/v/w.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_momentum(Grid g, int comp, int mode, T dx, T dy, T dz, T dt,
                                                     FieldView<T> u2, FieldView<T> v2, FieldView<T> w2,
                                                     FieldView<T> mu, FieldView<T> conv2, FieldView<T> diff,
                                                     FieldView<T> fc, FieldView<T> conv0, FieldView<T> conv1,
                                                     FieldView<T> csf0, FieldView<T> csf1, FieldView<T> csf2,
                                                     FieldView<T> rhs)
{
    using C = T;
    const Sp<C> sx = spacing(dx), sy = spacing(dy), sz = spacing(dz);
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z;
    const int P = g.P, maxn = g.maxn, maxl = g.maxl;
    if (i > maxl - 1 || j > P) return;
    bool in_range;
    if (comp == 0) in_range = (i <= maxl - 2);
    else if (comp == 1) in_range = true;
    else in_range = (k >= 1 && k <= maxn - 2);
    C cv = C(0), df = C(0);
    if (in_range) {
        if (comp == 0) {
            if (mode & 1 || mode & 4) cv = conv_u<C>(u2, v2, w2, mu, i, j, k, maxn, P, sx, sy, sz);
            if (mode & 2 || mode & 4) df = diff_uv<C>(u2, mu_faces_u<C>(mu, i, j, k, P), i, j, k, maxn, P, sx, sy, sz);
        } else if (comp == 1) {
            if (mode & 1 || mode & 4) cv = conv_v<C>(u2, v2, w2, mu, i, j, k, maxn, P, sx, sy, sz);
            if (mode & 2 || mode & 4) df = diff_uv<C>(v2, mu_faces_v<C>(mu, i, j, k, P), i, j, k, maxn, P, sx, sy, sz);
        } else {
            if (mode & 1 || mode & 4) cv = conv_w<C>(u2, v2, w2, mu, i, j, k, P, sx, sy, sz);
            if (mode & 2 || mode & 4) df = diff_w<C>(w2, mu_faces_w<C>(mu, i, j, k, P), i, j, k, P, sx, sy, sz);
        }
    }
    if (mode & 1) conv2(i, j, k) = cv;
    if (mode & 2) diff(i, j, k) = df;
    if (mode & 4) {
        // CSF face average, main_ns_hybrid.F90:629-641
        const int jp = wrapj(j + 1, P);
        C s2;
        if (comp == 0) s2 = (C(fc(i + 1, jp, k + 1)) + C(fc(i, jp, k + 1))) * C(0.5);
        else if (comp == 1) s2 = (C(fc(i, jp, k + 1)) + C(fc(i, j, k + 1))) * C(0.5);
        else s2 = (C(fc(i, jp, k + 1)) + C(fc(i, jp, k))) * C(0.5);
        csf2(i, j, k) = s2;
        conv2(i, j, k) = cv;
        // AB3 + Crank-Nicolson, main_ns_hybrid.F90:743-751
        const FieldView<T>& x2 = comp == 0 ? u2 : (comp == 1 ? v2 : w2);
        C r = C(x2(i, j, k)) -
              dt * ((C(23.0 / 12.0)) * cv - (C(4.0 / 3.0)) * C(conv1(i, j, k)) + (C(5.0 / 12.0)) * C(conv0(i, j, k))) +
              (dt * C(0.5)) * df +
              dt * ((C(23.0 / 12.0)) * s2 - (C(4.0 / 3.0)) * C(csf1(i, j, k)) + (C(5.0 / 12.0)) * C(csf0(i, j, k)));
        rhs(i, j, k) = r;
    }
}

// The fused assembly (mode 4) marching a tile of 64 x MS_TY columns up a
// chunk of levels: u2, v2, w2, the viscosity and the component's CSF force
// are staged plane by plane into LDS rings (the tile and the rows and
// columns its stencils reach: i-1 .. i+1, j-1 .. j+2), each value loaded from
// HBM once per block instead of once per stencil point, and the stencil
// functions read them there through MsRing.  Same arithmetic and values as
// k_momentum.
constexpr int MS_TY = 4, MS_R = MS_TY + 3, MS_W = BX + 2, MS_N = MS_R * MS_W, MS_NT = BX * MS_TY;
constexpr int MS_NS = (MS_N + MS_NT - 1) / MS_NT;

// A staged array: plane k in slot k & 3, row r for j = j0 - 1 + r (wrapped),
// column c for i = i0 - 1 + c.
template <typename C>
struct MsRing {
    const C* s;
    int i0, j0, P;
    __device__ __forceinline__ C operator()(int i, int j, int k) const
    {
        int r = j - j0 + 1;
        r = r < 0 ? r + P : (r >= MS_R ? r - P : r);
        return s[((k & 3) * MS_R + r) * MS_W + (i - i0 + 1)];
    }
};

template <int COMP, typename T, int D>
__global__ void __launch_bounds__(MS_NT) k_momentum_st(Grid g, int nbx, int nby, int ks, T dx, T dy, T dz, T dt,
                                                       FieldView<T> u2, FieldView<T> v2, FieldView<T> w2,
                                                       FieldView<T> mu, FieldView<T> fc, FieldView<T> conv0,
                                                       FieldView<T> conv1, FieldView<T> csf0, FieldView<T> csf1,
                                                       FieldView<T> conv2, FieldView<T> csf2, FieldView<T> rhs)
{
    using C = T;
    constexpr int NA = 5;  // u2, v2, w2, mu, fc
    __shared__ C st[NA][4][MS_R][MS_W];
    const Sp<C> sx = spacing(dx), sy = spacing(dy), sz = spacing(dz);
    const int t = xcd_tile(int(blockIdx.x), int(gridDim.x));
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int i0 = 1 + bx * BX, j0 = 1 + by * MS_TY;
    const int P = g.P, maxn = g.maxn, maxl = g.maxl;
    const int nz = COMP == 2 ? maxn : maxn - 1;
    const int kb = bz * ks, ke = min(kb + ks - 1, nz - 1);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int i = i0 + tx, j = j0 + ty;  // j <= P: the tile rows divide the period
    const bool mine = i <= maxl - 1;
    // staged element offsets (clamped columns, wrapped rows) and plane tops
    // of each array (u2, v2: maxn-2; w2: maxn-1; mu, fc: maxn)
    int sc[MS_NS], sr[MS_NS], sl[MS_NS];
    const FieldView<T> A[NA] = {u2, v2, w2, mu, fc};
    const int top[NA] = {maxn - 2, maxn - 2, maxn - 1, maxn, maxn};
#pragma unroll
    for (int n = 0; n < MS_NS; ++n) {
        const int e = min(tid + n * MS_NT, MS_N - 1);
        const int r = e / MS_W, c = e % MS_W;
        sl[n] = e;
        sc[n] = min(max(i0 - 1 + c, 0), maxl);
        sr[n] = wrapj(j0 - 1 + r, P);
    }
    // stage s: plane s of u2, v2, w2 and fc, plane s + 1 of mu
    auto stage_load = [&](C (*v)[MS_NS], int s) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int pl = min(max(a == 3 ? s + 1 : s, 0), top[a]);
            const long kz = long(pl) * A[a].sz;
#pragma unroll
            for (int n = 0; n < MS_NS; ++n) v[a][n] = A[a].p[sc[n] + long(sr[n]) * A[a].sy + kz];
        }
    };
    auto stage_store = [&](C (*v)[MS_NS], int s) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            C* q = &st[a][(a == 3 ? s + 1 : s) & 3][0][0];
#pragma unroll
            for (int n = 0; n < MS_NS; ++n) q[sl[n]] = v[a][n];
        }
    };
    const MsRing<C> U{&st[0][0][0][0], i0, j0, P}, V{&st[1][0][0][0], i0, j0, P}, W{&st[2][0][0][0], i0, j0, P},
        M{&st[3][0][0][0], i0, j0, P}, F{&st[4][0][0][0], i0, j0, P};
    {
        C v[NA][MS_NS];
#pragma unroll
        for (int s = kb - 1; s <= kb + 1; ++s) {
            stage_load(v, s);
            stage_store(v, s);
        }
    }
    C rv[D][NA][MS_NS];
#pragma unroll
    for (int d = 0; d < D; ++d) stage_load(rv[d], kb + 2 + d);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int kend = kb + (ke - kb + D) / D * D;
    for (int sb = kb; sb < kend; sb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int k = sb + d;
            stage_store(rv[d], k + 2);
            stage_load(rv[d], k + 2 + D);
            if (mine && k <= ke) {
                bool in_range;
                if (COMP == 0) in_range = (i <= maxl - 2);
                else if (COMP == 1) in_range = true;
                else in_range = (k >= 1 && k <= maxn - 2);
                C cv = C(0), df = C(0);
                if (in_range) {
                    if (COMP == 0) {
                        cv = conv_u<C>(U, V, W, M, i, j, k, maxn, P, sx, sy, sz);
                        df = diff_uv<C>(U, mu_faces_u<C>(M, i, j, k, P), i, j, k, maxn, P, sx, sy, sz);
                    } else if (COMP == 1) {
                        cv = conv_v<C>(U, V, W, M, i, j, k, maxn, P, sx, sy, sz);
                        df = diff_uv<C>(V, mu_faces_v<C>(M, i, j, k, P), i, j, k, maxn, P, sx, sy, sz);
                    } else {
                        cv = conv_w<C>(U, V, W, M, i, j, k, P, sx, sy, sz);
                        df = diff_w<C>(W, mu_faces_w<C>(M, i, j, k, P), i, j, k, P, sx, sy, sz);
                    }
                }
                // CSF face average, main_ns_hybrid.F90:629-641
                const int jp = wrapj(j + 1, P);
                C s2;
                if (COMP == 0) s2 = (C(F(i + 1, jp, k + 1)) + C(F(i, jp, k + 1))) * C(0.5);
                else if (COMP == 1) s2 = (C(F(i, jp, k + 1)) + C(F(i, j, k + 1))) * C(0.5);
                else s2 = (C(F(i, jp, k + 1)) + C(F(i, jp, k))) * C(0.5);
                csf2(i, j, k) = s2;
                conv2(i, j, k) = cv;
                // AB3 + Crank-Nicolson, main_ns_hybrid.F90:743-751
                const C xc = COMP == 0 ? U(i, j, k) : (COMP == 1 ? V(i, j, k) : W(i, j, k));
                C r = xc -
                      dt * ((C(23.0 / 12.0)) * cv - (C(4.0 / 3.0)) * C(conv1(i, j, k)) + (C(5.0 / 12.0)) * C(conv0(i, j, k))) +
                      (dt * C(0.5)) * df +
                      dt * ((C(23.0 / 12.0)) * s2 - (C(4.0 / 3.0)) * C(csf1(i, j, k)) + (C(5.0 / 12.0)) * C(csf0(i, j, k)));
                rhs(i, j, k) = r;
            }
            __syncthreads();
        }
    }
}

template <typename T>
hipError_t launch_viscosity(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi, FieldView<T> mu, int ilo,
                            int ihi, int shift)
{
    dim3 grd((ihi - ilo + 1 + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn + 1);
    hipLaunchKernelGGL(k_viscosity<T>, grd, dim3(BX, BY), 0, s, g, T(ph.sw), T(ph.Re), T(ph.mu_minus), T(ph.mu_plus),
                       phi, mu, ilo, ihi, shift);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_momentum(hipStream_t s, const Grid& g, const Phys& ph, int comp, int mode, FieldView<T> u2,
                           FieldView<T> v2, FieldView<T> w2, FieldView<T> mu, FieldView<T> conv2, FieldView<T> diff,
                           FieldView<T> f_csf, FieldView<T> conv0, FieldView<T> conv1, FieldView<T> csf0,
                           FieldView<T> csf1, FieldView<T> csf2, FieldView<T> rhs)
{
    const int nz = comp == 2 ? g.maxn : g.maxn - 1;
    // the fused assembly staged, where the tile rows divide the period
    if (mode == 4 && g.P % MS_TY == 0) {
        constexpr int D = 2;
        auto go = [&](auto kern) {
            static const int slots = resident_blocks(kern, MS_NT);
            const int nbx = (g.maxl - 1 + BX - 1) / BX, nby = g.P / MS_TY;
            const int ks = march_chunk(nbx * nby, nz, slots, 32);
            const int nbz = (nz + ks - 1) / ks;
            hipLaunchKernelGGL(kern, dim3(nbx * nby * nbz), dim3(BX, MS_TY), 0, s, g, nbx, nby, ks, T(ph.dx),
                               T(ph.dy), T(ph.dz), T(ph.dt), u2, v2, w2, mu, f_csf, conv0, conv1, csf0, csf1, conv2,
                               csf2, rhs);
            return hipGetLastError();
        };
        if (comp == 0) return go(k_momentum_st<0, T, D>);
        if (comp == 1) return go(k_momentum_st<1, T, D>);
        return go(k_momentum_st<2, T, D>);
    }
    dim3 grd((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, nz);
    hipLaunchKernelGGL(k_momentum<T>, grd, dim3(BX, BY), 0, s, g, comp, mode, T(ph.dx), T(ph.dy), T(ph.dz), T(ph.dt),
                       u2, v2, w2, mu, conv2, diff, f_csf, conv0, conv1, csf0, csf1, csf2, rhs);
    return hipGetLastError();
}

#define TPLS_INST_MOM(T)                                                                                          \
    template hipError_t launch_viscosity<T>(hipStream_t, const Grid&, const Phys&, FieldView<T>, FieldView<T>,   \
                                            int, int, int);                                                     \
    template hipError_t launch_momentum<T>(hipStream_t, const Grid&, const Phys&, int, int, FieldView<T>,        \
                                           FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, \
                                           FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, FieldView<T>, \
                                           FieldView<T>, FieldView<T>);
TPLS_INST_MOM(double)
TPLS_INST_MOM(float)

}  // namespace tpls
