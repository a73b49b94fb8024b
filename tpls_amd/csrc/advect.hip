// WENO5 level-set advection: weno.f90:18-215 and advect_phi.f90:19-119.
//
// The fused kernel computes conv2_phi = div(u * phi_face) on every p cell in
// one pass: the face values a cell needs are reconstructed in registers from
// phi2 (read with periodic y wrap, which is what the reference's phi_augaug
// plus exchange2d_augaug1/2/3 provide) and shared between the neighbouring
// cells of a block, instead of being written to and re-read from dfx/dfy/dfz
// arrays.  HBM traffic per cell: phi2, u2, v2,
// w2 read once, conv2_phi written once (40 B in FP64).
#include "tiles.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

// n / d in FP64 from an approximate reciprocal r of d: one Newton step on r,
// then one correction of the quotient with its exact residual (fused
// multiply-adds); within an ulp of the correctly rounded quotient, in six
// instructions instead of the ten of a correctly rounded division.
__device__ __forceinline__ double div_nr(double n, double d, double r)
{
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    const double q = n * r;
    return __builtin_fma(r, __builtin_fma(-d, q, n), q);
}

// WENO5 with one division: a_k = c_k / q_k^2, w_k = a_k / sum(a) rewritten
// over the common denominator sum_n c_n prod_{m!=n} q_m^2.  Same smoothness
// indicators and candidate stencils as weno.f90:190-213; rounding differs from
// the reference by a few ulp.
template <typename C>
__device__ __forceinline__ C weno5f(C v1, C v2, C v3, C v4, C v5)
{
    const C eps = C(TPLS_EPSLON);
    const C c1312 = C(13.0 / 12.0);
    C t1 = v1 - C(2) * v2 + v3;
    C t2 = v1 - C(4) * v2 + C(3) * v3;
    C q1 = eps + (c1312 * t1 * t1 + C(0.25) * t2 * t2);
    t1 = v2 - C(2) * v3 + v4;
    t2 = v2 - v4;
    C q2 = eps + (c1312 * t1 * t1 + C(0.25) * t2 * t2);
    t1 = v3 - C(2) * v4 + v5;
    t2 = C(3) * v3 - C(4) * v4 + v5;
    C q3 = eps + (c1312 * t1 * t1 + C(0.25) * t2 * t2);
    q1 *= q1;
    q2 *= q2;
    q3 *= q3;
    C n1 = C(0.1) * (q2 * q3);
    C n2 = C(0.6) * (q1 * q3);
    C n3 = C(0.3) * (q1 * q2);
    C e1 = v1 * C(1.0 / 3.0) - v2 * C(7.0 / 6.0) + v3 * C(11.0 / 6.0);
    C e2 = -(v2 * C(1.0 / 6.0)) + v3 * C(5.0 / 6.0) + v4 * C(1.0 / 3.0);
    C e3 = v3 * C(1.0 / 3.0) + v4 * C(5.0 / 6.0) - v5 * C(1.0 / 6.0);
    const C num = n1 * e1 + n2 * e2 + n3 * e3, den = n1 + n2 + n3;
    if constexpr (sizeof(C) == 4)
        return __fdividef(num, den);  // FP32 state: the hardware reciprocal (within 2 ulp)
    else
        return div_nr(num, den, __builtin_amdgcn_rcp(den));
}

// Face value on the u face I (between p cells I and I+1) at p row pj, level
// pk; uu = u at that face.  weno.f90:45-76 (WENO for 3 <= I <= maxl-3,
// first-order upwind otherwise, which also covers the inlet face of :86-100).
// The upwind direction selects the five stencil points and one WENO5
// evaluation follows, so lanes of mixed sign do not run both.
template <typename C, typename T>
__device__ __forceinline__ C face_x(const FieldView<T>& phi, int I, int pj, int pk, C uu, int maxl)
{
    const bool up = uu >= C(0);
    if (I >= 3 && I <= maxl - 3) {
        const int s = up ? 1 : -1, b = up ? I : I + 1;
        return weno5f<C>(phi(b - 2 * s, pj, pk), phi(b - s, pj, pk), phi(b, pj, pk), phi(b + s, pj, pk),
                         phi(b + 2 * s, pj, pk));
    }
    return C(phi(up ? I : I + 1, pj, pk));
}

// v face J (between p rows J and J+1), weno.f90:104-133: always WENO.
template <typename C, typename T>
__device__ __forceinline__ C face_y(const FieldView<T>& phi, int I, int J, int pk, C vv, int P)
{
    const int s = vv >= C(0) ? 1 : -1, b = vv >= C(0) ? J : J + 1;
    return weno5f<C>(phi(I, wrapj(b - 2 * s, P), pk), phi(I, wrapj(b - s, P), pk), phi(I, wrapj(b, P), pk),
                     phi(I, wrapj(b + s, P), pk), phi(I, wrapj(b + 2 * s, P), pk));
}

// w face K (between p levels K and K+1), weno.f90:135-167: WENO for
// 3 <= K <= maxn-3, central average next to the walls.
template <typename C, typename T>
__device__ __forceinline__ C face_z(const FieldView<T>& phi, int I, int pj, int K, C ww, int maxn)
{
    if (K >= 3 && K <= maxn - 3) {
        const int s = ww >= C(0) ? 1 : -1, b = ww >= C(0) ? K : K + 1;
        return weno5f<C>(phi(I, pj, b - 2 * s), phi(I, pj, b - s), phi(I, pj, b), phi(I, pj, b + s),
                         phi(I, pj, b + 2 * s));
    }
    return (C(phi(I, pj, K)) + C(phi(I, pj, K + 1))) / C(2);
}

// The advection kernel: a block of 64 x ADV_TY threads owns a tile of 64 x
// ADV_TY p columns and marches it up a chunk of levels.  phi2 is staged
// plane by plane into an LDS ring of four planes (levels k .. k+3 while level
// k is computed), each with the three-cell halo the WENO5 stencils need in x
// and y (ADV_W x ADV_H values); the next plane is loaded into registers one
// level ahead.  Each thread keeps the six values of its own column that the
// z stencils need (levels k-2 .. k+3) in registers, taken from the ring as
// planes arrive.  Per level the block reconstructs each x face of its rows
// (65 per row: the east face of every column and the west face of the
// first), each y face (ADV_TY + 1 per column) and each column's top face
// once, multiplies them by their face velocity, and forms conv2_phi from the
// flux differences (the bottom flux is the level below's top flux).  The
// velocities of the next level are loaded while this one is computed.  HBM
// traffic per cell: phi2, u2, v2, w2 read once (the halo re-reads of
// neighbouring tiles hit L2), conv2_phi written once.
constexpr int ADV_TY = 8;
constexpr int ADV_NT = BX * ADV_TY;
constexpr int ADV_W = BX + 6, ADV_H = ADV_TY + 6;
constexpr int ADV_NXF = (BX + 1) * ADV_TY;                     // x faces per level
constexpr int ADV_NYF = BX * (ADV_TY + 1);                     // y faces per level
constexpr int ADV_NS = (ADV_W * ADV_H + ADV_NT - 1) / ADV_NT;  // staged values per thread
constexpr int ADV_KS = 32;                                     // levels per chunk when unsized

__device__ __forceinline__ int wrap_any(int j, int P) { return ((j - 1) % P + P) % P + 1; }

// Face value between LDS positions c and c + st (a row or a column of a
// staged plane): WENO5 from the upwind side (weno.f90:45-76, 104-133).
template <typename C>
__device__ __forceinline__ C weno_lds(const C* a, int c, int st, bool up)
{
    const int b = up ? c : c + st, s = up ? st : -st;
    return weno5f<C>(a[b - 2 * s], a[b - s], a[b], a[b + s], a[b + 2 * s]);
}

// The z face between levels K and K+1 from a column's values z[m] = phi(K-2+m),
// m = 0..5 (weno.f90:135-167).
template <typename C>
__device__ __forceinline__ C zface(const C (&z)[6], int K, int n, C ww)
{
    if (K >= 3 && K <= n - 3) {
        const bool up = ww >= C(0);
        return weno5f<C>(up ? z[0] : z[5], up ? z[1] : z[4], up ? z[2] : z[3], up ? z[3] : z[2], up ? z[4] : z[1]);
    }
    return (z[2] + z[3]) / C(2);
}

template <typename T>
__global__ void __launch_bounds__(ADV_NT) k_advect_phi(Grid g, int nbx, int nby, int ks, T dx, T dy, T dz,
                                                       FieldView<T> phi, FieldView<T> u, FieldView<T> v,
                                                       FieldView<T> w, FieldView<T> conv)
{
    using C = T;
    __shared__ C ring[4][ADV_H * ADV_W];
    __shared__ C fx_s[ADV_TY][BX + 1];  // u * dfx on the x faces of the tile's rows
    __shared__ C fy_s[ADV_TY + 1][BX];  // v * dfy on the y faces of its columns
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % nbx, rest = t / nbx, by = rest % nby, bz = rest / nby;
    const int tid = threadIdx.x, tx = tid % BX, ty = tid / BX;
    const int L = g.maxl, n = g.maxn, P = g.P;
    const int i0 = 1 + bx * BX, j0 = 1 + by * ADV_TY;
    const int kb = 1 + bz * ks, ke = min(kb + ks - 1, n - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i <= L - 1 && j <= P;
    const int ic = min(i, L - 1), jm = wrap_any(j - 1, P);
    const long sz = phi.sz;
    const C rdx = C(1) / dx, rdy = C(1) / dy, rdz = C(1) / dz;
    // element offsets (level 0) of this thread's share of the staged window:
    // x from i0-3 clamped to the array, y rows j0-3 .. periodic
    long soff[ADV_NS];
#pragma unroll
    for (int m = 0; m < ADV_NS; ++m) {
        const int q = min(tid + m * ADV_NT, ADV_W * ADV_H - 1);
        const int li = q % ADV_W, lj = q / ADV_W;
        soff[m] = min(max(i0 - 3 + li, 0), L) + wrap_any(j0 - 3 + lj, P) * phi.sy;
    }
    auto fetch = [&](int k, C (&r)[ADV_NS]) {
        const long kz = long(min(max(k, 0), n)) * sz;
#pragma unroll
        for (int m = 0; m < ADV_NS; ++m) r[m] = phi.p[soff[m] + kz];
    };
    auto put = [&](int k, const C (&r)[ADV_NS]) {
        C* a = ring[k & 3];
#pragma unroll
        for (int m = 0; m < ADV_NS; ++m) {
            const int q = tid + m * ADV_NT;
            if (q < ADV_W * ADV_H) a[q] = r[m];
        }
    };
    // this thread's faces: x face tid and y face tid, and the faces beyond
    // ADV_NT: the 8 extra x faces go to the first lanes of the seventh
    // wavefront and the 64 extra y faces to the eighth, so that the two SIMDs
    // holding them evaluate 7 face reconstructions per level for the block
    // and the other two 6, instead of one wavefront doing 5 of its own.  LDS
    // positions and velocity offsets (advect_phi.f90:330-334: the x face I of
    // p row jr carries u(I, jr-1, k-1), the y face between p rows jr and jr+1
    // carries v(I, jr, k-1), the top face w(i, j-1, k))
    constexpr int XW = 6 * BX, YW = 7 * BX;  // first thread of the extra x / y faces
    static_assert(ADV_NXF - ADV_NT <= BX && ADV_NYF - ADV_NT <= BX && YW + BX == ADV_NT, "extra faces");
    const bool x2 = tid >= XW && tid < XW + (ADV_NXF - ADV_NT), y2 = tid >= YW && tid < YW + (ADV_NYF - ADV_NT);
    const int xf[2] = {tid, x2 ? ADV_NT + tid - XW : ADV_NXF - 1};
    const int yf[2] = {tid, y2 ? ADV_NT + tid - YW : ADV_NYF - 1};
    int xc[2], yc[2];
    bool xw[2];
    long uoff[2], voff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int fx = xf[h], fy = yf[h];
        const int Ix = i0 - 1 + fx % (BX + 1), rx = fx / (BX + 1);
        const int Iy = i0 + fy % BX, ry = fy / BX;
        xc[h] = (rx + 3) * ADV_W + (Ix - i0 + 3);
        xw[h] = Ix >= 3 && Ix <= L - 3;
        uoff[h] = min(Ix, L) + wrap_any(j0 + rx - 1, P) * u.sy;
        yc[h] = (ry + 2) * ADV_W + (Iy - i0 + 3);
        voff[h] = min(Iy, L) + wrap_any(j0 - 1 + ry, P) * v.sy;
    }
    const long woff = ic + jm * w.sy;
    C ux[2], vy[2], wt, ux_n[2] = {C(0), C(0)}, vy_n[2] = {C(0), C(0)}, wt_n = C(0);
    auto load_vel = [&](int k, C (&uu)[2], C (&vv)[2], C& ww) {
        const long kz = long(k - 1) * u.sz;
        uu[0] = u.p[uoff[0] + kz];
        uu[1] = u.p[uoff[1] + kz];
        vv[0] = v.p[voff[0] + long(k - 1) * v.sz];
        vv[1] = v.p[voff[1] + long(k - 1) * v.sz];
        ww = w.p[woff + long(k) * w.sz];
    };
    // prologue: planes kb .. kb+2 into the ring, the column window phi(kb-3 ..
    // kb+2) and the bottom flux of level kb (the face K = kb-1)
    C z[6];
    {
        C r[ADV_NS];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            fetch(kb + q, r);
            put(kb + q, r);
        }
        const long col = ic + wrap_any(j, P) * phi.sy;
#pragma unroll
        for (int m = 0; m < 6; ++m) z[m] = phi.p[col + long(min(max(kb - 3 + m, 0), n)) * sz];
    }
    const C wb = w.p[woff + long(kb - 1) * w.sz];
    C hb = wb * zface<C>(z, kb - 1, n, wb);
    C nxt[ADV_NS];
    fetch(kb + 3, nxt);
    load_vel(kb, ux, vy, wt);
    for (int k = kb; k <= ke; ++k) {
        // the ring slot of level k+3 held level k-1, last read before the
        // previous level's second barrier
        put(k + 3, nxt);
        if (k < ke) {
            fetch(k + 4, nxt);
            load_vel(k + 1, ux_n, vy_n, wt_n);
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 5; ++m) z[m] = z[m + 1];
        z[5] = ring[(k + 3) & 3][(ty + 3) * ADV_W + tx + 3];
        const C* pl = ring[k & 3];
        {
            const bool up = ux[0] >= C(0);
            const C f = xw[0] ? weno_lds<C>(pl, xc[0], 1, up) : pl[up ? xc[0] : xc[0] + 1];
            (&fx_s[0][0])[tid] = ux[0] * f;
        }
        if (x2) {
            const bool up = ux[1] >= C(0);
            const C f = xw[1] ? weno_lds<C>(pl, xc[1], 1, up) : pl[up ? xc[1] : xc[1] + 1];
            (&fx_s[0][0])[xf[1]] = ux[1] * f;
        }
        (&fy_s[0][0])[tid] = vy[0] * weno_lds<C>(pl, yc[0], ADV_W, vy[0] >= C(0));
        if (y2) (&fy_s[0][0])[yf[1]] = vy[1] * weno_lds<C>(pl, yc[1], ADV_W, vy[1] >= C(0));
        const C ht = wt * zface<C>(z, k, n, wt);
        __syncthreads();
        if (mine) {
            const C ax = fx_s[ty][tx + 1] - fx_s[ty][tx], ay = fy_s[ty + 1][tx] - fy_s[ty][tx], az = ht - hb;
            if constexpr (sizeof(C) == 4)
                conv(i, j, k) = ax * rdx + ay * rdy + az * rdz;  // FP32: reciprocal multiplies
            else
                conv(i, j, k) = div_nr(ax, dx, rdx) + div_nr(ay, dy, rdy) + div_nr(az, dz, rdz);
        }
        hb = ht;
        ux[0] = ux_n[0];
        ux[1] = ux_n[1];
        vy[0] = vy_n[0];
        vy[1] = vy_n[1];
        wt = wt_n;
    }
}

// WENOAppoxConvection as a standalone routine: dfx on I in 0..maxl-1, dfy on
// I in 1..maxl-1 (both K in 0..maxn-2), dfz on I in 1..maxl-1, K in 0..maxn-1.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_weno_faces(Grid g, FieldView<T> phi, FieldView<T> u,
                                                       FieldView<T> v, FieldView<T> w, FieldView<T> dfx,
                                                       FieldView<T> dfy, FieldView<T> dfz)
{
    using C = T;
    const int I = blockIdx.x * BX + threadIdx.x;
    const int J = blockIdx.y * BY + threadIdx.y + 1;
    const int K = blockIdx.z;
    if (I > g.maxl - 1 || J > g.P) return;
    const int pj = wrapj(J + 1, g.P);
    if (K <= g.maxn - 2) {
        dfx(I, J, K) = face_x<C>(phi, I, pj, K + 1, C(u(I, J, K)), g.maxl);
        if (I >= 1) dfy(I, J, K) = face_y<C>(phi, I, J, K + 1, C(v(I, J, K)), g.P);
    }
    if (I >= 1) dfz(I, J, K) = face_z<C>(phi, I, pj, K, C(w(I, J, K)), g.maxn);
}

template <typename T>
hipError_t launch_advect_phi(hipStream_t s, const Grid& g, const Phys& ph, FieldView<T> phi2,
                             FieldView<T> u2, FieldView<T> v2, FieldView<T> w2, FieldView<T> conv2_phi)
{
    const int nbx = (g.maxl - 1 + BX - 1) / BX, nby = (g.P + ADV_TY - 1) / ADV_TY, nz = g.maxn - 1;
    static const int slots = resident_blocks(k_advect_phi<T>, ADV_NT);
    const int ks = march_chunk(nbx * nby, nz, slots, ADV_KS), nbz = (nz + ks - 1) / ks;
    hipLaunchKernelGGL(k_advect_phi<T>, dim3(nbx * nby * nbz), dim3(ADV_NT), 0, s, g, nbx, nby, ks, T(ph.dx),
                       T(ph.dy), T(ph.dz), phi2, u2, v2, w2, conv2_phi);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_weno_faces(hipStream_t s, const Grid& g, FieldView<T> phi, FieldView<T> u,
                             FieldView<T> v, FieldView<T> w, FieldView<T> dfx, FieldView<T> dfy,
                             FieldView<T> dfz)
{
    dim3 blk(BX, BY, 1);
    dim3 grd((g.maxl + BX - 1) / BX, (g.P + BY - 1) / BY, g.maxn);
    hipLaunchKernelGGL(k_weno_faces<T>, grd, blk, 0, s, g, phi, u, v, w, dfx, dfy, dfz);
    return hipGetLastError();
}

template hipError_t launch_advect_phi<double>(hipStream_t, const Grid&, const Phys&, FieldView<double>,
                                              FieldView<double>, FieldView<double>, FieldView<double>,
                                              FieldView<double>);
template hipError_t launch_advect_phi<float>(hipStream_t, const Grid&, const Phys&, FieldView<float>,
                                             FieldView<float>, FieldView<float>, FieldView<float>,
                                             FieldView<float>);
template hipError_t launch_weno_faces<double>(hipStream_t, const Grid&, FieldView<double>, FieldView<double>,
                                              FieldView<double>, FieldView<double>, FieldView<double>,
                                              FieldView<double>, FieldView<double>);

}  // namespace tpls
