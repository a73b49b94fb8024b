// WENO5 level-set advection: weno.f90:18-215 and advect_phi.f90:19-119.
//
// The fused kernel computes conv2_phi = div(u * phi_face) on every p cell in
// one pass: the face values a cell needs are reconstructed in registers from
// phi2 (read with periodic y wrap, which is what the reference's phi_augaug
// plus exchange2d_augaug1/2/3 provide) and shared between the neighbouring
// cells of a block, instead of being written to and re-read from dfx/dfy/dfz
// arrays.  HBM traffic per cell: phi2, u2, v2,
// w2 read once, conv2_phi written once (40 B in FP64).
#include "tpls_kernels.hpp"

namespace tpls {

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
    return (n1 * e1 + n2 * e2 + n3 * e3) / (n1 + n2 + n3);
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

// Levels per thread of the advection kernel: the top face of one level is the
// bottom face of the next.
constexpr int ADV_KZ = 4;

// conv2_phi of ADV_KZ cells stacked in k.  Each face is reconstructed once per
// thread column: the west face is the east face of the lane to the left
// (shuffle), the south face the north face of the row below (LDS), the bottom
// face the previous level's top face; the first lane, the first row and the
// first level compute their own.  Lanes past the grid's edge work on clamped
// indices and store nothing, so every lane takes part in the exchanges.
template <typename T>
__global__ void __launch_bounds__(BX* BY) k_advect_phi(Grid g, T dx, T dy, T dz, FieldView<T> phi,
                                                       FieldView<T> u, FieldView<T> v, FieldView<T> w,
                                                       FieldView<T> conv)
{
    using C = T;
    __shared__ C fn_s[BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = blockIdx.x * BX + tx + 1;
    const int j = blockIdx.y * BY + ty + 1;
    const int k0 = blockIdx.z * ADV_KZ + 1;
    const int ic = min(i, g.maxl - 1), jc = min(j, g.P);
    const bool col = i <= g.maxl - 1 && j <= g.P;
    const int jm = wrapj(jc - 1, g.P);
    C ft = C(0), wt = C(0);
#pragma unroll
    for (int kk = 0; kk < ADV_KZ; ++kk) {
        const int k = min(k0 + kk, g.maxn - 1);
        // advect_phi.f90:330-334
        const C ue = u(ic, jm, k - 1);
        const C vn = v(ic, jc, k - 1), vs = v(ic, jm, k - 1);
        const C wb = kk ? wt : C(w(ic, jm, k - 1));
        wt = w(ic, jm, k);
        const C fe = face_x<C>(phi, ic, jc, k, ue, g.maxl);
        C uw = __shfl_up(ue, 1, 64), fw = __shfl_up(fe, 1, 64);
        if (tx == 0) {
            uw = u(ic - 1, jm, k - 1);
            fw = face_x<C>(phi, ic - 1, jc, k, uw, g.maxl);
        }
        const C fn = face_y<C>(phi, ic, jc, k, vn, g.P);
        fn_s[ty][tx] = fn;
        __syncthreads();
        const C fs = ty ? fn_s[ty - 1][tx] : face_y<C>(phi, ic, jc - 1, k, vs, g.P);
        __syncthreads();
        const C fb = kk ? ft : face_z<C>(phi, ic, jc, k - 1, wb, g.maxn);
        ft = face_z<C>(phi, ic, jc, k, wt, g.maxn);
        if (col && k0 + kk <= g.maxn - 1)
            conv(i, j, k) = (ue * fe - uw * fw) / dx + (vn * fn - vs * fs) / dy + (wt * ft - wb * fb) / dz;
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
    dim3 blk(BX, BY, 1);
    dim3 grd((g.maxl - 1 + BX - 1) / BX, (g.P + BY - 1) / BY, (g.maxn - 1 + ADV_KZ - 1) / ADV_KZ);
    hipLaunchKernelGGL(k_advect_phi<T>, grd, blk, 0, s, g, T(ph.dx), T(ph.dy), T(ph.dz), phi2, u2, v2, w2,
                       conv2_phi);
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
