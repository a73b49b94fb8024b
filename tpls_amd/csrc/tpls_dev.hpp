// Device-side building blocks of the TPLS stencil core on gfx950.
//
// Every field keeps the reference's Fortran shape and index bounds
// (main_ns_hybrid.F90:416-549), x fastest, with the x extent padded to a
// multiple of 16 elements so that each (j,k) row starts on a 128-byte line.
// A FieldView addresses element (i,j,k) of the Fortran array directly.
//
// The y direction is periodic with period P = maxm-1 over rows 1..P.  The
// reference keeps halo rows and refreshes them with MPI exchanges
// (mpi.f90:110-662) before every read; every read site on the hot path is
// preceded by such an exchange, so reading row wrapj(j) instead of halo row j
// gives identical values.  Kernels therefore never read halo rows.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tpls {

template <typename T>
struct FieldView {
    T* p;         // element (0,0,0) of the logical Fortran index space
    long sy, sz;  // strides in elements
    __host__ __device__ T& operator()(int i, int j, int k) const
    {
        return p[i + j * sy + k * sz];
    }
};

// Grid geometry shared by every kernel of a case.
struct Grid {
    int maxl, maxm, maxn;  // reference extents
    int P;                 // y period, maxm-1
};

__device__ __forceinline__ int wrapj(int j, int P)
{
    return j < 1 ? j + P : (j > P ? j - P : j);
}

// Single-precision literals of the reference widened to double:
// weno.f90:184 (epslon = 1.e-7) and the pi of levelset.f90:268,
// momentum_allflux.f90:489, main_ns_hybrid.F90:181.
#define TPLS_EPSLON ((double)1.e-7f)
#define TPLS_PI_SP ((double)3.1415926535f)

// Jiang-Shu WENO5 face value, weno.f90:175-215, in the reference's
// operation order.
template <typename C>
__device__ __forceinline__ C weno5(C v1, C v2, C v3, C v4, C v5)
{
    const C eps = (C)TPLS_EPSLON;
    C t1 = v1 - C(2) * v2 + v3;
    C t2 = v1 - C(4) * v2 + C(3) * v3;
    const C c1312 = C(13.0 / 12.0);
    C s1 = c1312 * t1 * t1 + C(0.25) * t2 * t2;
    t1 = v2 - C(2) * v3 + v4;
    t2 = v2 - v4;
    C s2 = c1312 * t1 * t1 + C(0.25) * t2 * t2;
    t1 = v3 - C(2) * v4 + v5;
    t2 = C(3) * v3 - C(4) * v4 + v5;
    C s3 = c1312 * t1 * t1 + C(0.25) * t2 * t2;
    C q1 = eps + s1, q2 = eps + s2, q3 = eps + s3;
    C a1 = C(0.1) / (q1 * q1);
    C a2 = C(0.6) / (q2 * q2);
    C a3 = C(0.3) / (q3 * q3);
    C sum = a1 + a2 + a3;
    C w1 = a1 / sum, w2 = a2 / sum, w3 = a3 / sum;
    t1 = v1 / C(3) - C(7) * v2 / C(6) + C(11) * v3 / C(6);
    t2 = -(v2 / C(6)) + C(5) * v3 / C(6) + v4 / C(3);
    C t3 = v3 / C(3) + C(5) * v4 / C(6) - v5 / C(6);
    return w1 * t1 + w2 * t2 + w3 * t3;
}

// Smoothed-Heaviside viscosity, momentum_allflux.f90:499-512 (the same
// expression main_ns_hybrid.F90:667-680 uses at the inlet/outlet planes).
template <typename C>
__device__ __forceinline__ C visc_of_phi(C pv, C sw, C Re, C mu_minus, C mu_plus)
{
    C H;
    if (pv < -sw)
        H = C(0);
    else if (pv > sw)
        H = C(1);
    else
        H = C(0.5) + (pv / (C(2) * sw)) + (C(1) / (C(2) * C(TPLS_PI_SP))) * sin(C(TPLS_PI_SP) * pv / sw);
    C temp = mu_plus * H + mu_minus * (C(1) - H);
    return temp / Re;
}

// Face viscosities of the staggered momentum grids: the averaging patterns of
// jacobi_iteration_allflux.f90:75-82 (u), 153-160 (v), 231-238 (w); the SOR
// sweeps and get_diffusion_all use the same ones.  mu is read with y wrap.
template <typename C>
struct MuFaces {
    C xp, xm, yp, ym, zp, zm;
};

template <typename C, typename T>
__device__ __forceinline__ MuFaces<C> mu_faces_u(const T& mu, int i, int j, int k, int P)
{
    const int j1 = wrapj(j + 1, P), j2 = wrapj(j + 2, P), j0 = wrapj(j, P);
    MuFaces<C> m;
    C a = mu(i, j1, k + 1), b = mu(i + 1, j1, k + 1);
    m.xp = b;
    m.xm = a;
    m.yp = (a + b + C(mu(i, j2, k + 1)) + C(mu(i + 1, j2, k + 1))) * C(0.25);
    m.ym = (a + b + C(mu(i, j0, k + 1)) + C(mu(i + 1, j0, k + 1))) * C(0.25);
    m.zp = (a + b + C(mu(i, j1, k + 2)) + C(mu(i + 1, j1, k + 2))) * C(0.25);
    m.zm = (a + b + C(mu(i, j1, k)) + C(mu(i + 1, j1, k))) * C(0.25);
    return m;
}

template <typename C, typename T>
__device__ __forceinline__ MuFaces<C> mu_faces_v(const T& mu, int i, int j, int k, int P)
{
    const int j1 = wrapj(j + 1, P), j0 = wrapj(j, P);
    MuFaces<C> m;
    C a = mu(i, j1, k + 1), b = mu(i, j0, k + 1);
    m.xp = (a + b + C(mu(i + 1, j1, k + 1)) + C(mu(i + 1, j0, k + 1))) * C(0.25);
    m.xm = (a + b + C(mu(i - 1, j1, k + 1)) + C(mu(i - 1, j0, k + 1))) * C(0.25);
    m.yp = a;
    m.ym = b;
    m.zp = (a + b + C(mu(i, j1, k + 2)) + C(mu(i, j0, k + 2))) * C(0.25);
    m.zm = (a + b + C(mu(i, j1, k)) + C(mu(i, j0, k))) * C(0.25);
    return m;
}

template <typename C, typename T>
__device__ __forceinline__ MuFaces<C> mu_faces_w(const T& mu, int i, int j, int k, int P)
{
    const int j1 = wrapj(j + 1, P), j2 = wrapj(j + 2, P), j0 = wrapj(j, P);
    MuFaces<C> m;
    C a = mu(i, j1, k), b = mu(i, j1, k + 1);
    m.xp = (a + b + C(mu(i + 1, j1, k)) + C(mu(i + 1, j1, k + 1))) * C(0.25);
    m.xm = (a + b + C(mu(i - 1, j1, k)) + C(mu(i - 1, j1, k + 1))) * C(0.25);
    m.yp = (a + b + C(mu(i, j2, k)) + C(mu(i, j2, k + 1))) * C(0.25);
    m.ym = (a + b + C(mu(i, j0, k)) + C(mu(i, j0, k + 1))) * C(0.25);
    m.zp = b;
    m.zm = a;
    return m;
}

// 1/d and a/d.  FP64: the correctly rounded division, the reference's
// operation.  FP32 (the storage-precision variant, held to tolerances against
// FP64): the hardware reciprocal, within an ulp, which takes one instruction
// where the correctly rounded FP32 division takes about ten.
template <typename C>
__device__ __forceinline__ C recip(C d)
{
    return C(1) / d;
}
template <>
__device__ __forceinline__ float recip<float>(float d)
{
    return __builtin_amdgcn_rcpf(d);
}
template <typename C>
__device__ __forceinline__ C sdiv(C a, C d)
{
    return a / d;
}
template <>
__device__ __forceinline__ float sdiv<float>(float a, float d)
{
    return a * __builtin_amdgcn_rcpf(d);
}

// (1/diag) * (...) of the Crank-Nicolson Helmholtz update,
// jacobi_iteration_allflux.f90:84-90 (neighbour order x+, x-, y+, y-, z+, z-).
// In two parts, so that a kernel that updates one cell twice with the same
// viscosity faces can form 1/diag once: helmholtz_rdiag, then helmholtz_rd
// with it; helmholtz is the two together.
template <typename C>
__device__ __forceinline__ C helmholtz_rdiag(const MuFaces<C>& m, C ax2, C ay2, C az2)
{
    C diag = C(1) + ax2 * (m.xp + m.xm) + ay2 * (m.yp + m.ym) + az2 * (m.zp + m.zm);
    return recip(diag);
}
template <typename C>
__device__ __forceinline__ C helmholtz_rd(C rd, const MuFaces<C>& m, C ax2, C ay2, C az2, C xp, C xm, C yp, C ym,
                                          C zp, C zm, C rhs)
{
    return rd * (ax2 * (m.xp * xp + m.xm * xm) + ay2 * (m.yp * yp + m.ym * ym) + az2 * (m.zp * zp + m.zm * zm) + rhs);
}
template <typename C>
__device__ __forceinline__ C helmholtz(const MuFaces<C>& m, C ax2, C ay2, C az2, C xp, C xm, C yp,
                                       C ym, C zp, C zm, C rhs)
{
    return helmholtz_rd(helmholtz_rdiag(m, ax2, ay2, az2), m, ax2, ay2, az2, xp, xm, yp, ym, zp, zm, rhs);
}

}  // namespace tpls
