! Fortran host side of the MI355X-native TPLS stencil core.
!
! These modules carry the reference's module and procedure names and dummy
! argument lists (weno.f90:18, advect_phi.f90:19, two_phase_levelset.f90:20,
! levelset.f90:17/97/241, cahn_hilliard_solver.f90:41, momentum_allflux.f90:22/224/474,
! jacobi_iteration_allflux.f90:21/102/195/285, sor_iteration_allflux.f90:19/
! 152/280, pressure.f90:16/59), so main_ns_hybrid.F90 compiles against them
! unchanged and its calls run the gfx950 kernels of libtpls_hip.so through
! the C ABI of include/tpls_core.h (module tpls_core_c below).
!
! The MPI arguments of do_advect_phi, do_tpls and dim (neighbours, communicator,
! stride datatypes) are accepted and ignored: the whole domain lives on one
! GPU and the halo exchanges are in-kernel periodic wraps.  A non-zero status
! from the C ABI stops the program with the library's message, as the
! reference's abort() does for a failing MPI or PETSc call.
!
! The pressure solve main_ns_hybrid.F90:896-913 goes through PETSc objects
! (KSPSolve and the callbacks of pressure_solver.F90).  Module tpls_state
! holds the solver-global data of state.F90 that those callbacks read, and
! module pressure_solver offers the calls main makes on its KSP object
! (KSPSolve, KSPGetResidualNorm, KSPGetIterationNumber,
! KSPGetConvergedReason, output_converged_reason, copy_pressure_out_of_petsc)
! with main's argument lists, so that block of main compiles as it stands;
! module tpls_pressure offers the whole sequence as one call.
!
! Module tpls_mpi holds the halo exchanges main calls between procedures
! (exchange2d, exchange2d_aug1, exchange2d_aug2 of mpi.f90) for the single
! rank, and module tpls_mirror attaches main's arrays to device copies that
! stay resident between calls (include/tpls_core.h, tpls_mirror_attach).

module tpls_core_c
  use iso_c_binding
  implicit none

  interface
    function tpls_last_error() bind(C, name="tpls_last_error")
      import :: c_ptr
      type(c_ptr) :: tpls_last_error
    end function

    integer(c_int) function c_weno(phi, u, v, w, sx, ex, sy, ey, l, n, dfx, dfy, dfz) &
        bind(C, name="tpls_wenoappoxconvection")
      import :: c_int, c_double
      real(c_double), intent(in) :: phi(*), u(*), v(*), w(*)
      integer(c_int), value :: sx, ex, sy, ey, l, n
      real(c_double), intent(inout) :: dfx(*), dfy(*), dfz(*)
    end function

    integer(c_int) function c_advect_phi(phi2, u2, v2, w2, conv2_phi, sx, ex, sy, ey, maxl, maxn, &
        ex_max, dx, dy, dz) bind(C, name="tpls_do_advect_phi")
      import :: c_int, c_double
      real(c_double), intent(inout) :: phi2(*), u2(*), v2(*), w2(*), conv2_phi(*)
      integer(c_int), value :: sx, ex, sy, ey, maxl, maxn, ex_max
      real(c_double), value :: dx, dy, dz
    end function

    integer(c_int) function c_tpls(phi2, phi3, c0, c1, c2, sx, ex, sy, ey, maxn, ex_max, its, &
        height, err2, dx, dz, dt) bind(C, name="tpls_do_tpls")
      import :: c_int, c_double
      real(c_double), intent(inout) :: phi2(*), phi3(*), c0(*), c1(*), c2(*)
      integer(c_int), value :: sx, ex, sy, ey, maxn, ex_max, its
      real(c_double), value :: height, dx, dz, dt
      real(c_double), intent(out) :: err2
    end function

    integer(c_int) function c_status(phi, d, ist, sg, sx, ex, sy, ey, n, h) &
        bind(C, name="tpls_get_levelset_status")
      import :: c_int, c_double
      real(c_double), intent(in) :: phi(*)
      real(c_double), intent(inout) :: d(*), sg(*)
      integer(c_int), intent(inout) :: ist(*)
      integer(c_int), value :: sx, ex, sy, ey, n
      real(c_double), value :: h
    end function

    integer(c_int) function c_reinit(phi, pr, d, ist, sg, sx, ex, sy, ey, n, ex_max, h, height, err) &
        bind(C, name="tpls_do_levelset_iteration")
      import :: c_int, c_double
      real(c_double), intent(in) :: phi(*), d(*), sg(*)
      real(c_double), intent(inout) :: pr(*)
      integer(c_int), intent(in) :: ist(*)
      integer(c_int), value :: sx, ex, sy, ey, n, ex_max
      real(c_double), value :: h, height
      real(c_double), intent(out) :: err
    end function

    integer(c_int) function c_csf(fx, fy, fz, fi, sx, ex, sy, ey, n, ex_max, dx, dy, dz, scap, sw) &
        bind(C, name="tpls_get_csf")
      import :: c_int, c_double
      real(c_double), intent(inout) :: fx(*), fy(*), fz(*)
      real(c_double), intent(in) :: fi(*)
      integer(c_int), value :: sx, ex, sy, ey, n, ex_max
      real(c_double), value :: dx, dy, dz, scap, sw
    end function

    integer(c_int) function c_diffusion(du, dv, dw, mu, u2, v2, w2, sx, sy, ex, ey, maxn, ex_max, &
        dx, dy, dz) bind(C, name="tpls_get_diffusion_all")
      import :: c_int, c_double
      real(c_double), intent(inout) :: du(*), dv(*), dw(*)
      real(c_double), intent(in) :: mu(*), u2(*), v2(*), w2(*)
      integer(c_int), value :: sx, sy, ex, ey, maxn, ex_max
      real(c_double), value :: dx, dy, dz
    end function

    integer(c_int) function c_conv(cu, cv, cw, mu, u2, v2, w2, sx, sy, ex, ey, maxn, ex_max, &
        dx, dy, dz) bind(C, name="tpls_get_conv_all")
      import :: c_int, c_double
      real(c_double), intent(inout) :: cu(*), cv(*), cw(*)
      real(c_double), intent(in) :: mu(*), u2(*), v2(*), w2(*)
      integer(c_int), value :: sx, sy, ex, ey, maxn, ex_max
      real(c_double), value :: dx, dy, dz
    end function

    integer(c_int) function c_viscosity(mu, phi, ex, ey, sx, sy, maxn, sw, re, mu_minus, mu_plus) &
        bind(C, name="tpls_get_viscosity")
      import :: c_int, c_double
      real(c_double), intent(inout) :: mu(*)
      real(c_double), intent(in) :: phi(*)
      integer(c_int), value :: ex, ey, sx, sy, maxn
      real(c_double), value :: sw, re, mu_minus, mu_plus
    end function

    integer(c_int) function c_jacobi_u(x, xo, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max) &
        bind(C, name="tpls_do_jacobi_u")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: xo(*), rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, ex_max
    end function
    integer(c_int) function c_jacobi_v(x, xo, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max) &
        bind(C, name="tpls_do_jacobi_v")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: xo(*), rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, ex_max
    end function
    integer(c_int) function c_jacobi_w(x, xo, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max) &
        bind(C, name="tpls_do_jacobi_w")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: xo(*), rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, ex_max
    end function

    integer(c_int) function c_difference(ua, ub, ex, ey, ez, sx, sy, sz, diff) &
        bind(C, name="tpls_get_difference")
      import :: c_int, c_double
      real(c_double), intent(in) :: ua(*), ub(*)
      integer(c_int), value :: ex, ey, ez, sx, sy, sz
      real(c_double), intent(out) :: diff
    end function

    integer(c_int) function c_sor_u(x, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max, flg, it) &
        bind(C, name="tpls_do_sor_u")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, ex_max, flg, it
    end function
    integer(c_int) function c_sor_v(x, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, it) &
        bind(C, name="tpls_do_sor_v")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, flg, it
    end function
    integer(c_int) function c_sor_w(x, rhs, mu, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, it) &
        bind(C, name="tpls_do_sor_w")
      import :: c_int, c_double
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: rhs(*), mu(*)
      real(c_double), value :: dx, dy, dz, dt
      integer(c_int), value :: ex, ey, sx, sy, maxn, flg, it
    end function

    integer(c_int) function c_source_pres(rhs, u3, v3, w3, ex, ey, sx, sy, maxn, dx, dy, dz, dt) &
        bind(C, name="tpls_get_source_pres")
      import :: c_int, c_double
      real(c_double), intent(inout) :: rhs(*)
      real(c_double), intent(in) :: u3(*), v3(*), w3(*)
      integer(c_int), value :: ex, ey, sx, sy, maxn
      real(c_double), value :: dx, dy, dz, dt
    end function

    integer(c_int) function c_uvw_pres(u3, v3, w3, p, ex, ey, sx, sy, maxn, ex_max, dx, dy, dz, dt, &
        u_inlet) bind(C, name="tpls_get_uvw_pres")
      import :: c_int, c_double
      real(c_double), intent(inout) :: u3(*), v3(*), w3(*)
      real(c_double), intent(in) :: p(*), u_inlet(*)
      integer(c_int), value :: ex, ey, sx, sy, maxn, ex_max
      real(c_double), value :: dx, dy, dz, dt
    end function

    integer(c_int) function c_pressure_solve(pres, rhs_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, &
        dt, rtol, abstol, dtol, maxits, its, rnorm, reason) bind(C, name="tpls_pressure_solve")
      import :: c_int, c_double
      real(c_double), intent(inout) :: pres(*)
      real(c_double), intent(in) :: rhs_p(*), u3(*), u_inlet(*)
      integer(c_int), value :: maxl, maxm, maxn, maxits
      real(c_double), value :: dx, dy, dz, dt, rtol, abstol, dtol
      integer(c_int), intent(out) :: its, reason
      real(c_double), intent(out) :: rnorm
    end function

    integer(c_int) function c_exchange2d(uu, ex, ey, ez, sx, sy, sz, kind) bind(C, name="tpls_exchange2d")
      import :: c_int, c_double
      real(c_double), intent(inout) :: uu(*)
      integer(c_int), value :: ex, ey, ez, sx, sy, sz, kind
    end function

    integer(c_int) function c_dim(maxl, maxn, dx, dy, dz, dt, epn, pe, phi, phi_old, conv2_phi, &
        conv_prev, phi_new, sx, ex, sy, ey, height, its, res) bind(C, name="tpls_dim")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: maxl, maxn, sx, ex, sy, ey, its
      real(c_double), value :: dx, dy, dz, dt, epn, pe, height
      real(c_double), intent(in) :: phi(*), phi_old(*), conv2_phi(*)
      real(c_double), intent(inout) :: conv_prev(*), phi_new(*)
      type(c_ptr), value :: res
    end function
  end interface

contains

  ! Stop with the library's message when a call fails.
  subroutine tpls_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    if (rc == 0) return
    p = tpls_last_error()
    n = 0
    if (c_associated(p)) then
      call c_f_pointer(p, s, [4096])
      do while (n < 4096)
        if (s(n + 1) == c_null_char) exit
        n = n + 1
      end do
      write(0, '(a, a, i0, a, 4096a1)') where, ': tpls status ', rc, ': ', s(1:n)
    else
      write(0, '(a, a, i0)') where, ': tpls status ', rc
    end if
    error stop 1
  end subroutine tpls_check
end module tpls_core_c


module weno
  use tpls_core_c
  implicit none
contains
  subroutine WENOAppoxConvection(phi, u, v, w, sx, ex, sy, ey, l, n, dfx, dfy, dfz)
    integer, intent(in) :: sx, ex, sy, ey, l, n
    double precision, intent(in) :: phi(sx-3:ex+3, sy-3:ey+3, 0:n)
    double precision, intent(in) :: u(sx-1:ex+1, sy-1:ey+1, 0:n-2)
    double precision, intent(in) :: v(sx-1:ex+1, sy-1:ey+1, 0:n-2)
    double precision, intent(in) :: w(sx-1:ex+1, sy-1:ey+1, 0:n-1)
    double precision, intent(out) :: dfx(sx-1:ex+1, sy-1:ey+1, 0:n-2)
    double precision, intent(out) :: dfy(sx-1:ex+1, sy-1:ey+1, 0:n-2)
    double precision, intent(out) :: dfz(sx-1:ex+1, sy-1:ey+1, 0:n-1)
    call tpls_check(c_weno(phi, u, v, w, sx, ex, sy, ey, l, n, dfx, dfy, dfz), 'WENOAppoxConvection')
  end subroutine WENOAppoxConvection
end module weno


module advect_phi
  use tpls_core_c
  implicit none
contains
  subroutine do_advect_phi(phi2, u2, v2, w2, conv2_phi, sx, ex, sy, ey, sz_p, ez_p, sz_uv, sz_w, &
      ez_uv, ez_w, maxl, maxn, ex_max, neighbours, comm2d_quasiperiodic, dx, dy, dz, &
      stride_p_augaug1_xz, stride_p_augaug1_yz, stride_p_augaug2_xz, stride_p_augaug2_yz, &
      stride_p_augaug3_xz, stride_p_augaug3_yz, stride_uv_xz, stride_uv_yz, stride_w_xz, stride_w_yz)
    integer, intent(in) :: sx, ex, sy, ey, sz_p, ez_p, sz_uv, sz_w, ez_uv, ez_w
    integer, intent(in) :: maxl, maxn, ex_max
    integer, intent(in) :: neighbours(6)
    integer, intent(in) :: comm2d_quasiperiodic
    integer, intent(in) :: stride_p_augaug1_xz, stride_p_augaug1_yz
    integer, intent(in) :: stride_p_augaug2_xz, stride_p_augaug2_yz
    integer, intent(in) :: stride_p_augaug3_xz, stride_p_augaug3_yz
    integer, intent(in) :: stride_uv_xz, stride_uv_yz, stride_w_xz, stride_w_yz
    double precision, intent(in) :: dx, dy, dz
    double precision, intent(inout) :: phi2(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: u2(sx-1:ex+1, sy-1:ey+1, sz_uv:ez_uv)
    double precision, intent(inout) :: v2(sx-1:ex+1, sy-1:ey+1, sz_uv:ez_uv)
    double precision, intent(inout) :: w2(sx-1:ex+1, sy-1:ey+1, sz_w:ez_w)
    double precision, intent(out) :: conv2_phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    if (sz_p /= 0 .or. ez_p /= maxn .or. sz_uv /= 0 .or. ez_uv /= maxn - 2 .or. sz_w /= 0 &
        .or. ez_w /= maxn - 1) call tpls_check(-2_c_int, 'do_advect_phi: z bounds')
    call tpls_check(c_advect_phi(phi2, u2, v2, w2, conv2_phi, sx, ex, sy, ey, maxl, maxn, ex_max, &
        dx, dy, dz), 'do_advect_phi')
  end subroutine do_advect_phi
end module advect_phi


module levelset
  use tpls_core_c
  implicit none
contains
  subroutine get_levelset_status(phi, D_mx, istatus, sign_mx, sx, ex, sy, ey, n, h)
    integer, intent(in) :: sx, ex, sy, ey, n
    double precision, intent(in) :: h
    double precision, intent(in) :: phi(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: sign_mx(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: D_mx(sx-1:ex+1, sy-1:ey+1, 0:n)
    integer, intent(out) :: istatus(sx-1:ex+1, sy-1:ey+1, 0:n)
    call tpls_check(c_status(phi, D_mx, istatus, sign_mx, sx, ex, sy, ey, n, h), 'get_levelset_status')
  end subroutine get_levelset_status

  subroutine do_levelset_iteration(phi, phi_reinit, D_mx, istatus, sign_mx, sx, ex, sy, ey, n, &
      ex_max, h, height, error_val)
    integer, intent(in) :: sx, ex, sy, ey, n, ex_max
    integer, intent(in) :: istatus(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(in) :: h, height
    double precision, intent(in) :: phi(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(in) :: sign_mx(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(in) :: D_mx(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(inout) :: phi_reinit(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: error_val
    call tpls_check(c_reinit(phi, phi_reinit, D_mx, istatus, sign_mx, sx, ex, sy, ey, n, ex_max, h, &
        height, error_val), 'do_levelset_iteration')
  end subroutine do_levelset_iteration

  subroutine get_csf(fx_csf, fy_csf, fz_csf, fi, sx, ex, sy, ey, n, ex_max, dx, dy, dz, scap, &
      smooth_width)
    integer, intent(in) :: sx, ex, sy, ey, n, ex_max
    double precision, intent(in) :: dx, dy, dz, scap, smooth_width
    double precision, intent(in) :: fi(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: fx_csf(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: fy_csf(sx-1:ex+1, sy-1:ey+1, 0:n)
    double precision, intent(out) :: fz_csf(sx-1:ex+1, sy-1:ey+1, 0:n)
    call tpls_check(c_csf(fx_csf, fy_csf, fz_csf, fi, sx, ex, sy, ey, n, ex_max, dx, dy, dz, scap, &
        smooth_width), 'get_csf')
  end subroutine get_csf
end module levelset


module two_phase_levelset
  use tpls_core_c
  implicit none
contains
  subroutine do_tpls(phi2, phi3, conv0_phi, conv1_phi, conv2_phi, sx, ex, sy, ey, sz_p, ez_p, maxn, &
      ex_max, stride_p_xz, stride_p_yz, max_iteration_levelset, neighbours, comm2d_quasiperiodic, &
      height, err2, dx, dz, dt)
    integer, intent(in) :: sx, ex, sy, ey, sz_p, ez_p
    integer, intent(in) :: maxn, ex_max
    integer, intent(in) :: stride_p_xz, stride_p_yz
    integer, intent(in) :: max_iteration_levelset
    integer, intent(in) :: neighbours(6)
    integer, intent(in) :: comm2d_quasiperiodic
    double precision, intent(in) :: dx, dz, dt
    double precision, intent(in) :: height
    double precision, intent(inout) :: err2
    double precision, intent(inout) :: phi2(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: conv0_phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: conv1_phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: conv2_phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: phi3(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    if (sz_p /= 0 .or. ez_p /= maxn) call tpls_check(-2_c_int, 'do_tpls: z bounds')
    call tpls_check(c_tpls(phi2, phi3, conv0_phi, conv1_phi, conv2_phi, sx, ex, sy, ey, maxn, ex_max, &
        max_iteration_levelset, height, err2, dx, dz, dt), 'do_tpls')
  end subroutine do_tpls
end module two_phase_levelset


module cahn_hilliard_solver
  use tpls_core_c
  implicit none
contains
  subroutine dim(maxl, maxn, dx, dy, dz, dt, epn, Pe, phi, phiOld, conv2_phi, &
      ConvPrevPhi, phiNew, sx, ex, sy, ey, sz_p, ez_p, height, &
      stride_p_xz, stride_p_yz, neighbours, comm2d_quasiperiodic, &
      max_iteration_dim)
    integer, intent(in) :: maxl, maxn
    integer, intent(in) :: sx, ex, sy, ey, sz_p, ez_p
    integer, intent(in) :: stride_p_xz, stride_p_yz
    integer, intent(in) :: neighbours(6)
    integer, intent(in) :: comm2d_quasiperiodic
    integer, intent(in) :: max_iteration_dim
    double precision, intent(in) :: dx, dy, dz, dt
    double precision, intent(in) :: Pe, epn
    double precision, intent(in) :: height
    double precision, intent(in) :: phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(in) :: phiOld(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(in) :: conv2_phi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: ConvPrevPhi(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    double precision, intent(inout) :: phiNew(sx-1:ex+1, sy-1:ey+1, sz_p:ez_p)
    if (sz_p /= 0 .or. ez_p /= maxn) call tpls_check(-2_c_int, 'dim: z bounds')
    call tpls_check(c_dim(maxl, maxn, dx, dy, dz, dt, epn, Pe, phi, phiOld, conv2_phi, ConvPrevPhi, &
        phiNew, sx, ex, sy, ey, height, max_iteration_dim, c_null_ptr), 'dim')
  end subroutine dim
end module cahn_hilliard_solver


module momentum_allflux
  use tpls_core_c
  implicit none
contains
  subroutine get_diffusion_all(diff_u, diff_v, diff_w, viscosity, u2, v2, w2, sx, sy, ex, ey, maxn, &
      ex_max, dx, dy, dz)
    integer, intent(in) :: sx, sy, ex, ey, maxn, ex_max
    double precision, intent(in) :: u2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: diff_u(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: v2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: diff_v(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: w2(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(inout) :: diff_w(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz
    call tpls_check(c_diffusion(diff_u, diff_v, diff_w, viscosity, u2, v2, w2, sx, sy, ex, ey, maxn, &
        ex_max, dx, dy, dz), 'get_diffusion_all')
  end subroutine get_diffusion_all

  subroutine get_conv_all(conv_u2, conv_v2, conv_w2, viscosity_aug, u2, v2, w2, sx, sy, ex, ey, &
      maxn, ex_max, dx, dy, dz)
    integer, intent(in) :: sx, ex, sy, ey, ex_max, maxn
    double precision, intent(in) :: dx, dy, dz
    double precision, intent(inout) :: conv_u2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: u2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: conv_v2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: v2(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: conv_w2(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: w2(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: viscosity_aug(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    call tpls_check(c_conv(conv_u2, conv_v2, conv_w2, viscosity_aug, u2, v2, w2, sx, sy, ex, ey, maxn, &
        ex_max, dx, dy, dz), 'get_conv_all')
  end subroutine get_conv_all

  subroutine get_viscosity(viscosity, phi, ex, ey, sx, sy, maxn, smooth_width, Re, mu_minus, mu_plus)
    integer, intent(in) :: sx, ex, sy, ey, maxn
    double precision, intent(inout) :: viscosity(sx-1:ex+1, sy-1:ey+1, 0:maxn)
    double precision, intent(in) :: phi(sx-1:ex+1, sy-1:ey+1, 0:maxn)
    double precision, intent(in) :: smooth_width, Re, mu_minus, mu_plus
    call tpls_check(c_viscosity(viscosity, phi, ex, ey, sx, sy, maxn, smooth_width, Re, mu_minus, &
        mu_plus), 'get_viscosity')
  end subroutine get_viscosity
end module momentum_allflux


module jacobi_iteration_allflux
  use tpls_core_c
  implicit none
contains
  subroutine do_jacobi_u(u3, u3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max)
    integer, intent(in) :: sx, sy, ex, ey, maxn, ex_max
    double precision, intent(inout) :: u3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: u3_old(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_jacobi_u(u3, u3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, &
        ex_max), 'do_jacobi_u')
  end subroutine do_jacobi_u

  subroutine do_jacobi_v(v3, v3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max)
    integer, intent(in) :: sx, sy, ex, ey, maxn, ex_max
    double precision, intent(inout) :: v3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: v3_old(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_jacobi_v(v3, v3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, &
        ex_max), 'do_jacobi_v')
  end subroutine do_jacobi_v

  subroutine do_jacobi_w(w3, w3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max)
    integer, intent(in) :: sx, sy, ex, ey, maxn, ex_max
    double precision, intent(inout) :: w3(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: w3_old(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_jacobi_w(w3, w3_old, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, &
        ex_max), 'do_jacobi_w')
  end subroutine do_jacobi_w

  subroutine get_difference(ua, ub, ex, ey, ez, sx, sy, sz, diff)
    integer, intent(in) :: ex, ey, ez, sx, sy, sz
    double precision, dimension(sx-1:ex+1, sy-1:ey+1, sz:ez), intent(in) :: ua, ub
    double precision, intent(out) :: diff
    call tpls_check(c_difference(ua, ub, ex, ey, ez, sx, sy, sz, diff), 'get_difference')
  end subroutine get_difference
end module jacobi_iteration_allflux


module sor_iteration_allflux
  use tpls_core_c
  implicit none
contains
  subroutine do_sor_u(u3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max, flg, &
      iteration_time)
    integer, intent(in) :: sx, sy, ex, ey, flg, maxn, iteration_time, ex_max
    double precision, intent(inout) :: u3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_sor_u(u3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max, flg, &
        iteration_time), 'do_sor_u')
  end subroutine do_sor_u

  subroutine do_sor_v(v3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, iteration_time)
    integer, intent(in) :: sx, sy, ex, ey, flg, maxn, iteration_time
    double precision, intent(inout) :: v3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_sor_v(v3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, &
        iteration_time), 'do_sor_v')
  end subroutine do_sor_v

  subroutine do_sor_w(w3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, iteration_time)
    integer, intent(in) :: sx, sy, ex, ey, flg, maxn, iteration_time
    double precision, intent(inout) :: w3(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: RHS(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: viscosity(sx-2:ex+2, sy-2:ey+2, 0:maxn)
    double precision, intent(in) :: dx, dy, dz, dt
    call tpls_check(c_sor_w(w3, RHS, viscosity, dx, dy, dz, dt, ex, ey, sx, sy, maxn, flg, &
        iteration_time), 'do_sor_w')
  end subroutine do_sor_w
end module sor_iteration_allflux


module pressure
  use tpls_core_c
  implicit none
contains
  subroutine get_source_pres(rhs, u3, v3, w3, ex, ey, sx, sy, maxn, dx, dy, dz, dt)
    integer, intent(in) :: ex, ey, sx, sy, maxn
    double precision, intent(in) :: dx, dy, dz, dt
    double precision, intent(inout) :: rhs(sx-1:ex+1, sy-1:ey+1, 0:maxn)
    double precision, intent(in) :: u3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: v3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(in) :: w3(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    call tpls_check(c_source_pres(rhs, u3, v3, w3, ex, ey, sx, sy, maxn, dx, dy, dz, dt), &
        'get_source_pres')
  end subroutine get_source_pres

  subroutine get_uvw_pres(u3, v3, w3, p, ex, ey, sx, sy, maxn, ex_max, dx, dy, dz, dt, u_inlet)
    integer, intent(in) :: ex, ey, sx, sy, maxn, ex_max
    double precision, intent(in) :: dx, dy, dz, dt
    double precision, intent(in) :: p(sx-1:ex+1, sy-1:ey+1, 0:maxn)
    double precision, intent(inout) :: u3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: v3(sx-1:ex+1, sy-1:ey+1, 0:maxn-2)
    double precision, intent(inout) :: w3(sx-1:ex+1, sy-1:ey+1, 0:maxn-1)
    double precision, intent(in) :: u_inlet(0:maxn-2)
    call tpls_check(c_uvw_pres(u3, v3, w3, p, ex, ey, sx, sy, maxn, ex_max, dx, dy, dz, dt, u_inlet), &
        'get_uvw_pres')
  end subroutine get_uvw_pres
end module pressure


! The KSPSolve sequence of main_ns_hybrid.F90:896-913 (initial guess from
! pres, compute_rhs, compute_matrix, KSPSolve, iteration count, residual norm,
! converged reason, copy_pressure_out_of_petsc) as one call.  pres holds the
! initial guess on entry and the solution with its ghost planes on return.
! The solver-global data of state.F90:12-53, which the reference's PETSc
! callbacks (copy_pressure_in_to_petsc, compute_rhs, compute_matrix) read, in
! plain Fortran types: main_ns_hybrid.F90 allocates and fills these arrays
! as it does with the reference's module of the same name.
module tpls_state
  implicit none
  integer :: maxl, maxm, maxn
  integer :: global_dim_x, global_dim_y, global_dim_z
  double precision, allocatable, dimension(:, :, :) :: pres
  double precision :: dx, dy, dz, dt
  double precision, allocatable, dimension(:, :, :) :: u3
  double precision, allocatable, dimension(:, :, :) :: RHS_p
  double precision, allocatable, dimension(:) :: u_inlet
end module tpls_state


! The KSP calls of main_ns_hybrid.F90:896-913 and pressure_solver.F90's
! output_converged_reason (:88-145) and copy_pressure_out_of_petsc
! (:444-486), with their argument lists.  KSPSolve solves the system that
! compute_matrix / compute_rhs assemble (pressure_solver.F90:234-441) from
! tpls_state's RHS_p, u3, u_inlet and spacings, starting from pres as
! copy_pressure_in_to_petsc does, with conjugate gradients on the GPU (the
! CG, no-preconditioner configuration of the PETSc options database,
! constant null space, KSPConvergedReason codes); the solution stays in the
! KSP object until copy_pressure_out_of_petsc writes it, ghost planes
! included, into pres.  tpls_ksp_create replaces the KSP set-up of
! configure_mpi_petsc.F90:143-183, taking the tolerances that set-up reads
! from the options database.
module pressure_solver
  use tpls_core_c
  use tpls_state
  implicit none
  private
  public :: tpls_ksp, tpls_ksp_create, PETSC_NULL_OBJECT, KSPSolve, KSPGetResidualNorm, &
      KSPGetIterationNumber, KSPGetConvergedReason, output_converged_reason, copy_pressure_out_of_petsc

  integer(8), parameter :: PETSC_NULL_OBJECT = 0

  type tpls_ksp
     double precision :: rtol = 1.d-5, abstol = 1.d-50, dtol = 1.d5
     integer :: maxits = 10000
     integer :: its = 0, reason = 0
     double precision :: rnorm = 0.d0
     double precision, allocatable :: x(:, :, :)
  end type tpls_ksp

contains

  subroutine tpls_ksp_create(ksp, rtol, abstol, dtol, maxits)
    type(tpls_ksp), intent(out) :: ksp
    double precision, intent(in) :: rtol, abstol, dtol
    integer, intent(in) :: maxits
    ksp%rtol = rtol
    ksp%abstol = abstol
    ksp%dtol = dtol
    ksp%maxits = maxits
  end subroutine tpls_ksp_create

  subroutine KSPSolve(ksp, b, x, ierr)
    type(tpls_ksp), intent(inout) :: ksp
    integer(8), intent(in) :: b, x
    integer, intent(out) :: ierr
    if (.not. allocated(ksp%x)) allocate(ksp%x(0:maxl, 0:maxm, 0:maxn))
    ksp%x = pres
    ierr = c_pressure_solve(ksp%x, RHS_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt, ksp%rtol, &
        ksp%abstol, ksp%dtol, ksp%maxits, ksp%its, ksp%rnorm, ksp%reason)
  end subroutine KSPSolve

  subroutine KSPGetResidualNorm(ksp, rnorm, ierr)
    type(tpls_ksp), intent(in) :: ksp
    double precision, intent(out) :: rnorm
    integer, intent(out) :: ierr
    rnorm = ksp%rnorm
    ierr = 0
  end subroutine KSPGetResidualNorm

  subroutine KSPGetIterationNumber(ksp, its, ierr)
    type(tpls_ksp), intent(in) :: ksp
    integer, intent(out) :: its, ierr
    its = ksp%its
    ierr = 0
  end subroutine KSPGetIterationNumber

  subroutine KSPGetConvergedReason(ksp, reason, ierr)
    type(tpls_ksp), intent(in) :: ksp
    integer, intent(out) :: reason, ierr
    reason = ksp%reason
    ierr = 0
  end subroutine KSPGetConvergedReason

  ! The name of PETSc's KSPConvergedReason code, printed as the reference does.
  subroutine output_converged_reason(ksp)
    type(tpls_ksp), intent(inout) :: ksp
    character(29) :: converged_reason
    select case (ksp%reason)
    case (-10); converged_reason = 'KSP_DIVERGED_INDEFINITE_MAT'
    case (-9); converged_reason = 'KSP_DIVERGED_NAN'
    case (-8); converged_reason = 'KSP_DIVERGED_INDEFINITE_PC'
    case (-7); converged_reason = 'KSP_DIVERGED_NONSYMMETRIC'
    case (-6); converged_reason = 'KSP_DIVERGED_BREAKDOWN_BICG'
    case (-5); converged_reason = 'KSP_DIVERGED_BREAKDOWN'
    case (-4); converged_reason = 'KSP_DIVERGED_DTOL'
    case (-3); converged_reason = 'KSP_DIVERGED_ITS'
    case (-2); converged_reason = 'KSP_DIVERGED_NULL'
    case (-1); converged_reason = 'UNUSED'
    case (0); converged_reason = 'KSP_CONVERGED_ITERATING'
    case (1); converged_reason = 'KSP_CONVERGED_RTOL_NORMAL'
    case (2); converged_reason = 'KSP_CONVERGED_RTOL'
    case (3); converged_reason = 'KSP_CONVERGED_ATOL'
    case (4); converged_reason = 'KSP_CONVERGED_ITS'
    case (5); converged_reason = 'KSP_CONVERGED_CG_NEG_CURVE'
    case (6); converged_reason = 'KSP_CONVERGED_CG_CONSTRAINED'
    case (7); converged_reason = 'KSP_CONVERGED_STEP_LENGTH'
    case (8); converged_reason = 'KSP_CONVERGED_HAPPY_BREAKDOWN'
    case (9); converged_reason = 'KSP_CONVERGED_ATOL_NORMAL'
    case default; converged_reason = ' '
    end select
    write(*, *) 'reason =', converged_reason
  end subroutine output_converged_reason

  subroutine copy_pressure_out_of_petsc(ksp, ierr)
    type(tpls_ksp), intent(in) :: ksp
    integer, intent(out) :: ierr
    if (.not. allocated(ksp%x)) then
       ierr = -1
       return
    end if
    pres = ksp%x
    ierr = 0
  end subroutine copy_pressure_out_of_petsc
end module pressure_solver


module tpls_pressure
  use tpls_core_c
  implicit none
contains
  subroutine tpls_ksp_solve(pres, RHS_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt, rtol, abstol, &
      dtol, maxits, its, rnorm, reason)
    integer, intent(in) :: maxl, maxm, maxn, maxits
    double precision, intent(in) :: dx, dy, dz, dt, rtol, abstol, dtol
    double precision, intent(inout) :: pres(0:maxl, 0:maxm, 0:maxn)
    double precision, intent(in) :: RHS_p(0:maxl, 0:maxm, 0:maxn)
    double precision, intent(in) :: u3(0:maxl, 0:maxm, 0:maxn-2)
    double precision, intent(in) :: u_inlet(0:maxn-2)
    integer, intent(out) :: its, reason
    double precision, intent(out) :: rnorm
    call tpls_check(c_pressure_solve(pres, RHS_p, u3, u_inlet, maxl, maxm, maxn, dx, dy, dz, dt, rtol, &
        abstol, dtol, maxits, its, rnorm, reason), 'tpls_ksp_solve')
  end subroutine tpls_ksp_solve
end module tpls_pressure


! The data-file procedures of io.F90:47-169 with their dummy argument lists,
! on the library's parallel reader and writer (include/tpls_io.h): a drop-in
! for the reference's tpls_io procedures of the same names (the module's
! other procedures write NetCDF and channel outputs and stay the reference's).
! The statuses are the reference's; a file the library cannot open or parse
! stops the program with its message, as the Fortran runtime would.
module tpls_io_native
  use iso_c_binding
  use tpls_core_c, only: tpls_check
  implicit none
  private
  public :: save_data, load_data, load_data_into_array
  interface
    integer(c_int) function c_save_data(filename, description, data, lx, ly, mx, my, nx, ny) &
        bind(C, name="tpls_save_data")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: filename(*), description(*)
      real(c_double), intent(in) :: data(*)
      integer(c_int), value :: lx, ly, mx, my, nx, ny
    end function
    integer(c_int) function c_data_bounds(filename, bounds) bind(C, name="tpls_data_bounds")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: filename(*)
      integer(c_int), intent(out) :: bounds(6)
    end function
    integer(c_int) function c_load_data(filename, data, bounds, ierr) bind(C, name="tpls_load_data")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: filename(*)
      type(c_ptr), intent(out) :: data
      integer(c_int), intent(out) :: bounds(6), ierr
    end function
    integer(c_int) function c_load_data_into_array(filename, data, lx, ly, mx, my, nx, ny, ierr) &
        bind(C, name="tpls_load_data_into_array")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: filename(*)
      real(c_double), intent(inout) :: data(*)
      integer(c_int), value :: lx, ly, mx, my, nx, ny
      integer(c_int), intent(out) :: ierr
    end function
    subroutine c_free(p) bind(C, name="tpls_free")
      import :: c_ptr
      type(c_ptr), value :: p
    end subroutine
  end interface
contains
  subroutine save_data(filename, description, data, lx, ly, mx, my, nx, ny)
    character(len=*), intent(in) :: filename
    character(len=*), intent(in) :: description
    integer, intent(in) :: lx, ly, mx, my, nx, ny
    double precision, intent(in) :: data(lx:ly, mx:my, nx:ny)
    call tpls_check(c_save_data(trim(filename) // c_null_char, description // c_null_char, data, &
        lx, ly, mx, my, nx, ny), 'save_data')
  end subroutine save_data

  subroutine load_data(filename, data, lx, ly, mx, my, nx, ny, ierr)
    character(len=*), intent(in) :: filename
    double precision, allocatable, dimension(:,:,:), intent(out) :: data
    integer, intent(out) :: lx, ly, mx, my, nx, ny, ierr
    type(c_ptr) :: p
    integer(c_int) :: b(6), e
    real(c_double), pointer :: flat(:)
    integer :: n
    call tpls_check(c_load_data(trim(filename) // c_null_char, p, b, e), 'load_data')
    lx = b(1); ly = b(2); mx = b(3); my = b(4); nx = b(5); ny = b(6)
    allocate(data(lx:ly, mx:my, nx:ny))
    n = size(data)
    if (n > 0) then
      call c_f_pointer(p, flat, [n])
      data = reshape(flat, shape(data))
    end if
    call c_free(p)
    ierr = e
  end subroutine load_data

  subroutine load_data_into_array(filename, data, lx, ly, mx, my, nx, ny, ierr)
    character(len=*), intent(in) :: filename
    double precision, allocatable, dimension(:,:,:), intent(inout) :: data
    integer, intent(in) :: lx, ly, mx, my, nx, ny
    integer, intent(out) :: ierr
    integer(c_int) :: e
    call tpls_check(c_load_data_into_array(trim(filename) // c_null_char, data, lx, ly, mx, my, nx, ny, e), &
        'load_data_into_array')
    ierr = e
  end subroutine load_data_into_array
end module tpls_io_native


! The halo exchanges of mpi.f90 that main_ns_hybrid.F90 calls between the
! procedures above (exchange2d :119, exchange2d_aug1 :215, exchange2d_aug2
! :294), with their argument lists, for the single rank whose y neighbours
! are itself: tpls_exchange2d copies the y halo rows, on the device when the
! array is attached to the mirror and its device copy is the newer.  The
! stride datatypes, neighbours and communicator are accepted and ignored.
module tpls_mpi
  use tpls_core_c
  implicit none
  integer :: master_id
  parameter(master_id=0)
contains
  logical function is_master(my_id)
    integer, intent(in) :: my_id
    is_master = (master_id == my_id)
  end function is_master

  ! mpi.f90:26-49 with one process in x and y: sx = 1, ex = maxl-1, sy = 1,
  ! ey = maxm-1 (the same arithmetic with dims = 1, coords = 0)
  subroutine mpi_decomp_2d(sx, ex, sy, ey, n_local_x, n_local_y, maxl, maxm, coords, dims, Ndim)
    integer, intent(inout) :: sx, ex, sy, ey, n_local_x, n_local_y
    integer, intent(in) :: maxl, maxm, Ndim
    integer, intent(in) :: coords(Ndim), dims(Ndim)
    n_local_x = (maxl - 1) / dims(1)
    sx = coords(1) * n_local_x + 1
    ex = sx + n_local_x - 1
    n_local_y = (maxm - 1) / dims(2)
    sy = coords(2) * n_local_y + 1
    ey = sy + n_local_y - 1
  end subroutine mpi_decomp_2d

  ! mpi.f90:52-66 on one rank: the y neighbours are the rank itself, the x
  ! and z neighbours none (-1); the exchanges above read none of them
  subroutine get_mpi_neighbours(neighbours, comm3d)
    integer :: neighbours(6), comm3d
    neighbours = -1
    neighbours(1) = 0
    neighbours(3) = 0
  end subroutine get_mpi_neighbours

  ! mpi.f90's stride datatypes (get_stride_p*, :69-604) describe the halo
  ! planes for MPI; the single-rank exchanges take the bounds instead, so
  ! the handles are set to 0
  subroutine get_stride_p(stride_p_yz, stride_p_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_xz, stride_p_yz, ex, ey, ez, sx, sy, sz
    stride_p_xz = 0
    stride_p_yz = 0
  end subroutine get_stride_p
  subroutine get_stride_p_aug1(stride_p_aug1_yz, stride_p_aug1_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_aug1_xz, stride_p_aug1_yz, ex, ey, ez, sx, sy, sz
    stride_p_aug1_xz = 0
    stride_p_aug1_yz = 0
  end subroutine get_stride_p_aug1
  subroutine get_stride_p_aug2(stride_p_aug2_yz, stride_p_aug2_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_aug2_xz, stride_p_aug2_yz, ex, ey, ez, sx, sy, sz
    stride_p_aug2_xz = 0
    stride_p_aug2_yz = 0
  end subroutine get_stride_p_aug2
  subroutine get_stride_p_augaug1(stride_p_augaug1_yz, stride_p_augaug1_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_augaug1_xz, stride_p_augaug1_yz, ex, ey, ez, sx, sy, sz
    stride_p_augaug1_xz = 0
    stride_p_augaug1_yz = 0
  end subroutine get_stride_p_augaug1
  subroutine get_stride_p_augaug2(stride_p_augaug2_yz, stride_p_augaug2_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_augaug2_xz, stride_p_augaug2_yz, ex, ey, ez, sx, sy, sz
    stride_p_augaug2_xz = 0
    stride_p_augaug2_yz = 0
  end subroutine get_stride_p_augaug2
  subroutine get_stride_p_augaug3(stride_p_augaug2_yz, stride_p_augaug2_xz, sx, ex, sy, ey, sz, ez)
    integer :: stride_p_augaug2_xz, stride_p_augaug2_yz, ex, ey, ez, sx, sy, sz
    stride_p_augaug2_xz = 0
    stride_p_augaug2_yz = 0
  end subroutine get_stride_p_augaug3

  subroutine exchange2d(uu, stride_p_xz, stride_p_yz, neighbours, ex, ey, ez, sx, sy, sz, comm_topology)
    integer, intent(in) :: ex, ey, ez, sx, sy, sz
    integer, intent(in) :: stride_p_xz, stride_p_yz
    integer, intent(in) :: neighbours(6), comm_topology
    double precision, dimension(sx-1:ex+1, sy-1:ey+1, sz:ez), intent(inout) :: uu
    call tpls_check(c_exchange2d(uu, ex, ey, ez, sx, sy, sz, 0), 'exchange2d')
  end subroutine exchange2d

  subroutine exchange2d_aug1(uu, stride_p_aug1_xz, stride_p_aug1_yz, neighbours, ex, ey, ez, sx, sy, sz, &
      comm_topology)
    integer, intent(in) :: ex, ey, ez, sx, sy, sz
    integer, intent(in) :: stride_p_aug1_xz, stride_p_aug1_yz
    integer, intent(in) :: neighbours(6), comm_topology
    double precision, dimension(sx-2:ex+2, sy-2:ey+2, sz:ez), intent(inout) :: uu
    call tpls_check(c_exchange2d(uu, ex, ey, ez, sx, sy, sz, 1), 'exchange2d_aug1')
  end subroutine exchange2d_aug1

  subroutine exchange2d_aug2(uu, stride_p_aug2_xz, stride_p_aug2_yz, neighbours, ex, ey, ez, sx, sy, sz, &
      comm_topology)
    integer, intent(in) :: ex, ey, ez, sx, sy, sz
    integer, intent(in) :: stride_p_aug2_xz, stride_p_aug2_yz
    integer, intent(in) :: neighbours(6), comm_topology
    double precision, dimension(sx-2:ex+2, sy-2:ey+2, sz:ez), intent(inout) :: uu
    call tpls_check(c_exchange2d(uu, ex, ey, ez, sx, sy, sz, 2), 'exchange2d_aug2')
  end subroutine exchange2d_aug2
end module tpls_mpi


! Device copies of main's arrays that stay resident between calls
! (tpls_mirror_attach in include/tpls_core.h).  main_ns_hybrid.F90 attaches
! its arrays once after allocating them (main_ns_hybrid.F90:416-549) and
! detaches them before deallocating them or reading files into them; its
! statements between calls read and write the arrays as before.
module tpls_mirror
  use iso_c_binding
  use tpls_core_c, only: tpls_check
  implicit none
  private
  public :: mirror_attach, mirror_detach, mirror_detach_all, mirror_flush, mirror_counters

  interface mirror_attach
    module procedure attach3, attach1
  end interface
  interface mirror_detach
    module procedure detach3, detach1
  end interface

  interface
    integer(c_int) function c_attach(p, n) bind(C, name="tpls_mirror_attach")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: p
      integer(c_long_long), value :: n
    end function
    integer(c_int) function c_detach(p) bind(C, name="tpls_mirror_detach")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function c_detach_all() bind(C, name="tpls_mirror_detach_all")
      import :: c_int
    end function
    integer(c_int) function c_flush() bind(C, name="tpls_mirror_flush")
      import :: c_int
    end function
    integer(c_int) function c_counters(out) bind(C, name="tpls_mirror_counters")
      import :: c_int, c_long_long
      integer(c_long_long), intent(out) :: out(6)
    end function
  end interface

contains
  subroutine attach3(a)
    double precision, intent(inout), target, contiguous :: a(:, :, :)
    call tpls_check(c_attach(c_loc(a), int(size(a), c_long_long) * 8_c_long_long), 'mirror_attach')
  end subroutine attach3

  subroutine attach1(a)
    double precision, intent(inout), target, contiguous :: a(:)
    call tpls_check(c_attach(c_loc(a), int(size(a), c_long_long) * 8_c_long_long), 'mirror_attach')
  end subroutine attach1

  subroutine detach3(a)
    double precision, intent(inout), target, contiguous :: a(:, :, :)
    call tpls_check(c_detach(c_loc(a)), 'mirror_detach')
  end subroutine detach3

  subroutine detach1(a)
    double precision, intent(inout), target, contiguous :: a(:)
    call tpls_check(c_detach(c_loc(a)), 'mirror_detach')
  end subroutine detach1

  subroutine mirror_detach_all()
    call tpls_check(c_detach_all(), 'mirror_detach_all')
  end subroutine mirror_detach_all

  subroutine mirror_flush()
    call tpls_check(c_flush(), 'mirror_flush')
  end subroutine mirror_flush

  ! bytes uploaded, bytes downloaded, faults, uploads, downloads, arrays attached
  subroutine mirror_counters(out)
    integer(8), intent(out) :: out(6)
    integer(c_long_long) :: c(6)
    call tpls_check(c_counters(c), 'mirror_counters')
    out = c
  end subroutine mirror_counters
end module tpls_mirror
