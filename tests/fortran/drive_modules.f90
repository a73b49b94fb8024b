! Test driver: calls the Fortran host modules (tpls_amd/fortran/tpls_modules.f90)
! exactly as main_ns_hybrid.F90 calls the reference modules, one procedure per
! run, with inputs read from a stream file and outputs written to another.
!
!   drive_modules <in> <out>
!
! Input stream: int32 routine id, int32 maxl, maxm, maxn, then the
! procedure's arguments in the reference's order: arrays as raw column-major
! values of their reference shape (single rank sx = 1, ex = maxl-1, sy = 1,
! ey = maxm-1), integers as int32, reals as float64.  Output stream: the
! procedure's output arrays and scalars in argument order.
program drive_modules
  use weno
  use advect_phi
  use levelset
  use two_phase_levelset
  use cahn_hilliard_solver
  use momentum_allflux
  use jacobi_iteration_allflux
  use sor_iteration_allflux
  use pressure
  use tpls_state, only: st_maxl => maxl, st_maxm => maxm, st_maxn => maxn, pres, RHS_p, u3, u_inlet, &
      st_dx => dx, st_dy => dy, st_dz => dz, st_dt => dt
  use pressure_solver
  implicit none

  integer :: rid, maxl, maxm, maxn, fin, fout
  character(len=512) :: fn_in, fn_out
  integer :: i1, i2, i3, i4, i5, i6, i7, i8
  double precision :: d1, d2, d3, d4, d5, s1
  double precision, allocatable :: p1(:,:,:), p2(:,:,:), p3(:,:,:), p4(:,:,:), p5(:,:,:)
  double precision, allocatable :: a1(:,:,:), a2(:,:,:), a3(:,:,:), a4(:,:,:), a5(:,:,:), a6(:,:,:)
  double precision, allocatable :: mu(:,:,:), aa(:,:,:), inl(:)
  integer, allocatable :: ist(:,:,:)
  integer :: nb(6)
  type(tpls_ksp) :: ksp
  integer :: ierr, its
  double precision :: rnorm

  call get_command_argument(1, fn_in)
  call get_command_argument(2, fn_out)
  open(newunit=fin, file=trim(fn_in), access='stream', form='unformatted', status='old')
  open(newunit=fout, file=trim(fn_out), access='stream', form='unformatted', status='replace')
  read(fin) rid, maxl, maxm, maxn
  nb = -1

  select case (rid)
  case (1)   ! WENOAppoxConvection(phi,u,v,w,sx,ex,sy,ey,l,n,dfx,dfy,dfz)
    call get_augaug(aa); call get_uv(a1); call get_uv(a2); call get_w(a3)
    read(fin) i1, i2, i3, i4, i5, i6
    call get_uv(a4); call get_uv(a5); call get_w(a6)
    call WENOAppoxConvection(aa, a1, a2, a3, i1, i2, i3, i4, i5, i6, a4, a5, a6)
    write(fout) a4, a5, a6
  case (2)   ! get_levelset_status(phi,D_mx,istatus,sign_mx,sx,ex,sy,ey,n,h)
    call get_p(p1); call get_p(p2); call get_pi(ist); call get_p(p3)
    read(fin) i1, i2, i3, i4, i5, d1
    call get_levelset_status(p1, p2, ist, p3, i1, i2, i3, i4, i5, d1)
    write(fout) p2, ist, p3
  case (3)   ! do_levelset_iteration(phi,phi_reinit,D_mx,istatus,sign_mx,sx,ex,sy,ey,n,ex_max,h,height,error_val)
    call get_p(p1); call get_p(p2); call get_p(p3); call get_pi(ist); call get_p(p4)
    read(fin) i1, i2, i3, i4, i5, i6, d1, d2
    call do_levelset_iteration(p1, p2, p3, ist, p4, i1, i2, i3, i4, i5, i6, d1, d2, s1)
    write(fout) p2, s1
  case (4)   ! get_csf(fx,fy,fz,fi,sx,ex,sy,ey,n,ex_max,dx,dy,dz,scap,smooth_width)
    call get_p(p1); call get_p(p2); call get_p(p3); call get_p(p4)
    read(fin) i1, i2, i3, i4, i5, i6, d1, d2, d3, d4, d5
    call get_csf(p1, p2, p3, p4, i1, i2, i3, i4, i5, i6, d1, d2, d3, d4, d5)
    write(fout) p1, p2, p3
  case (5)   ! get_viscosity(viscosity,phi,ex,ey,sx,sy,maxn,smooth_width,Re,mu_minus,mu_plus)
    call get_p(p1); call get_p(p2)
    read(fin) i1, i2, i3, i4, i5, d1, d2, d3, d4
    call get_viscosity(p1, p2, i1, i2, i3, i4, i5, d1, d2, d3, d4)
    write(fout) p1
  case (6, 7)   ! get_conv_all / get_diffusion_all(c_u,c_v,c_w,mu,u2,v2,w2,sx,sy,ex,ey,maxn,ex_max,dx,dy,dz)
    call get_uv(a1); call get_uv(a2); call get_w(a3); call get_aug(mu)
    call get_uv(a4); call get_uv(a5); call get_w(a6)
    read(fin) i1, i2, i3, i4, i5, i6, d1, d2, d3
    if (rid == 6) then
      call get_conv_all(a1, a2, a3, mu, a4, a5, a6, i1, i2, i3, i4, i5, i6, d1, d2, d3)
    else
      call get_diffusion_all(a1, a2, a3, mu, a4, a5, a6, i1, i2, i3, i4, i5, i6, d1, d2, d3)
    end if
    write(fout) a1, a2, a3
  case (8, 9, 10)   ! do_jacobi_u/v/w(x,x_old,RHS,viscosity,dx,dy,dz,dt,ex,ey,sx,sy,maxn,ex_max)
    if (rid == 10) then
      call get_w(a1); call get_w(a2); call get_w(a3)
    else
      call get_uv(a1); call get_uv(a2); call get_uv(a3)
    end if
    call get_aug(mu)
    read(fin) d1, d2, d3, d4, i1, i2, i3, i4, i5, i6
    if (rid == 8) call do_jacobi_u(a1, a2, a3, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6)
    if (rid == 9) call do_jacobi_v(a1, a2, a3, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6)
    if (rid == 10) call do_jacobi_w(a1, a2, a3, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6)
    write(fout) a1
  case (11)   ! do_sor_u(u3,RHS,viscosity,dx,dy,dz,dt,ex,ey,sx,sy,maxn,ex_max,flg,iteration_time)
    call get_uv(a1); call get_uv(a2); call get_aug(mu)
    read(fin) d1, d2, d3, d4, i1, i2, i3, i4, i5, i6, i7, i8
    call do_sor_u(a1, a2, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6, i7, i8)
    write(fout) a1
  case (12, 13)   ! do_sor_v/w(x,RHS,viscosity,dx,dy,dz,dt,ex,ey,sx,sy,maxn,flg,iteration_time)
    if (rid == 13) then
      call get_w(a1); call get_w(a2)
    else
      call get_uv(a1); call get_uv(a2)
    end if
    call get_aug(mu)
    read(fin) d1, d2, d3, d4, i1, i2, i3, i4, i5, i6, i7
    if (rid == 12) call do_sor_v(a1, a2, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6, i7)
    if (rid == 13) call do_sor_w(a1, a2, mu, d1, d2, d3, d4, i1, i2, i3, i4, i5, i6, i7)
    write(fout) a1
  case (14)   ! get_source_pres(rhs,u3,v3,w3,ex,ey,sx,sy,maxn,dx,dy,dz,dt)
    call get_p(p1); call get_uv(a1); call get_uv(a2); call get_w(a3)
    read(fin) i1, i2, i3, i4, i5, d1, d2, d3, d4
    call get_source_pres(p1, a1, a2, a3, i1, i2, i3, i4, i5, d1, d2, d3, d4)
    write(fout) p1
  case (15)   ! get_uvw_pres(u3,v3,w3,p,ex,ey,sx,sy,maxn,ex_max,dx,dy,dz,dt,u_inlet)
    call get_uv(a1); call get_uv(a2); call get_w(a3); call get_p(p1)
    read(fin) i1, i2, i3, i4, i5, i6, d1, d2, d3, d4
    allocate(inl(0:maxn-2))
    read(fin) inl
    call get_uvw_pres(a1, a2, a3, p1, i1, i2, i3, i4, i5, i6, d1, d2, d3, d4, inl)
    write(fout) a1, a2, a3
  case (16)   ! do_advect_phi(phi2,u2,v2,w2,conv2_phi, bounds, maxl,maxn,ex_max, MPI args, dx,dy,dz, strides)
    call get_p(p1); call get_uv(a1); call get_uv(a2); call get_w(a3); call get_p(p2)
    read(fin) d1, d2, d3
    call do_advect_phi(p1, a1, a2, a3, p2, 1, maxl - 1, 1, maxm - 1, 0, maxn, 0, 0, maxn - 2, &
        maxn - 1, maxl, maxn, maxl - 1, nb, -1, d1, d2, d3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1)
    write(fout) p2
  case (17)   ! do_tpls(phi2,phi3,conv0,conv1,conv2, bounds, maxn,ex_max, strides, its, MPI args, height,err2,dx,dz,dt)
    call get_p(p1); call get_p(p2); call get_p(p3); call get_p(p4); call get_p(p5)
    read(fin) i1, d1, d2, d3, d4
    s1 = 0.d0
    call do_tpls(p1, p2, p3, p4, p5, 1, maxl - 1, 1, maxm - 1, 0, maxn, maxn, maxl - 1, -1, -1, &
        i1, nb, -1, d1, s1, d2, d3, d4)
    write(fout) p2, s1
  case (18)   ! dim(maxl,maxn,dx,dy,dz,dt,epn,Pe,phi,phiOld,conv2_phi,ConvPrevPhi,phiNew, bounds, height, strides, MPI args, its)
    call get_p(p1); call get_p(p2); call get_p(p3); call get_p(p4); call get_p(p5)
    read(fin) i1, d1, d2, d3, d4, d5
    call dim(maxl, maxn, d1, d1, d1, d2, d3, d4, p1, p2, p3, p4, p5, 1, maxl - 1, 1, maxm - 1, 0, maxn, &
        d5, -1, -1, nb, -1, i1)
    write(fout) p4, p5
  case (19)   ! main_ns_hybrid.F90:896-913: pres = pres_old, KSPSolve, the prints, copy_pressure_out_of_petsc
    st_maxl = maxl; st_maxm = maxm; st_maxn = maxn
    allocate(pres(0:maxl, 0:maxm, 0:maxn), RHS_p(0:maxl, 0:maxm, 0:maxn), u3(0:maxl, 0:maxm, 0:maxn-2))
    allocate(u_inlet(0:maxn-2))
    call get_p(p1)
    read(fin) RHS_p, u3, u_inlet
    read(fin) st_dx, st_dy, st_dz, st_dt, d1
    call tpls_ksp_create(ksp, d1, 1.d-50, 1.d5, 10000)
    pres = p1
    call KSPSolve(ksp, PETSC_NULL_OBJECT, PETSC_NULL_OBJECT, ierr)
    if (ierr /= 0) error stop 3
    call output_converged_reason(ksp)
    call KSPGetResidualNorm(ksp, rnorm, ierr)
    if (ierr /= 0) error stop 4
    write(*, *) 'approximate, preconditioned, residual norm', rnorm
    call KSPGetIterationNumber(ksp, its, ierr)
    if (ierr /= 0) error stop 5
    write(*, *) 'iterations =', its
    call copy_pressure_out_of_petsc(ksp, ierr)
    if (ierr /= 0) error stop 6
    call KSPGetConvergedReason(ksp, i1, ierr)
    write(fout) pres, its, i1, rnorm
  case default
    write(0, *) 'drive_modules: unknown routine id', rid
    error stop 2
  end select
  close(fin)
  close(fout)

contains

  subroutine get_p(a)
    double precision, allocatable, intent(out) :: a(:,:,:)
    allocate(a(0:maxl, 0:maxm, 0:maxn))
    read(fin) a
  end subroutine
  subroutine get_pi(a)
    integer, allocatable, intent(out) :: a(:,:,:)
    allocate(a(0:maxl, 0:maxm, 0:maxn))
    read(fin) a
  end subroutine
  subroutine get_uv(a)
    double precision, allocatable, intent(out) :: a(:,:,:)
    allocate(a(0:maxl, 0:maxm, 0:maxn-2))
    read(fin) a
  end subroutine
  subroutine get_w(a)
    double precision, allocatable, intent(out) :: a(:,:,:)
    allocate(a(0:maxl, 0:maxm, 0:maxn-1))
    read(fin) a
  end subroutine
  subroutine get_aug(a)
    double precision, allocatable, intent(out) :: a(:,:,:)
    allocate(a(-1:maxl+1, -1:maxm+1, 0:maxn))
    read(fin) a
  end subroutine
  subroutine get_augaug(a)
    double precision, allocatable, intent(out) :: a(:,:,:)
    allocate(a(-2:maxl+2, -2:maxm+2, 0:maxn))
    read(fin) a
  end subroutine
end program drive_modules
