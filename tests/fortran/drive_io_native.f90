! The known-answer checks of test/io_test.f90:71-165 run through module
! tpls_io_native (the library's reader and writer behind io.F90's procedure
! signatures): data (i+1)(j+2)(k+3) on bounds 1:9, 0:10, 1:11, saved, loaded
! both ways, and the bound mismatch statuses 1, 2, 2, 3, 3.  Prints one line
! per check and "ALL PASSED" at the end; stops with status 1 on a failure.
program drive_io_native
  use tpls_io_native
  implicit none
  integer, parameter :: lx = 1, ly = 9, mx = 0, my = 10, nx = 1, ny = 11
  double precision, allocatable :: out_data(:,:,:), in_data(:,:,:)
  integer :: i, j, k, ierr, b(6), fails
  character(len=*), parameter :: f = 'drive_io_native.tmp'

  fails = 0
  allocate(out_data(lx:ly, mx:my, nx:ny))
  do k = nx, ny
    do j = mx, my
      do i = lx, ly
        out_data(i, j, k) = (i + 1) * (j + 2) * (k + 3)
      end do
    end do
  end do
  call save_data(f, 'Test data', out_data, lx, ly, mx, my, nx, ny)
  call load_data(f, in_data, b(1), b(2), b(3), b(4), b(5), b(6), ierr)
  call check('load_data status', ierr == 0)
  call check('load_data bounds', all(b == [lx, ly, mx, my, nx, ny]))
  call check('load_data values', all(in_data == out_data))
  deallocate(in_data)
  allocate(in_data(lx:ly, mx:my, nx:ny))
  in_data = 0
  call load_data_into_array(f, in_data, lx, ly, mx, my, nx, ny, ierr)
  call check('load_data_into_array status', ierr == 0)
  call check('load_data_into_array values', all(in_data == out_data))
  call load_data_into_array(f, in_data, lx, ly + 1, mx, my, nx, ny, ierr)
  call check('mismatch l', ierr == 1)
  call load_data_into_array(f, in_data, lx, ly, mx + 1, my, nx, ny, ierr)
  call check('mismatch m lower', ierr == 2)
  call load_data_into_array(f, in_data, lx, ly, mx, my + 1, nx, ny, ierr)
  call check('mismatch m upper', ierr == 2)
  call load_data_into_array(f, in_data, lx, ly, mx, my, nx + 1, ny, ierr)
  call check('mismatch n lower', ierr == 3)
  call load_data_into_array(f, in_data, lx, ly, mx, my, nx, ny + 1, ierr)
  call check('mismatch n upper', ierr == 3)
  open(unit=7, file=f, status='old')
  close(7, status='delete')
  if (fails > 0) stop 1
  write(*, '(a)') 'ALL PASSED'
contains
  subroutine check(what, ok)
    character(len=*), intent(in) :: what
    logical, intent(in) :: ok
    if (ok) then
      write(*, '(a, a)') 'pass ', what
    else
      write(*, '(a, a)') 'FAIL ', what
      fails = fails + 1
    end if
  end subroutine check
end program drive_io_native
