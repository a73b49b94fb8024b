! Fortran list-directed I/O in the record layout of a 3-D data file (the
! statements of io.F90:60-72 and io.F90:98-109), for the interoperability
! tests of tests/test_io.py:
!   drive_io write RAW DAT lx ly mx my nx ny
!       RAW holds the values as raw doubles (i fastest); DAT is written with
!       one list-directed record per value, as save_data writes it
!   drive_io read DAT RAW
!       DAT is read with the list-directed reads of load_data; RAW gets the
!       six bounds (int32) and the values (raw doubles)
program drive_io
  implicit none
  character(len=1024) :: mode, src, dst, arg
  integer :: b(6), i, j, k, q
  double precision, allocatable :: a(:,:,:)

  call get_command_argument(1, mode)
  call get_command_argument(2, src)
  call get_command_argument(3, dst)
  if (trim(mode) == 'write') then
     do q = 1, 6
        call get_command_argument(3 + q, arg)
        read(arg, *) b(q)
     end do
     allocate(a(b(1):b(2), b(3):b(4), b(5):b(6)))
     open(unit=10, file=src, access='stream', form='unformatted', status='old', action='read')
     read(10) a
     close(10)
     open(unit=11, file=dst, status='replace')
     write(11,*) 'DATA: ', 'written by drive_io'
     write(11,*) 'DIMENSIONS: Lmin Lmax Mmin Mmax Nmin Nmax'
     write(11,*) b(1), b(2), b(3), b(4), b(5), b(6)
     do k = b(5), b(6)
        do j = b(3), b(4)
           do i = b(1), b(2)
              write(11,*) a(i,j,k)
           end do
        end do
     end do
     close(11)
  else if (trim(mode) == 'read') then
     open(unit=10, file=src, action='read', status='old')
     read(10,*)
     read(10,*)
     read(10,*) b(1), b(2), b(3), b(4), b(5), b(6)
     allocate(a(b(1):b(2), b(3):b(4), b(5):b(6)))
     do k = b(5), b(6)
        do j = b(3), b(4)
           do i = b(1), b(2)
              read(10,*) a(i,j,k)
           end do
        end do
     end do
     close(10)
     open(unit=11, file=dst, access='stream', form='unformatted', status='replace')
     write(11) b
     write(11) a
     close(11)
  else
     write(*,'(A)') 'usage: drive_io write RAW DAT lx ly mx my nx ny | drive_io read DAT RAW'
     stop 2
  end if
end program drive_io
