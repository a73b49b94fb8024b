! Checker for the option-file tests: drives the reference's own options_utils
! module (options_utils.F90, compiled from the reference checkout by
! oracle/Makefile into oracle/_ref/) and prints what it does.
!
!   ref_options_driver load FILE
!       load_options(FILE); prints "ierr N" and "num N", then one line
!       "name|value|" per option (trailing blanks trimmed) and, for names
!       starting i_, d_ or l_, the value get_integer_option (I0),
!       get_double_option (bits as Z16.16) or get_logical_option (L1) reads.
!   ref_options_driver put FILE
!       reads "kind name value" lines from standard input (kind i: integer,
!       d: double given as 16 hex digits of its bits, l: T or F, s: string,
!       the rest of the line), stores each with put_integer_option,
!       put_double_option, put_logical_option or put_string_option, then
!       save_options(FILE) with the two header lines "ref header one" and
!       "ref header two".
program ref_options_driver
  use options_utils
  implicit none
  character(len=80), dimension(:,:), allocatable :: options
  character(len=256), dimension(2) :: hdr
  character(len=1024) :: mode, file, line, name
  character(len=256) :: rest
  integer :: num_options, ierr, i, iv, st, sp
  integer(kind=8) :: bits
  double precision :: dv
  logical :: lv

  call get_command_argument(1, mode)
  call get_command_argument(2, file)
  if (trim(mode) == 'load') then
     call load_options(file, options, num_options, ierr)
     write(*,'(A,I0)') 'ierr ', ierr
     if (ierr == option_file_not_found_err) stop
     write(*,'(A,I0)') 'num ', num_options
     do i = 1, num_options
        write(*,'(A,A,A,A)') trim(options(i,1)), '|', trim(options(i,2)), '|'
        if (options(i,1)(1:2) == 'i_') then
           call get_integer_option(options, num_options, trim(options(i,1)), iv, ierr)
           write(*,'(A,I0,A,I0)') 'int ', iv, ' ', ierr
        else if (options(i,1)(1:2) == 'd_') then
           call get_double_option(options, num_options, trim(options(i,1)), dv, ierr)
           write(*,'(A,Z16.16,A,I0)') 'dbl ', transfer(dv, bits), ' ', ierr
        else if (options(i,1)(1:2) == 'l_') then
           call get_logical_option(options, num_options, trim(options(i,1)), lv, ierr)
           write(*,'(A,L1,A,I0)') 'log ', lv, ' ', ierr
        end if
     end do
  else if (trim(mode) == 'put') then
     allocate(options(1000,2))
     num_options = 0
     do
        read(*,'(A)',iostat=st) line
        if (st /= 0) exit
        if (len_trim(line) == 0) cycle
        rest = adjustl(line(2:))
        sp = index(rest, ' ')
        name = rest(:sp-1)
        rest = rest(sp+1:)
        select case (line(1:1))
        case ('i')
           read(rest, *) iv
           call put_integer_option(options, num_options, trim(name), iv)
        case ('d')
           read(rest, '(Z16)') bits
           dv = transfer(bits, dv)
           call put_double_option(options, num_options, trim(name), dv)
        case ('l')
           lv = rest(1:1) == 'T'
           call put_logical_option(options, num_options, trim(name), lv)
        case default
           call put_string_option(options, num_options, trim(name), trim(rest))
        end select
     end do
     hdr(1) = 'ref header one'
     hdr(2) = 'ref header two'
     call save_options(file, options, num_options, hdr, 2)
  else
     write(*,'(A)') 'usage: ref_options_driver load FILE | put FILE < lines'
     stop 2
  end if
end program ref_options_driver
