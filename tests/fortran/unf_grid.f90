! Fortran side of the output-format tests: the grid files that a Fortran
! runtime writes and reads with the reference's statements.
!
!   unf_grid write <file> <nr> <nt> <np>
!     writes a synthetic grid with the record layout of output_grid_real8
!     (output_2_test.f90:33-41 and :74-80): the four integers, the three
!     cell-centre lists, the cell array.
!   unf_grid read <file> <nr> <nt> <np>
!     reads a file with the statements of initarr_2_new.f90:60-66 and
!     prints the four integers, then every value of the three lists and of
!     the array, one per line.
program unf_grid
  implicit none
  character(len=512) :: mode, fname, arg
  integer :: nr, nt, np, ir, it, ip, id
  real(8), allocatable :: d1(:), d2(:), d3(:), a(:,:,:)

  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  call get_command_argument(3, arg)
  read(arg, *) nr
  call get_command_argument(4, arg)
  read(arg, *) nt
  call get_command_argument(5, arg)
  read(arg, *) np
  allocate(d1(nr), d2(nt), d3(np), a(nr, nt, np))

  if (trim(mode) == 'write') then
     do ir = 1, nr
        d1(ir) = 1.5d0 * ir + 0.0625d0
     end do
     do it = 1, nt
        d2(it) = 0.01d0 * it
     end do
     do ip = 1, np
        d3(ip) = 0.5d0 * ip
     end do
     do ip = 1, np
        do it = 1, nt
           do ir = 1, nr
              a(ir, it, ip) = 1000.d0 * ir + 10.d0 * it + 0.25d0 * ip + 0.125d0
           end do
        end do
     end do
     open(unit=15, file=trim(fname), form='unformatted')
     write(15) nr, nt, np, 1
     write(15) (d1(ir), ir=1, nr)
     write(15) (d2(it), it=1, nt)
     write(15) (d3(ip), ip=1, np)
     write(15) (((a(ir, it, ip), ir=1, nr), it=1, nt), ip=1, np)
     close(15)
  else
     open(unit=14, file=trim(fname), status='old', form='unformatted')
     read(14) ir, it, ip, id
     read(14) d1
     read(14) d2
     read(14) d3
     read(14) a
     close(14)
     write(*, '(4I8)') ir, it, ip, id
     write(*, '(ES25.17)') d1
     write(*, '(ES25.17)') d2
     write(*, '(ES25.17)') d3
     write(*, '(ES25.17)') a
  end if
end program unf_grid
