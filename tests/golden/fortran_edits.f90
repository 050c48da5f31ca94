! E15.7 / G15.7 edits and the E9.2 frequency tag of the result-file names, printed by the
! compiler (expected strings of tests/test_obsgrid.py; run make_fortran_edits.sh)
program fortran_edits
implicit none
real(8) :: x(9) = (/1.d0, -2.5d-5, 0.d0, 0.5d0, 12345.678d0, 9999999.6d0, 0.099999996d0, -180.d0, 1.d0/44714.d0/)
integer :: i
character(9) :: num
do i = 1, 9
  write(*,'(a,e15.7,a)') '[', x(i), ']'
  write(*,'(a,g15.7,a)') '[', x(i), ']'
end do
write(num,'(e9.2)') 1.d0/44714.d0
write(*,'(a,a,a)') '[', num, ']'
end program fortran_edits
