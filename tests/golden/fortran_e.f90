! Writes tests/golden/fortran_e.txt: values through the edit descriptors of the
! reference's HisD and material files (E14.6E3, E18.10E3, E31.23E3), one value
! per line as "<hex of the double> <E14.6E3>|<E18.10E3>|<E31.23E3>".
! Build and run (ROCm flang):  flang fortran_e.f90 -o fe && ./fe > fortran_e.txt
program fortran_e
  use iso_fortran_env, only: int64, real64
  implicit none
  integer, parameter :: n = 40
  real(real64) :: v(n)
  integer(int64) :: bits
  integer :: i
  integer(int64) :: s
  v(1:12) = [1.0d0, -2.5d-7, 0.0d0, 123456.789d0, -9.9999999999995d0, 1.0d-300, 1.0d300, &
             3.0517578125d-5, 0.1d0, -166.0d0, 0.5d0, 9.99999949d0]
  s = 88172645463325252_int64
  do i = 13, n
    s = ieor(s, ishft(s, 13)); s = ieor(s, ishft(s, -7)); s = ieor(s, ishft(s, 17))
    v(i) = (dble(mod(abs(s), 2000001_int64)) - 1.0d6) * 10.0d0**(mod(i * 7, 31) - 15) / 7.0d0
  end do
  do i = 1, n
    bits = transfer(v(i), bits)
    write(*, "(Z16.16,1x,E14.6E3,'|',E18.10E3,'|',E31.23E3)") bits, v(i), v(i), v(i)
  end do
end program fortran_e
