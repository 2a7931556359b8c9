! Exercises the Fortran drop-in (qg_helmholtz_hip over libpequod_hip.so).
!   helm_drop_in args   : argument checking through the C ABI only (no GPU
!                         work): an unsupported grid is rejected with
!                         PQ_ERR_ARG and a message.  Prints "args ok".
!   helm_drop_in solve  : allocate / helm / deallocate on a 33 x 33 x 2
!                         discrete eigenfunction sin(p pi i/N) sin(q pi j/N),
!                         whose exact solution is rhs / (lambda_p + lambda_q
!                         - ss_l); prints the max relative interior error.
program helm_drop_in

  use iso_c_binding
  use pequod_hip_c
  use qg_parallel, only : decomposition
  use qg_helmholtz_hip, only : helm, helm_cache, allocate, deallocate

  implicit none

  character(len = 16) :: mode
  type(c_ptr) :: h
  integer(c_int) :: rc
  real(c_double) :: ss(2), m2z(4), z2m(4)

  call get_command_argument(1, mode)
  if (trim(mode) == "args") then
    ss = (/ 0.0d0, 0.5d0 /)
    m2z = (/ 1.0d0, 1.0d0, 1.0d0, -1.0d0 /)
    z2m = (/ 0.5d0, 0.5d0, 0.5d0, -0.5d0 /)
    rc = pq_inv_create(h, 100_c_int, 2_c_int, ss, m2z, z2m, PQ_PREC_FP64, 1_c_int)
    if (rc /= PQ_ERR_ARG) stop 1
    if (len(pq_error_message()) == 0) stop 2
    write(*, "(a)") "args ok"
  else
    call solve_case()
  end if

contains

  subroutine solve_case()
    integer, parameter :: g = 33, nl = 2, p = 3, q = 5
    real(kind = kind(1.0d0)), parameter :: pi = 3.14159265358979323846d0
    type(decomposition), target :: decomp
    type(helm_cache) :: hc
    real(kind = kind(1.0d0)) :: zeta(g, g, nl), expect(g, g, nl), lam_p, lam_q, err, scl
    real(kind = kind(1.0d0)) :: s(nl)
    integer :: i, j, l, n

    n = g - 1
    s = (/ 0.0d0, 0.5d0 /)
    decomp%layers = nl
    decomp%ii = g
    decomp%jj = g
    lam_p = -4.0d0 * sin(p * pi / (2.0d0 * n)) ** 2
    lam_q = -4.0d0 * sin(q * pi / (2.0d0 * n)) ** 2
    do l = 1, nl
      do j = 1, g
        do i = 1, g
          zeta(i, j, l) = sin(p * pi * (i - 1) / n) * sin(q * pi * (j - 1) / n)
          expect(i, j, l) = zeta(i, j, l) / (lam_p + lam_q - s(l))
        end do
      end do
    end do
    call allocate(hc, decomp, s)
    call helm(hc, zeta)
    call deallocate(hc)
    err = 0.0d0
    scl = maxval(abs(expect))
    do l = 1, nl
      err = max(err, maxval(abs(zeta(2:g - 1, 2:g - 1, l) - expect(2:g - 1, 2:g - 1, l))) / scl)
    end do
    write(*, "(a, es12.4)") "max rel err ", err
  end subroutine solve_case

end program helm_drop_in
