! Test scaffold: the two names of the reference's qg_parallel module
! (qg3p_base.F90:108-173, 383-403) that qg_helmholtz_hip uses, so that the
! drop-in module can be compiled and run standalone.  Only the fields the
! drop-in reads are present.
module qg_parallel

  implicit none

  private

  public :: decomposition, pabort

  type decomposition
    integer :: n = 1
    logical :: periodic_j = .false.
    integer :: layers = 1
    integer :: ii = 0, jj = 0
  end type decomposition

contains

  subroutine pabort(error, filename, line)
    character(len = *), intent(in) :: error
    character(len = *), intent(in) :: filename
    integer, intent(in) :: line

    write(0, "(a, a, a, i0, a, a)") "pabort: ", filename, ":", line, ": ", error
    stop 3
  end subroutine pabort

end module qg_parallel
