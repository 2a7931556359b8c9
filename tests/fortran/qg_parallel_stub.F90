! Test scaffold: the names of the reference's qg_parallel module
! (qg3p_base.F90:108-199, 383-403) that qg_helmholtz_hip and qg_cabaret_hip
! use, so that the drop-in modules can be compiled and run standalone.  Only
! the fields the drop-ins read are present.
module qg_parallel

  implicit none

  private

  public :: decomposition, parallel_data, parallel_flux_data, pabort
  integer, parameter, public :: FLUX_II = 1, FLUX_JJ = 2

  type decomposition
    integer :: n = 1
    logical :: periodic_j = .false.
    integer :: layers = 1
    integer :: ii = 0, jj = 0
  end type decomposition

  type parallel_data
    type(decomposition), pointer :: decomp => null()
    real(kind = kind(1.0D0)), dimension(:, :, :), pointer :: data => null()
  end type parallel_data

  type parallel_flux_data
    type(decomposition), pointer :: decomp => null()
    integer :: flux_dir = FLUX_II
    real(kind = kind(1.0D0)), dimension(:, :, :), pointer :: data => null()
  end type parallel_flux_data

contains

  subroutine pabort(error, filename, line)
    character(len = *), intent(in) :: error
    character(len = *), intent(in) :: filename
    integer, intent(in) :: line

    write(0, "(a, a, a, i0, a, a)") "pabort: ", filename, ":", line, ": ", error
    stop 3
  end subroutine pabort

end module qg_parallel
