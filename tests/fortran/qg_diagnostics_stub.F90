! Test scaffold: the two types of the reference's qg_diagnostics module
! (qg3p_diagnostics.F90:25-55) that qg_cabaret_hip reads, with only the
! fields it reads.
module qg_diagnostics

  implicit none

  private

  public :: stratification, velocity_boundary_condition

  type stratification
    integer :: layers = 1
    real(kind = kind(1.0D0)), dimension(:), pointer :: s1 => null()
    real(kind = kind(1.0D0)), dimension(:), pointer :: s2 => null()
  end type stratification

  type velocity_boundary_condition
    integer :: bc_type = 1
    real(kind = kind(1.0D0)) :: alpha = 0.0D0
  end type velocity_boundary_condition

end module qg_diagnostics
