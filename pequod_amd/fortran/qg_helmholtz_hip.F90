! qg_helmholtz_hip -- drop-in replacement of the reference's qg_helmholtz
! module (qg3p_helmholtz.F90) on one MI355X.  It exports the same four names
! that qg_aux imports (qg3p_aux.F90:13-17):
!
!   use qg_helmholtz_hip, only : solve_helmholtz => helm, &
!     & helmholtz_solver_cache => helm_cache
!   use qg_helmholtz_hip, only : allocate, deallocate
!
! allocate(helm_c, decomp, ss) / helm(helm_c, zeta) / deallocate(helm_c) keep
! the reference's signatures and semantics: helm solves
! (Delta_5 - ss_l) psi_l = zeta_l in place with homogeneous Dirichlet values
! and the reference's output boundary values (qg3p_helmholtz.F90:614-769).
! Errors abort through the reference's pabort (qg3p_base.F90:383-403), as its
! labort macro does.  One domain, closed basin (decomp%n == 1, not periodic
! in j), G - 1 = 2^a or 3 * 2^a.
module qg_helmholtz_hip

  use iso_c_binding
  use pequod_hip_c
  use qg_parallel, only : decomposition, pabort

  implicit none

  private

  integer, parameter :: real_8 = kind(1.0D0)

  public :: helm
  public :: helm_cache
  public :: allocate, deallocate

  ! decomp and ss are the reference's public components that qg_aux reads
  ! (correction_function_mass, qg3p_aux.F90:196-230); h is the library's
  ! handle.
  type helm_cache
    type(c_ptr) :: h = c_null_ptr
    type(decomposition), pointer :: decomp => null()
    real(kind = real_8), dimension(:), pointer :: ss => null()
  end type helm_cache

  interface allocate
    module procedure allocate_helm_c
  end interface allocate

  interface deallocate
    module procedure deallocate_helm_c
  end interface deallocate

contains

  ! Allocate a helm cache (qg3p_helmholtz.F90:292-386).
  subroutine allocate_helm_c(helm_c, decomp, ss)
    type(helm_cache), intent(out) :: helm_c
    type(decomposition), target, intent(in) :: decomp
    real(kind = real_8), dimension(decomp%layers), intent(in) :: ss

    integer(c_int) :: rc

    if (decomp%n /= 1) then
      call pabort("qg_helmholtz_hip: one domain only (run with one MPI process)", &
        & __FILE__, __LINE__)
    end if
    if (decomp%periodic_j) then
      call pabort("qg_helmholtz_hip: closed basin only", __FILE__, __LINE__)
    end if
    helm_c%decomp => decomp
    allocate(helm_c%ss(decomp%layers))
    helm_c%ss = ss
    rc = pq_helm_create(helm_c%h, int(decomp%ii, c_int), int(decomp%jj, c_int), &
      & int(decomp%layers, c_int), ss, PQ_PREC_FP64)
    if (rc /= PQ_OK) then
      call pabort("qg_helmholtz_hip: " // pq_error_message(), __FILE__, __LINE__)
    end if
  end subroutine allocate_helm_c

  ! Deallocate a helm cache (qg3p_helmholtz.F90:389).
  subroutine deallocate_helm_c(helm_c)
    type(helm_cache), intent(inout) :: helm_c

    integer(c_int) :: rc

    rc = pq_helm_destroy(helm_c%h)
    helm_c%h = c_null_ptr
    nullify(helm_c%decomp)
    if (associated(helm_c%ss)) deallocate(helm_c%ss)
  end subroutine deallocate_helm_c

  ! On input zeta is the RHS, on output the solution (qg3p_helmholtz.F90:614).
  ! zeta is assumed-shape as in the reference; it reaches the C ABI's
  ! assumed-size dummy by sequence association, so a contiguous actual (the
  ! cache's phi, the reference's only caller) is passed in place and a
  ! non-contiguous one through the compiler's copy-in / copy-out.
  subroutine helm(helm_c, zeta)
    type(helm_cache), intent(inout) :: helm_c
    real(kind = real_8), dimension(:, :, :), intent(inout) :: zeta

    integer(c_int) :: rc

    if (size(zeta, 1) /= helm_c%decomp%ii .or. size(zeta, 2) /= helm_c%decomp%jj &
      & .or. size(zeta, 3) /= helm_c%decomp%layers) then
      call pabort("qg_helmholtz_hip: zeta shape differs from the cache's decomposition", &
        & __FILE__, __LINE__)
    end if

    rc = pq_helm_solve(helm_c%h, zeta, int(size(zeta, 1), c_int), &
      & int(size(zeta, 2), c_int))
    if (rc /= PQ_OK) then
      call pabort("qg_helmholtz_hip: " // pq_error_message(), __FILE__, __LINE__)
    end if
  end subroutine helm

end module qg_helmholtz_hip
