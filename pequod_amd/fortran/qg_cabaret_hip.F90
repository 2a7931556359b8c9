! qg_cabaret_hip -- drop-in replacement of the reference's CABARET time-step
! routines and the rel / zeta boundary routines of qg_aux (qg3p_aux.F90:32-39,
! 643-1642) on one MI355X, over the C ABI of libpequod_hip.so.  It exports the
! same names with the same arguments, so that qg3.F90's calls
! (qg3.F90:598-807, 952) link against it unchanged:
!
!   use qg_cabaret_hip, only : cabaret_adapt_dt, cabaret_cfl, &
!     & cabaret_predictor, cabaret_beta, cabaret_forcing, &
!     & cabaret_cell_boundaries, cabaret_extrapolator, cabaret_flux, &
!     & cabaret_corrector, rel_from_zeta_and_psi, update_zeta_bc
!
! Arrays follow the library's host-pointer convention: an array registered
! with pq_mirror_register (its data array, from its first element) is used
! in place on the device -- the state stays resident and the host pulls it
! when it needs it; an unregistered array is copied to the device before the
! call and, when the routine writes it, back afterwards.  The kernels evaluate
! each routine's statements in the reference's order without FP contraction,
! so the results equal the reference's bit for bit.  Errors abort through the
! reference's pabort (qg3p_base.F90:383-403).  One domain (decomp%n == 1),
! closed basin, WIND_FORCING (config.h), no BACKGROUND_VSHEAR.
module qg_cabaret_hip

  use iso_c_binding
  use pequod_hip_c
  use qg_parallel, only : decomposition, parallel_data, parallel_flux_data, pabort
  use qg_diagnostics, only : stratification, velocity_boundary_condition

  implicit none

  private

  integer, parameter :: real_8 = kind(1.0D0)

  public :: cabaret_adapt_dt, cabaret_cfl, cabaret_predictor, cabaret_beta, &
    & cabaret_forcing, cabaret_cell_boundaries, cabaret_extrapolator, &
    & cabaret_flux, cabaret_corrector
  public :: rel_from_zeta_and_psi, update_zeta_bc

  interface cabaret_cell_boundaries
    module procedure cabaret_cell_boundaries_velocity, &
      & cabaret_cell_boundaries_time_integrated
  end interface cabaret_cell_boundaries

  ! One array argument of a call: the device address the kernel uses and,
  ! for an unregistered host array, the temporary device copy.
  type dev_arg
    type(c_ptr) :: host = c_null_ptr
    type(c_ptr) :: d = c_null_ptr
    type(c_ptr) :: tmp = c_null_ptr
    integer(c_size_t) :: bytes = 0
  end type dev_arg

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len = *), intent(in) :: what

    if (rc /= PQ_OK) call pabort("qg_cabaret_hip: " // what // ": " // pq_error_message(), &
      & __FILE__, __LINE__)
  end subroutine check

  ! The device address of x(1..n): its mirror, or a temporary copy.
  subroutine arg_open(a, x, n)
    type(dev_arg), intent(out) :: a
    real(kind = real_8), dimension(*), target, intent(in) :: x
    integer, intent(in) :: n

    a%host = c_loc(x)
    a%bytes = int(n, c_size_t) * 8_c_size_t
    a%d = pq_mirror_device_ptr(a%host)
    if (.not. c_associated(a%d)) then
      call check(pq_malloc(a%tmp, a%bytes), "device copy")
      call check(pq_memcpy_h2d(a%tmp, a%host, a%bytes), "copy to the device")
      a%d = a%tmp
    end if
  end subroutine arg_open

  ! After the call: written arrays go back to their host copies (temporary
  ! copies only; a mirror stays on the device), and temporaries are freed.
  subroutine arg_close(a, written)
    type(dev_arg), intent(inout) :: a
    logical, intent(in) :: written

    if (c_associated(a%tmp)) then
      if (written) call check(pq_memcpy_d2h(a%host, a%tmp, a%bytes), "copy to the host")
      call check(pq_free(a%tmp), "free")
      a%tmp = c_null_ptr
    end if
  end subroutine arg_close

  subroutine one_domain(decomp)
    type(decomposition), intent(in) :: decomp

    if (decomp%n /= 1) call pabort("qg_cabaret_hip: one domain only", __FILE__, __LINE__)
    if (decomp%periodic_j) call pabort("qg_cabaret_hip: closed basin only", __FILE__, __LINE__)
  end subroutine one_domain

  ! cabaret_cfl (qg3p_aux.F90:920-943)
  function cabaret_cfl(dt, ui, uj) result(cfl)
    real(kind = real_8), intent(in) :: dt
    type(parallel_flux_data), intent(in) :: ui
    type(parallel_flux_data), intent(in) :: uj
    real(kind = real_8) :: cfl

    type(dev_arg) :: a_ui, a_uj
    real(c_double) :: c

    if (.not. dt > 0.0D0) call pabort("qg_cabaret_hip: cabaret_cfl needs dt > 0", __FILE__, __LINE__)
    call one_domain(ui%decomp)
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_uj, uj%data, size(uj%data))
    call check(pq_cabaret_cfl_dev(int(ui%decomp%ii, c_int), int(ui%decomp%layers, c_int), dt, &
      & a_ui%d, a_uj%d, c, c_null_ptr), "cabaret_cfl")
    call arg_close(a_ui, .false.)
    call arg_close(a_uj, .false.)
    cfl = c
  end function cabaret_cfl

  ! cabaret_adapt_dt (qg3p_aux.F90:946-990): dt from the target Courant number,
  ! at most max_dt_increase times the old dt and at most max_dt, unchanged
  ! when max |u| is zero to within ten spacings; then the time weights.
  subroutine cabaret_adapt_dt(target_cfl, max_dt, max_dt_increase, ui, uj, dt, tweight_a, &
    & tweight_u)
    real(kind = real_8), intent(in) :: target_cfl
    real(kind = real_8), intent(in) :: max_dt
    real(kind = real_8), intent(in) :: max_dt_increase
    type(parallel_flux_data), intent(in) :: ui
    type(parallel_flux_data), intent(in) :: uj
    real(kind = real_8), intent(inout) :: dt
    real(kind = real_8), intent(out) :: tweight_a
    real(kind = real_8), intent(out) :: tweight_u

    real(kind = real_8) :: old_dt, max_u

    old_dt = dt
    max_u = cabaret_cfl(1.0D0, ui, uj)
    if (abs(max_u - 0.0D0) > 10.0D0 * max(spacing(max_u), epsilon(0.0D0))) then
      dt = min(target_cfl / max_u, old_dt * max_dt_increase, max_dt)
    end if
    tweight_a = 0.5D0 * dt / old_dt
    tweight_u = dt / (old_dt + dt)
  end subroutine cabaret_adapt_dt

  ! cabaret_predictor (qg3p_aux.F90:993-1020)
  subroutine cabaret_predictor(dt, zeta_old, res, zeta)
    real(kind = real_8), intent(in) :: dt
    real(kind = real_8), dimension(:, :, :), contiguous, target, intent(in) :: zeta_old
    real(kind = real_8), dimension(:, :, :), contiguous, target, intent(in) :: res
    type(parallel_data), intent(inout) :: zeta

    type(dev_arg) :: a_zo, a_res, a_z

    call one_domain(zeta%decomp)
    call arg_open(a_zo, zeta_old, size(zeta_old))
    call arg_open(a_res, res, size(res))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call check(pq_cabaret_predictor_dev(int(zeta%decomp%ii, c_int), int(zeta%decomp%layers, c_int), &
      & dt, a_zo%d, a_res%d, a_z%d, c_null_ptr), "cabaret_predictor")
    call arg_close(a_zo, .false.)
    call arg_close(a_res, .false.)
    call arg_close(a_z, .true.)
  end subroutine cabaret_predictor

  ! cabaret_beta (qg3p_aux.F90:1023-1061)
  subroutine cabaret_beta(dt, tweight_a, beta, ui, zeta, res_beta_o)
    real(kind = real_8), intent(in) :: dt
    real(kind = real_8), intent(in) :: tweight_a
    type(parallel_data), intent(inout) :: zeta
    real(kind = real_8), dimension(zeta%decomp%layers), target, intent(in) :: beta
    type(parallel_flux_data), intent(in) :: ui
    real(kind = real_8), dimension(:, :, :), contiguous, target, intent(inout) :: res_beta_o

    type(dev_arg) :: a_b, a_ui, a_z, a_rb

    call one_domain(zeta%decomp)
    call arg_open(a_b, beta, size(beta))
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call arg_open(a_rb, res_beta_o, size(res_beta_o))
    call check(pq_cabaret_beta_dev(int(zeta%decomp%ii, c_int), int(zeta%decomp%layers, c_int), &
      & dt, tweight_a, a_b%d, a_ui%d, a_z%d, a_rb%d, c_null_ptr), "cabaret_beta")
    call arg_close(a_b, .false.)
    call arg_close(a_ui, .false.)
    call arg_close(a_z, .true.)
    call arg_close(a_rb, .true.)
  end subroutine cabaret_beta

  ! cabaret_forcing with WIND_FORCING (qg3p_aux.F90:1065-1164)
  subroutine cabaret_forcing(dt, force_wind, viscosity, visc_bot, rel_p, zeta)
    real(kind = real_8), intent(in) :: dt
    real(kind = real_8), dimension(:, :), contiguous, target, intent(in) :: force_wind
    real(kind = real_8), intent(in) :: viscosity
    real(kind = real_8), intent(in) :: visc_bot
    type(parallel_data), intent(in) :: rel_p
    type(parallel_data), intent(inout) :: zeta

    type(dev_arg) :: a_fw, a_rel, a_z

    call one_domain(zeta%decomp)
    call arg_open(a_fw, force_wind, size(force_wind))
    call arg_open(a_rel, rel_p%data, size(rel_p%data))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call check(pq_cabaret_forcing_dev(int(zeta%decomp%ii, c_int), int(zeta%decomp%layers, c_int), &
      & dt, a_fw%d, viscosity, visc_bot, a_rel%d, a_z%d, c_null_ptr), "cabaret_forcing")
    call arg_close(a_fw, .false.)
    call arg_close(a_rel, .false.)
    call arg_close(a_z, .true.)
  end subroutine cabaret_forcing

  ! cabaret_cell_boundaries_velocity (qg3p_aux.F90:1168-1231)
  subroutine cabaret_cell_boundaries_velocity(psi, ui, uj)
    type(parallel_data), intent(in) :: psi
    type(parallel_flux_data), intent(inout) :: ui
    type(parallel_flux_data), intent(inout) :: uj

    type(dev_arg) :: a_psi, a_ui, a_uj

    call one_domain(psi%decomp)
    call arg_open(a_psi, psi%data, size(psi%data))
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_uj, uj%data, size(uj%data))
    call check(pq_cabaret_cell_boundaries_velocity_dev(int(psi%decomp%ii, c_int), &
      & int(psi%decomp%layers, c_int), a_psi%d, a_ui%d, a_uj%d, c_null_ptr), &
      & "cabaret_cell_boundaries_velocity")
    call arg_close(a_psi, .false.)
    call arg_close(a_ui, .true.)
    call arg_close(a_uj, .true.)
  end subroutine cabaret_cell_boundaries_velocity

  ! cabaret_cell_boundaries_time_integrated (qg3p_aux.F90:1235-1313)
  subroutine cabaret_cell_boundaries_time_integrated(tweight_u, psi, ui, uj, uimo, ujmo)
    real(kind = real_8), intent(in) :: tweight_u
    type(parallel_data), intent(in) :: psi
    type(parallel_flux_data), intent(inout) :: ui
    type(parallel_flux_data), intent(inout) :: uj
    type(parallel_flux_data), intent(inout) :: uimo
    type(parallel_flux_data), intent(inout) :: ujmo

    type(dev_arg) :: a_psi, a_ui, a_uj, a_uimo, a_ujmo

    call one_domain(psi%decomp)
    call arg_open(a_psi, psi%data, size(psi%data))
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_uj, uj%data, size(uj%data))
    call arg_open(a_uimo, uimo%data, size(uimo%data))
    call arg_open(a_ujmo, ujmo%data, size(ujmo%data))
    call check(pq_cabaret_cell_boundaries_dev(int(psi%decomp%ii, c_int), &
      & int(psi%decomp%layers, c_int), tweight_u, a_psi%d, a_ui%d, a_uj%d, a_uimo%d, a_ujmo%d, &
      & c_null_ptr), "cabaret_cell_boundaries_time_integrated")
    call arg_close(a_psi, .false.)
    call arg_close(a_ui, .true.)
    call arg_close(a_uj, .true.)
    call arg_close(a_uimo, .true.)
    call arg_close(a_ujmo, .true.)
  end subroutine cabaret_cell_boundaries_time_integrated

  ! cabaret_extrapolator with CABARET_LIMITER (qg3p_aux.F90:1317-1544)
  subroutine cabaret_extrapolator(dt, zeta, zeta_old, ui, uj, zeta_fi, zeta_fj)
    real(kind = real_8), intent(in) :: dt
    type(parallel_data), intent(in) :: zeta
    real(kind = real_8), dimension(:, :, :), contiguous, target, intent(in) :: zeta_old
    type(parallel_flux_data), intent(in) :: ui
    type(parallel_flux_data), intent(in) :: uj
    type(parallel_flux_data), intent(inout) :: zeta_fi
    type(parallel_flux_data), intent(inout) :: zeta_fj

    type(dev_arg) :: a_z, a_zo, a_ui, a_uj, a_fi, a_fj

    call one_domain(zeta%decomp)
    call arg_open(a_z, zeta%data, size(zeta%data))
    call arg_open(a_zo, zeta_old, size(zeta_old))
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_uj, uj%data, size(uj%data))
    call arg_open(a_fi, zeta_fi%data, size(zeta_fi%data))
    call arg_open(a_fj, zeta_fj%data, size(zeta_fj%data))
    call check(pq_cabaret_extrapolator_dev(int(zeta%decomp%ii, c_int), &
      & int(zeta%decomp%layers, c_int), dt, a_z%d, a_zo%d, a_ui%d, a_uj%d, a_fi%d, a_fj%d, &
      & c_null_ptr, c_null_ptr), "cabaret_extrapolator")
    call arg_close(a_z, .false.)
    call arg_close(a_zo, .false.)
    call arg_close(a_ui, .false.)
    call arg_close(a_uj, .false.)
    call arg_close(a_fi, .true.)
    call arg_close(a_fj, .true.)
  end subroutine cabaret_extrapolator

  ! cabaret_flux (qg3p_aux.F90:1547-1594)
  subroutine cabaret_flux(ui, uj, zeta_fi, zeta_fj, flux_i, flux_j)
    type(parallel_flux_data), intent(in) :: ui
    type(parallel_flux_data), intent(in) :: uj
    type(parallel_flux_data), intent(in) :: zeta_fi
    type(parallel_flux_data), intent(in) :: zeta_fj
    type(parallel_flux_data), intent(inout) :: flux_i
    type(parallel_flux_data), intent(inout) :: flux_j

    type(dev_arg) :: a_ui, a_uj, a_fi, a_fj, a_xi, a_xj

    call one_domain(ui%decomp)
    call arg_open(a_ui, ui%data, size(ui%data))
    call arg_open(a_uj, uj%data, size(uj%data))
    call arg_open(a_fi, zeta_fi%data, size(zeta_fi%data))
    call arg_open(a_fj, zeta_fj%data, size(zeta_fj%data))
    call arg_open(a_xi, flux_i%data, size(flux_i%data))
    call arg_open(a_xj, flux_j%data, size(flux_j%data))
    call check(pq_cabaret_flux_dev(int(ui%decomp%ii, c_int), int(ui%decomp%layers, c_int), &
      & a_ui%d, a_uj%d, a_fi%d, a_fj%d, a_xi%d, a_xj%d, c_null_ptr), "cabaret_flux")
    call arg_close(a_ui, .false.)
    call arg_close(a_uj, .false.)
    call arg_close(a_fi, .false.)
    call arg_close(a_fj, .false.)
    call arg_close(a_xi, .true.)
    call arg_close(a_xj, .true.)
  end subroutine cabaret_flux

  ! cabaret_corrector (qg3p_aux.F90:1597-1642)
  subroutine cabaret_corrector(dt, flux_i, flux_j, res, zeta)
    real(kind = real_8), intent(in) :: dt
    type(parallel_data), intent(inout) :: zeta
    type(parallel_flux_data), intent(in) :: flux_i
    type(parallel_flux_data), intent(in) :: flux_j
    real(kind = real_8), dimension(:, :, :), contiguous, target, intent(inout) :: res

    type(dev_arg) :: a_xi, a_xj, a_res, a_z

    call one_domain(zeta%decomp)
    call arg_open(a_xi, flux_i%data, size(flux_i%data))
    call arg_open(a_xj, flux_j%data, size(flux_j%data))
    call arg_open(a_res, res, size(res))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call check(pq_cabaret_corrector_dev(int(zeta%decomp%ii, c_int), int(zeta%decomp%layers, c_int), &
      & dt, a_xi%d, a_xj%d, a_res%d, a_z%d, c_null_ptr), "cabaret_corrector")
    call arg_close(a_xi, .false.)
    call arg_close(a_xj, .false.)
    call arg_close(a_res, .true.)
    call arg_close(a_z, .true.)
  end subroutine cabaret_corrector

  ! rel_from_zeta_and_psi with apply_rel_bc (qg3p_aux.F90:738-915)
  subroutine rel_from_zeta_and_psi(strat, vbc, zeta, psi, rel)
    type(stratification), intent(in) :: strat
    type(velocity_boundary_condition), intent(in) :: vbc
    type(parallel_data), intent(in) :: zeta
    type(parallel_data), intent(in) :: psi
    type(parallel_data), intent(inout) :: rel

    type(dev_arg) :: a_s1, a_s2, a_z, a_psi, a_rel

    call one_domain(zeta%decomp)
    call arg_open(a_s1, strat%s1, size(strat%s1))
    call arg_open(a_s2, strat%s2, size(strat%s2))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call arg_open(a_psi, psi%data, size(psi%data))
    call arg_open(a_rel, rel%data, size(rel%data))
    call check(pq_rel_from_zeta_and_psi_dev(int(zeta%decomp%ii, c_int), &
      & int(zeta%decomp%layers, c_int), a_s1%d, a_s2%d, int(vbc%bc_type, c_int), vbc%alpha, &
      & a_z%d, a_psi%d, a_rel%d, c_null_ptr), "rel_from_zeta_and_psi")
    call arg_close(a_s1, .false.)
    call arg_close(a_s2, .false.)
    call arg_close(a_z, .false.)
    call arg_close(a_psi, .false.)
    call arg_close(a_rel, .true.)
  end subroutine rel_from_zeta_and_psi

  ! update_zeta_bc (qg3p_aux.F90:643-734)
  subroutine update_zeta_bc(strat, rel, psi, zeta)
    type(stratification), intent(in) :: strat
    type(parallel_data), intent(in) :: rel
    type(parallel_data), intent(in) :: psi
    type(parallel_data), intent(inout) :: zeta

    type(dev_arg) :: a_s1, a_s2, a_rel, a_psi, a_z

    call one_domain(zeta%decomp)
    call arg_open(a_s1, strat%s1, size(strat%s1))
    call arg_open(a_s2, strat%s2, size(strat%s2))
    call arg_open(a_rel, rel%data, size(rel%data))
    call arg_open(a_psi, psi%data, size(psi%data))
    call arg_open(a_z, zeta%data, size(zeta%data))
    call check(pq_update_zeta_bc_dev(int(zeta%decomp%ii, c_int), int(zeta%decomp%layers, c_int), &
      & a_s1%d, a_s2%d, a_rel%d, a_psi%d, a_z%d, c_null_ptr), "update_zeta_bc")
    call arg_close(a_s1, .false.)
    call arg_close(a_s2, .false.)
    call arg_close(a_rel, .false.)
    call arg_close(a_psi, .false.)
    call arg_close(a_z, .true.)
  end subroutine update_zeta_bc

end module qg_cabaret_hip
