#include "config.h"

! Fixture driver for the reference's CABARET time step.  Built in this
! container only (make_reference_fixtures.py) against the reference's own
! modules, unchanged; it calls their routines in the order of the driver's
! time loop (qg3.F90:393-459 for the START_PERTURB start, 590-967 for a
! step) on one serial domain, starting from a given PV field instead of
! init_rand, and writes the state after chosen steps.
!
! Input (stream, native endian): int32 g, nsteps, adapt, bc_type, nout,
! out_steps(nout); float64 dt, beta(L), viscosity, visc_bot, alpha,
! target_cfl, max_dt, max_dt_increase, h(L), s1(L-1), s2(L-1),
! force_wind(g, g), zeta0(g, g, L) (the sdata interior field; its boundary
! is set to zero as the reference's start does).
! Output: per requested step, int32 step, float64 dt, tweight_a,
! tweight_u, model_time, then psi, zeta, rel, zeta_old (data arrays,
! (g+2, g+2, L)), zeta_fi, ui, uimo, flux_i ((g+3, g+2, L)), zeta_fj, uj,
! ujmo, flux_j ((g+2, g+3, L)), res, res_beta ((g, g, L)); and a second
! file with dt after every step.
program ref_driver

  use qg_aux
  use qg_diagnostics
  use qg_helmholtz
  use qg_laplace_bc
  use qg_parallel
  use qg_parameters, only : ii_m, jj_m, layers, real_8

  implicit none

  character(len = 1024) :: fin, fout, fdt
  integer :: g, nsteps, adapt, bc_type, nout, step, k, u_in, u_out, u_dt
  integer, allocatable :: out_steps(:)
  real(kind = real_8) :: dt, viscosity, visc_bot, alpha, target_cfl, max_dt, &
    & max_dt_increase, model_time, tweight_a, tweight_u
  real(kind = real_8), dimension(layers) :: beta, h
  real(kind = real_8), dimension(layers - 1) :: s1, s2
  real(kind = real_8), allocatable :: zeta0(:, :, :)
  real(kind = real_8), dimension(ii_m, jj_m) :: force_wind
  real(kind = real_8), dimension(ii_m, jj_m, layers) :: res, res_beta
  real(kind = real_8), dimension(ii_m + 2, jj_m + 2, layers) :: zeta_old

  type(decomposition) :: decomp
  type(stratification) :: strat
  type(velocity_boundary_condition) :: vbc
  type(helmholtz_cache) :: hcache
  type(parallel_data) :: psi, rel, zeta
  type(parallel_flux_data) :: zeta_fi, zeta_fj, ui, uj, uimo, ujmo, flux_i, flux_j

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, fdt)

  open(newunit = u_in, file = trim(fin), access = "stream", form = "unformatted", status = "old")
  read(u_in) g, nsteps, adapt, bc_type, nout
  allocate(out_steps(nout))
  read(u_in) out_steps
  read(u_in) dt, beta, viscosity, visc_bot, alpha, target_cfl, max_dt, max_dt_increase
  read(u_in) h, s1, s2
  force_wind = 0.0D0
  read(u_in) force_wind(:g, :g)
  allocate(zeta0(g, g, layers))
  read(u_in) zeta0
  close(u_in)

  call parallel_init()
  call allocate(decomp, iis = (/g/), jjs = (/g/), layers = layers, periodic_j = .false.)
  call allocate(psi, decomp)
  call allocate(rel, decomp)
  call allocate(zeta, decomp)
  call allocate(zeta_fi, decomp, FLUX_II)
  call allocate(zeta_fj, decomp, FLUX_JJ)
  call allocate(ui, decomp, FLUX_II)
  call allocate(uj, decomp, FLUX_JJ)
  call allocate(uimo, decomp, FLUX_II)
  call allocate(ujmo, decomp, FLUX_JJ)
  call allocate(flux_i, decomp, FLUX_II)
  call allocate(flux_j, decomp, FLUX_JJ)

  call allocate(strat, h, s1, s2)
  select case(bc_type)
    case(BC_PARTIAL_SLIP)
      call initialise(vbc, BC_PARTIAL_SLIP, alpha = alpha)
    case default
      call initialise(vbc, bc_type)
  end select
  call allocate(hcache, decomp, strat, vbc)

  ! START_PERTURB with the given field in place of init_rand
  call zero(zeta)
  zeta%sdata(:, :, :) = zeta0
  call apply_dirichlet_bc(0.0D0, zeta)
  call pv_inversion(strat, vbc, zeta, hcache, psi)
  call rel_from_zeta_and_psi(strat, vbc, zeta, psi, rel)
  call update_zeta_bc(strat, rel, psi, zeta)
  call zero(zeta_fi)
  call zero(zeta_fj)
  call zero(ui)
  call zero(uj)
  call zero(uimo)
  call zero(ujmo)
  call zero(flux_i)
  call zero(flux_j)
  res = 0.0D0
  res_beta = 0.0D0
  zeta_old = 0.0D0
  call set(zeta_old, zeta)
  model_time = dt / 2.0D0
  tweight_a = 0.5D0
  tweight_u = 0.5D0

  open(newunit = u_out, file = trim(fout), access = "stream", form = "unformatted", status = "replace")
  open(newunit = u_dt, file = trim(fdt), access = "stream", form = "unformatted", status = "replace")
  if(any(out_steps == 0)) call dump(0)

  do step = 1, nsteps
    if(adapt /= 0) then
      call cabaret_adapt_dt(target_cfl, max_dt, max_dt_increase, ui, uj, dt, tweight_a, tweight_u)
    end if
    call cabaret_predictor(dt, zeta_old, res, zeta)
    call cabaret_beta(dt, tweight_a, beta, ui, zeta, res_beta)
    call pv_inversion(strat, vbc, zeta, hcache, psi)
    call rel_from_zeta_and_psi(strat, vbc, zeta, psi, rel)
    call update_zeta_bc(strat, rel, psi, zeta)
    model_time = model_time + 0.5D0 * dt
    call cabaret_forcing(dt, force_wind, viscosity, visc_bot, rel, zeta)
    call cabaret_cell_boundaries(tweight_u, psi, ui, uj, uimo, ujmo)
    call cabaret_extrapolator(dt, zeta, zeta_old, ui, uj, zeta_fi, zeta_fj)
    call cabaret_flux(ui, uj, zeta_fi, zeta_fj, flux_i, flux_j)
    call cabaret_corrector(dt, flux_i, flux_j, res, zeta)
    call set(zeta_old, zeta)
    model_time = model_time + 0.5D0 * dt
    write(u_dt) dt
    if(any(out_steps == step)) call dump(step)
  end do
  close(u_out)
  close(u_dt)
  call parallel_finalise()

contains

  subroutine dump(s)
    integer, intent(in) :: s
    write(u_out) s
    write(u_out) dt, tweight_a, tweight_u, model_time
    write(u_out) psi%data, zeta%data, rel%data, zeta_old(:g + 2, :g + 2, :)
    write(u_out) zeta_fi%data, ui%data, uimo%data, flux_i%data
    write(u_out) zeta_fj%data, uj%data, ujmo%data, flux_j%data
    write(u_out) res(:g, :g, :), res_beta(:g, :g, :)
  end subroutine dump

end program ref_driver
