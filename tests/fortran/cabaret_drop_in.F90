! Exercises the Fortran CABARET drop-in (qg_cabaret_hip over libpequod_hip.so)
! with the reference's call sequence of one time step (qg3.F90:598-952,
! without the PV inversion: psi is an input).
!   cabaret_drop_in copy     IN OUT : host arrays, copied to the device and
!                                     back by every call (unregistered);
!   cabaret_drop_in resident IN OUT : every array registered as a device
!                                     mirror first, so the calls work in
!                                     place; pulled back at the end.
! IN holds, as raw little-endian values: g, layers, bc_type (int32); dt,
! target_cfl, max_dt, max_dt_increase, viscosity, visc_bot, alpha (float64);
! s1, s2, beta (layers); force_wind (g, g); zeta_old, zeta, psi, rel (g+2,
! g+2, layers); res, res_beta (g, g, layers); ui, uimo,
! zeta_fi, flux_i (g+3, g+2, layers); uj, ujmo, zeta_fj, flux_j (g+2, g+3,
! layers).  OUT receives dt, tweight_a, tweight_u, cfl, then zeta, rel, res,
! res_beta, ui, uj, uimo, ujmo, zeta_fi, zeta_fj, flux_i, flux_j.
program cabaret_drop_in

  use iso_c_binding
  use pequod_hip_c
  use qg_parallel, only : decomposition, parallel_data, parallel_flux_data, FLUX_II, FLUX_JJ
  use qg_diagnostics, only : stratification, velocity_boundary_condition
  use qg_cabaret_hip

  implicit none

  integer, parameter :: r8 = kind(1.0D0)
  character(len = 16) :: mode
  character(len = 512) :: fin, fout
  integer(c_int) :: g, nl, bc
  real(r8) :: dt, target_cfl, max_dt, max_dt_increase, visc, visc_bot, alpha
  real(r8) :: tw_a, tw_u, cfl
  type(decomposition), target :: decomp
  type(stratification) :: strat
  type(velocity_boundary_condition) :: vbc
  real(r8), dimension(:), pointer :: beta
  real(r8), dimension(:, :), allocatable, target :: fw
  real(r8), dimension(:, :, :), allocatable, target :: zeta_old, res, res_beta
  type(parallel_data) :: zeta, psi, rel
  type(parallel_flux_data) :: ui, uj, uimo, ujmo, zfi, zfj, fli, flj
  integer :: u
  logical :: resident

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  resident = trim(mode) == "resident"

  open(newunit = u, file = trim(fin), access = "stream", form = "unformatted", status = "old")
  read(u) g, nl, bc
  read(u) dt, target_cfl, max_dt, max_dt_increase, visc, visc_bot, alpha
  decomp%ii = g
  decomp%jj = g
  decomp%layers = nl
  strat%layers = nl
  allocate(strat%s1(nl), strat%s2(nl), beta(nl), fw(g, g))
  read(u) strat%s1, strat%s2, beta, fw
  vbc%bc_type = bc
  vbc%alpha = alpha
  allocate(zeta_old(g + 2, g + 2, nl), res(g, g, nl), res_beta(g, g, nl))
  call node(zeta)
  call node(psi)
  call node(rel)
  read(u) zeta_old, zeta%data, psi%data, rel%data, res, res_beta
  call face(ui, FLUX_II)
  call face(uimo, FLUX_II)
  call face(zfi, FLUX_II)
  call face(fli, FLUX_II)
  call face(uj, FLUX_JJ)
  call face(ujmo, FLUX_JJ)
  call face(zfj, FLUX_JJ)
  call face(flj, FLUX_JJ)
  read(u) ui%data, uimo%data, zfi%data, fli%data, uj%data, ujmo%data, zfj%data, flj%data
  close(u)

  if (resident) then
    call reg(zeta_old, size(zeta_old))
    call reg(res, size(res))
    call reg(res_beta, size(res_beta))
    call reg(zeta%data, size(zeta%data))
    call reg(psi%data, size(psi%data))
    call reg(rel%data, size(rel%data))
    call reg(ui%data, size(ui%data))
    call reg(uj%data, size(uj%data))
    call reg(uimo%data, size(uimo%data))
    call reg(ujmo%data, size(ujmo%data))
    call reg(zfi%data, size(zfi%data))
    call reg(zfj%data, size(zfj%data))
    call reg(fli%data, size(fli%data))
    call reg(flj%data, size(flj%data))
  end if

  ! one step in the reference's order (qg3.F90:598-952)
  call cabaret_adapt_dt(target_cfl, max_dt, max_dt_increase, ui, uj, dt, tw_a, tw_u)
  call cabaret_predictor(dt, zeta_old, res, zeta)
  call cabaret_beta(dt, tw_a, beta, ui, zeta, res_beta)
  call rel_from_zeta_and_psi(strat, vbc, zeta, psi, rel)
  call update_zeta_bc(strat, rel, psi, zeta)
  cfl = cabaret_cfl(dt, ui, uj)
  call cabaret_forcing(dt, fw, visc, visc_bot, rel, zeta)
  call cabaret_cell_boundaries(tw_u, psi, ui, uj, uimo, ujmo)
  call cabaret_extrapolator(dt, zeta, zeta_old, ui, uj, zfi, zfj)
  call cabaret_flux(ui, uj, zfi, zfj, fli, flj)
  call cabaret_corrector(dt, fli, flj, res, zeta)

  if (resident) then
    call pull(zeta%data)
    call pull(rel%data)
    call pull(res)
    call pull(res_beta)
    call pull(ui%data)
    call pull(uj%data)
    call pull(uimo%data)
    call pull(ujmo%data)
    call pull(zfi%data)
    call pull(zfj%data)
    call pull(fli%data)
    call pull(flj%data)
  end if

  open(newunit = u, file = trim(fout), access = "stream", form = "unformatted", &
    & status = "replace")
  write(u) dt, tw_a, tw_u, cfl
  write(u) zeta%data, rel%data, res, res_beta, ui%data, uj%data, uimo%data, ujmo%data, &
    & zfi%data, zfj%data, fli%data, flj%data
  close(u)
  write(*, "(a)") "cabaret drop-in ok"

contains

  subroutine node(p)
    type(parallel_data), intent(out) :: p

    p%decomp => decomp
    allocate(p%data(g + 2, g + 2, nl))
  end subroutine node

  subroutine face(f, dir)
    type(parallel_flux_data), intent(out) :: f
    integer, intent(in) :: dir

    f%decomp => decomp
    f%flux_dir = dir
    if (dir == FLUX_II) then
      allocate(f%data(g + 3, g + 2, nl))
    else
      allocate(f%data(g + 2, g + 3, nl))
    end if
  end subroutine face

  subroutine reg(x, n)
    real(r8), dimension(*), target, intent(in) :: x
    integer, intent(in) :: n

    if (pq_mirror_register(c_loc(x), int(n, c_size_t) * 8_c_size_t) /= PQ_OK) stop 4
  end subroutine reg

  subroutine pull(x)
    real(r8), dimension(*), target, intent(inout) :: x

    if (pq_mirror_pull(c_loc(x)) /= PQ_OK) stop 5
  end subroutine pull

end program cabaret_drop_in
