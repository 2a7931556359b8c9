! pequod_hip_c.F90 -- iso_c_binding interfaces of libpequod_hip.so (the C ABI
! declared in include/pequod_hip.h), for a Fortran host.
!
! Arrays are passed as the reference stores them (column-major, i fastest);
! a g x g view of an sdata slice is (first element, ld1, ld2).  Every function
! returns the library status (PQ_OK = 0); pq_error_message() returns the
! calling thread's last error text.
module pequod_hip_c

  use iso_c_binding

  implicit none

  private

  integer(c_int), parameter, public :: PQ_OK = 0, PQ_ERR_ARG = 1, &
    & PQ_ERR_HIP = 2, PQ_ERR_NAN = 3, PQ_WARN_LARGE = 4, PQ_ERR_NOMEM = 5
  integer(c_int), parameter, public :: PQ_PREC_FP64 = 0, PQ_PREC_FP32 = 1, &
    & PQ_PREC_MIXED = 2

  public :: pq_error_message
  public :: pq_helm_create, pq_helm_destroy, pq_helm_solve
  public :: pq_inv_create, pq_inv_destroy, pq_inv_run, pq_inv_status
  public :: pq_mirror_register, pq_mirror_unregister, pq_mirror_push, &
    & pq_mirror_pull
  public :: pq_model_stream
  public :: pq_mirror_device_ptr, pq_malloc, pq_free, pq_memcpy_h2d, pq_memcpy_d2h
  public :: pq_cabaret_cfl_dev, pq_cabaret_predictor_dev, pq_cabaret_beta_dev, &
    & pq_cabaret_forcing_dev, pq_cabaret_cell_boundaries_dev, &
    & pq_cabaret_cell_boundaries_velocity_dev, pq_cabaret_extrapolator_dev, &
    & pq_cabaret_flux_dev, pq_cabaret_corrector_dev, pq_rel_from_zeta_and_psi_dev, &
    & pq_update_zeta_bc_dev

  interface
    function pq_last_error() bind(C, name = "pq_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function pq_last_error

    ! qg3p_helmholtz.F90:292-295 (allocate), 614 (helm), 389 (deallocate)
    function pq_helm_create(h, ii, jj, layers, ss, precision) &
      & bind(C, name = "pq_helm_create") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: ii, jj, layers, precision
      real(c_double), dimension(*), intent(in) :: ss
      integer(c_int) :: rc
    end function pq_helm_create

    function pq_helm_create_periodic(h, ii, jj, layers, ss, precision) &
      & bind(C, name = "pq_helm_create_periodic") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: ii, jj, layers, precision
      real(c_double), dimension(*), intent(in) :: ss
      integer(c_int) :: rc
    end function pq_helm_create_periodic

    function pq_helm_destroy(h) bind(C, name = "pq_helm_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function pq_helm_destroy

    function pq_helm_solve(h, zeta, ld1, ld2) &
      & bind(C, name = "pq_helm_solve") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*), intent(inout) :: zeta
      integer(c_int), value :: ld1, ld2
      integer(c_int) :: rc
    end function pq_helm_solve

    ! qg3p_aux.F90:152-193 (allocate_helmholtz_cache), 452-479 (pv_inversion)
    function pq_inv_create(h, g, layers, ss, modes_to_z, z_to_modes, &
      & precision, batch_max) bind(C, name = "pq_inv_create") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: g, layers, precision, batch_max
      real(c_double), dimension(*), intent(in) :: ss, modes_to_z, z_to_modes
      integer(c_int) :: rc
    end function pq_inv_create

    ! the periodic channel (qg3p_aux.F90:152-193, 233-379, 452-479, 524-575)
    function pq_chan_create(h, gi, gj, layers, ss, modes_to_z, z_to_modes, &
      & bc_type, alpha, precision, batch_max) bind(C, name = "pq_chan_create") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: gi, gj, layers, bc_type, precision, batch_max
      real(c_double), dimension(*), intent(in) :: ss, modes_to_z, z_to_modes
      real(c_double), value :: alpha
      integer(c_int) :: rc
    end function pq_chan_create

    function pq_inv_destroy(h) bind(C, name = "pq_inv_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function pq_inv_destroy

    function pq_inv_run(h, zeta, ldz1, ldz2, psi, ldp1, ldp2) &
      & bind(C, name = "pq_inv_run") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*), intent(in) :: zeta
      real(c_double), dimension(*), intent(inout) :: psi
      integer(c_int), value :: ldz1, ldz2, ldp1, ldp2
      integer(c_int) :: rc
    end function pq_inv_run

    function pq_inv_status(h, psic) bind(C, name = "pq_inv_status") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*), intent(out) :: psic
      integer(c_int) :: rc
    end function pq_inv_status

    ! Device mirrors: keep model state resident between calls.
    function pq_mirror_register(host, bytes) &
      & bind(C, name = "pq_mirror_register") result(rc)
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: host
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function pq_mirror_register

    function pq_mirror_unregister(host) &
      & bind(C, name = "pq_mirror_unregister") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: host
      integer(c_int) :: rc
    end function pq_mirror_unregister

    function pq_mirror_push(host) bind(C, name = "pq_mirror_push") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: host
      integer(c_int) :: rc
    end function pq_mirror_push

    function pq_mirror_pull(host) bind(C, name = "pq_mirror_pull") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: host
      integer(c_int) :: rc
    end function pq_mirror_pull

    ! qg3p_diagnostics.F90:778-863 (energy)
    function pq_diag_create(d, g, layers, h, s2) bind(C, name = "pq_diag_create") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: d
      integer(c_int), value :: g, layers
      real(c_double), dimension(*), intent(in) :: h, s2
      integer(c_int) :: rc
    end function pq_diag_create

    function pq_diag_destroy(d) bind(C, name = "pq_diag_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: d
      integer(c_int) :: rc
    end function pq_diag_destroy

    function pq_diag_energy(d, psi, ld1, ld2, ekin, epot) &
      & bind(C, name = "pq_diag_energy") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: d
      real(c_double), dimension(*), intent(in) :: psi
      integer(c_int), value :: ld1, ld2
      real(c_double), dimension(*), intent(out) :: ekin
      real(c_double), intent(out) :: epot
      integer(c_int) :: rc
    end function pq_diag_energy

    function pq_model_stream(m) bind(C, name = "pq_model_stream") result(s)
      import :: c_ptr
      type(c_ptr), value :: m
      type(c_ptr) :: s
    end function pq_model_stream

    ! device memory and mirrors
    function pq_mirror_device_ptr(host) bind(C, name = "pq_mirror_device_ptr") result(d)
      import :: c_ptr
      type(c_ptr), value :: host
      type(c_ptr) :: d
    end function pq_mirror_device_ptr

    function pq_malloc(d, bytes) bind(C, name = "pq_malloc") result(rc)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), intent(out) :: d
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function pq_malloc

    function pq_free(d) bind(C, name = "pq_free") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: d
      integer(c_int) :: rc
    end function pq_free

    function pq_memcpy_h2d(d, host, bytes) bind(C, name = "pq_memcpy_h2d") result(rc)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: d, host
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function pq_memcpy_h2d

    function pq_memcpy_d2h(host, d, bytes) bind(C, name = "pq_memcpy_d2h") result(rc)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: host, d
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function pq_memcpy_d2h

    ! CABARET and rel / zeta boundary stages on device arrays of the
    ! reference's shapes (include/pequod_hip.h)
    function pq_cabaret_cfl_dev(g, layers, dt, ui, uj, cfl, stream) &
      & bind(C, name = "pq_cabaret_cfl_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt
      type(c_ptr), value :: ui, uj, stream
      real(c_double), intent(out) :: cfl
      integer(c_int) :: rc
    end function pq_cabaret_cfl_dev

    function pq_cabaret_predictor_dev(g, layers, dt, zeta_old, res, zeta, stream) &
      & bind(C, name = "pq_cabaret_predictor_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt
      type(c_ptr), value :: zeta_old, res, zeta, stream
      integer(c_int) :: rc
    end function pq_cabaret_predictor_dev

    function pq_cabaret_beta_dev(g, layers, dt, tweight_a, beta, ui, zeta, res_beta, stream) &
      & bind(C, name = "pq_cabaret_beta_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt, tweight_a
      type(c_ptr), value :: beta, ui, zeta, res_beta, stream
      integer(c_int) :: rc
    end function pq_cabaret_beta_dev

    function pq_cabaret_forcing_dev(g, layers, dt, force_wind, viscosity, visc_bot, rel, zeta, &
      & stream) bind(C, name = "pq_cabaret_forcing_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt, viscosity, visc_bot
      type(c_ptr), value :: force_wind, rel, zeta, stream
      integer(c_int) :: rc
    end function pq_cabaret_forcing_dev

    function pq_cabaret_cell_boundaries_dev(g, layers, tweight_u, psi, ui, uj, uimo, ujmo, &
      & stream) bind(C, name = "pq_cabaret_cell_boundaries_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: tweight_u
      type(c_ptr), value :: psi, ui, uj, uimo, ujmo, stream
      integer(c_int) :: rc
    end function pq_cabaret_cell_boundaries_dev

    function pq_cabaret_cell_boundaries_velocity_dev(g, layers, psi, ui, uj, stream) &
      & bind(C, name = "pq_cabaret_cell_boundaries_velocity_dev") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: g, layers
      type(c_ptr), value :: psi, ui, uj, stream
      integer(c_int) :: rc
    end function pq_cabaret_cell_boundaries_velocity_dev

    function pq_cabaret_extrapolator_dev(g, layers, dt, zeta, zeta_old, ui, uj, zeta_fi, &
      & zeta_fj, scratch, stream) bind(C, name = "pq_cabaret_extrapolator_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt
      type(c_ptr), value :: zeta, zeta_old, ui, uj, zeta_fi, zeta_fj, scratch, stream
      integer(c_int) :: rc
    end function pq_cabaret_extrapolator_dev

    function pq_cabaret_flux_dev(g, layers, ui, uj, zeta_fi, zeta_fj, flux_i, flux_j, stream) &
      & bind(C, name = "pq_cabaret_flux_dev") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: g, layers
      type(c_ptr), value :: ui, uj, zeta_fi, zeta_fj, flux_i, flux_j, stream
      integer(c_int) :: rc
    end function pq_cabaret_flux_dev

    function pq_cabaret_corrector_dev(g, layers, dt, flux_i, flux_j, res, zeta, stream) &
      & bind(C, name = "pq_cabaret_corrector_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers
      real(c_double), value :: dt
      type(c_ptr), value :: flux_i, flux_j, res, zeta, stream
      integer(c_int) :: rc
    end function pq_cabaret_corrector_dev

    function pq_rel_from_zeta_and_psi_dev(g, layers, s1, s2, bc_type, alpha, zeta, psi, rel, &
      & stream) bind(C, name = "pq_rel_from_zeta_and_psi_dev") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g, layers, bc_type
      real(c_double), value :: alpha
      type(c_ptr), value :: s1, s2, zeta, psi, rel, stream
      integer(c_int) :: rc
    end function pq_rel_from_zeta_and_psi_dev

    function pq_update_zeta_bc_dev(g, layers, s1, s2, rel, psi, zeta, stream) &
      & bind(C, name = "pq_update_zeta_bc_dev") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: g, layers
      type(c_ptr), value :: s1, s2, rel, psi, zeta, stream
      integer(c_int) :: rc
    end function pq_update_zeta_bc_dev
  end interface

  interface
    function c_strlen(s) bind(C, name = "strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  ! The library's last error message for this thread.
  function pq_error_message() result(msg)
    character(len = :), allocatable :: msg

    type(c_ptr) :: p
    character(kind = c_char), dimension(:), pointer :: chars
    integer :: n, i

    p = pq_last_error()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len = n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function pq_error_message

end module pequod_hip_c
