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
