! ISO_C_BINDING interface to the qgpu C ABI (include/qgpu.h), for qdyn's
! Fortran host.  Each routine binds one extern "C" entry point; arrays are
! qdyn's own x(:), v(:), d(:) (real(8), length 3*natom) passed by address.
module qgpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: QGPU_FP64 = 0, QGPU_MIXED = 1

  interface
    integer(c_int) function qgpu_create(input_files, n_instances, precision, handle) &
        bind(C, name='qgpu_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: input_files(*)
      integer(c_int), value :: n_instances, precision
      type(c_ptr), intent(out) :: handle
    end function

    integer(c_int) function qgpu_destroy(handle) bind(C, name='qgpu_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function qgpu_make_pair_lists(handle) bind(C, name='qgpu_make_pair_lists')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function qgpu_pot_energy(handle) bind(C, name='qgpu_pot_energy')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function qgpu_step(handle, nsteps) bind(C, name='qgpu_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function qgpu_get_state(handle, k, x, v) bind(C, name='qgpu_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(out) :: x(*), v(*)
    end function

    integer(c_int) function qgpu_set_state(handle, k, x, v) bind(C, name='qgpu_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(in) :: x(*), v(*)
    end function

    integer(c_int) function qgpu_get_gradient(handle, k, d) bind(C, name='qgpu_get_gradient')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(out) :: d(*)
    end function

    integer(c_int) function qgpu_energies(handle, k, e, cap) bind(C, name='qgpu_energies')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k, cap
      real(c_double), intent(out) :: e(*)
    end function

    integer(c_int) function qgpu_qenergies(handle, k, e, cap) bind(C, name='qgpu_qenergies')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k, cap
      real(c_double), intent(out) :: e(*)
    end function

    type(c_ptr) function qgpu_last_error() bind(C, name='qgpu_last_error')
      import :: c_ptr
    end function
  end interface

contains

  ! The last error message as a Fortran string (for die(msg)).
  function qgpu_error_message() result(msg)
    character(len=256) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ''
    p = qgpu_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

end module qgpu
