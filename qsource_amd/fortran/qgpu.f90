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

    ! pot_energy_nonbonds (src/md.f90:16261-16344) at x; d accumulates,
    ! e(7) = pp, pw, ww (el, vdW) and LRF, eq(4, nstates) = qp, qw (el, vdW)
    integer(c_int) function qgpu_pot_energy_nonbonds(handle, k, x, d, e, eq, eq_cap) &
        bind(C, name='qgpu_pot_energy_nonbonds')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k, eq_cap
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: d(*)
      real(c_double), intent(out) :: e(*), eq(*)
    end function

    ! shake(xx, x) (src/md.f90:17581-17679): x constrained in place
    integer(c_int) function qgpu_shake(handle, k, xx, x, niter) bind(C, name='qgpu_shake')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(in) :: xx(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), intent(out) :: niter
    end function

    ! failure record of instance k: bits (1 list regrowth, 2 shell capacity,
    ! 4 SHAKE, 8 NaN temperature) and the step it was found at
    integer(c_int) function qgpu_instance_status(handle, k, bits, at_step) &
        bind(C, name='qgpu_instance_status')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      integer(c_int), intent(out) :: bits, at_step
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

    integer(c_int) function qgpu_natom(handle, k, natom) bind(C, name='qgpu_natom')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      integer(c_int), intent(out) :: natom
    end function

    integer(c_int) function qgpu_temperature(handle, k, temp) bind(C, name='qgpu_temperature')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(out) :: temp
    end function

    integer(c_int) function qgpu_theta_corr(handle, k, tc, n) bind(C, name='qgpu_theta_corr')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k
      real(c_double), intent(out) :: tc(*)
      integer(c_int), intent(out) :: n
    end function

    ! energy file frames kept on the device (put_ene every iene steps)
    integer(c_int) function qgpu_set_energy_interval(handle, iene, max_frames) &
        bind(C, name='qgpu_set_energy_interval')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: iene, max_frames
    end function

    integer(c_int) function qgpu_energy_series(handle, k, out, cap, nframes) &
        bind(C, name='qgpu_energy_series')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: k, cap
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nframes
    end function

    ! monitored group pairs (nonbond_monitor): 3 + 2 nstates values per pair
    integer(c_int) function qgpu_monitor(h, k, out, cap, n) bind(C, name='qgpu_monitor')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: k, cap
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: n
    end function

    ! restart (write_xfin) and DCD trajectory (trj_create/trj_write) writers
    integer(c_int) function qgpu_write_restart(path, natom, x, v, nshell, theta_corr) &
        bind(C, name='qgpu_write_restart')
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: natom, nshell
      real(c_double), intent(in) :: x(*), v(*), theta_corr(*)
    end function

    ! periodic box: restart with the box record, and the box of instance k
    integer(c_int) function qgpu_write_restart_box(path, natom, x, v, box) &
        bind(C, name='qgpu_write_restart_box')
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: natom
      real(c_double), intent(in) :: x(*), v(*), box(6)
    end function

    integer(c_int) function qgpu_box(h, k, box, is_box) bind(C, name='qgpu_box')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: k
      real(c_double), intent(out) :: box(6)
      integer(c_int), intent(out) :: is_box
    end function

    ! constant pressure: out(1) tries, out(2) accepted, then each move's
    ! verdict (0 rejected, 1 accepted, 2 accepted by the random draw)
    integer(c_int) function qgpu_volume_moves(h, k, out, cap, tries) bind(C, name='qgpu_volume_moves')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: k, cap
      integer(c_int), intent(out) :: out(*)
      integer(c_int), intent(out) :: tries
    end function

    integer(c_int) function qgpu_trj_create(input_path, trj_path, trj) bind(C, name='qgpu_trj_create')
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: input_path(*)
      type(c_ptr), value :: trj_path
      type(c_ptr), intent(out) :: trj
    end function

    integer(c_int) function qgpu_trj_write(trj, x) bind(C, name='qgpu_trj_write')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: trj
      real(c_double), intent(in) :: x(*)
    end function

    integer(c_int) function qgpu_trj_close(trj) bind(C, name='qgpu_trj_close')
      import :: c_int, c_ptr
      type(c_ptr), value :: trj
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
