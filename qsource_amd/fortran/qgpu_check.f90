! Link check of the Fortran binding: creates an engine from the qdyn input
! named on the command line, runs the requested number of steps and prints
! the potential energy, or prints the error message qgpu reports.
program qgpu_check
  use qgpu
  implicit none
  character(len=512), target :: arg
  character(kind=c_char, len=513), target :: cname
  type(c_ptr) :: files(1), h
  integer(c_int) :: rc
  real(c_double) :: e(30)
  integer :: nsteps
  character(len=32) :: sarg

  call get_command_argument(1, arg)
  call get_command_argument(2, sarg)
  nsteps = 0
  if (len_trim(sarg) > 0) read(sarg, *) nsteps
  cname = trim(arg) // c_null_char
  files(1) = c_loc(cname)
  rc = qgpu_create(files, 1_c_int, QGPU_FP64, h)
  if (rc /= 0) then
    write(*, '(a,a)') 'qgpu error: ', trim(qgpu_error_message())
    stop 1
  end if
  rc = qgpu_step(h, int(nsteps, c_int))
  if (rc /= 0) then
    write(*, '(a,a)') 'qgpu error: ', trim(qgpu_error_message())
    stop 1
  end if
  rc = qgpu_energies(h, 0_c_int, e, 30_c_int)
  write(*, '(a,f16.6)') 'potential ', e(29)
  rc = qgpu_destroy(h)
end program qgpu_check
