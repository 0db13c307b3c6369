﻿!mod$ v1 sum:c7f872a0786f8194
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module qgpu
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::qgpu_fp64=0_4
integer(4),parameter::qgpu_mixed=1_4
interface
function qgpu_create(input_files,n_instances,precision,handle) bind(c,name="qgpu_create")
import::c_ptr
type(c_ptr),intent(in)::input_files(1_8:*)
integer(4),value::n_instances
integer(4),value::precision
type(c_ptr),intent(out)::handle
integer(4)::qgpu_create
end
end interface
interface
function qgpu_destroy(handle) bind(c,name="qgpu_destroy")
import::c_ptr
type(c_ptr),value::handle
integer(4)::qgpu_destroy
end
end interface
interface
function qgpu_make_pair_lists(handle) bind(c,name="qgpu_make_pair_lists")
import::c_ptr
type(c_ptr),value::handle
integer(4)::qgpu_make_pair_lists
end
end interface
interface
function qgpu_pot_energy(handle) bind(c,name="qgpu_pot_energy")
import::c_ptr
type(c_ptr),value::handle
integer(4)::qgpu_pot_energy
end
end interface
interface
function qgpu_pot_energy_nonbonds(handle,k,x,d,e,eq,eq_cap) bind(c,name="qgpu_pot_energy_nonbonds")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(in)::x(1_8:*)
real(8),intent(inout)::d(1_8:*)
real(8),intent(out)::e(1_8:*)
real(8),intent(out)::eq(1_8:*)
integer(4),value::eq_cap
integer(4)::qgpu_pot_energy_nonbonds
end
end interface
interface
function qgpu_shake(handle,k,xx,x,niter) bind(c,name="qgpu_shake")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(in)::xx(1_8:*)
real(8),intent(inout)::x(1_8:*)
integer(4),intent(out)::niter
integer(4)::qgpu_shake
end
end interface
interface
function qgpu_instance_status(handle,k,bits,at_step) bind(c,name="qgpu_instance_status")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
integer(4),intent(out)::bits
integer(4),intent(out)::at_step
integer(4)::qgpu_instance_status
end
end interface
interface
function qgpu_step(handle,nsteps) bind(c,name="qgpu_step")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::nsteps
integer(4)::qgpu_step
end
end interface
interface
function qgpu_get_state(handle,k,x,v) bind(c,name="qgpu_get_state")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::x(1_8:*)
real(8),intent(out)::v(1_8:*)
integer(4)::qgpu_get_state
end
end interface
interface
function qgpu_set_state(handle,k,x,v) bind(c,name="qgpu_set_state")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(in)::x(1_8:*)
real(8),intent(in)::v(1_8:*)
integer(4)::qgpu_set_state
end
end interface
interface
function qgpu_get_gradient(handle,k,d) bind(c,name="qgpu_get_gradient")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::d(1_8:*)
integer(4)::qgpu_get_gradient
end
end interface
interface
function qgpu_energies(handle,k,e,cap) bind(c,name="qgpu_energies")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::e(1_8:*)
integer(4),value::cap
integer(4)::qgpu_energies
end
end interface
interface
function qgpu_qenergies(handle,k,e,cap) bind(c,name="qgpu_qenergies")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::e(1_8:*)
integer(4),value::cap
integer(4)::qgpu_qenergies
end
end interface
interface
function qgpu_natom(handle,k,natom) bind(c,name="qgpu_natom")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
integer(4),intent(out)::natom
integer(4)::qgpu_natom
end
end interface
interface
function qgpu_temperature(handle,k,temp) bind(c,name="qgpu_temperature")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::temp
integer(4)::qgpu_temperature
end
end interface
interface
function qgpu_theta_corr(handle,k,tc,n) bind(c,name="qgpu_theta_corr")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::tc(1_8:*)
integer(4),intent(out)::n
integer(4)::qgpu_theta_corr
end
end interface
interface
function qgpu_set_energy_interval(handle,iene,max_frames) bind(c,name="qgpu_set_energy_interval")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::iene
integer(4),value::max_frames
integer(4)::qgpu_set_energy_interval
end
end interface
interface
function qgpu_energy_series(handle,k,out,cap,nframes) bind(c,name="qgpu_energy_series")
import::c_ptr
type(c_ptr),value::handle
integer(4),value::k
real(8),intent(out)::out(1_8:*)
integer(4),value::cap
integer(4),intent(out)::nframes
integer(4)::qgpu_energy_series
end
end interface
interface
function qgpu_monitor(h,k,out,cap,n) bind(c,name="qgpu_monitor")
import::c_ptr
type(c_ptr),value::h
integer(4),value::k
real(8),intent(out)::out(1_8:*)
integer(4),value::cap
integer(4),intent(out)::n
integer(4)::qgpu_monitor
end
end interface
interface
function qgpu_write_restart(path,natom,x,v,nshell,theta_corr) bind(c,name="qgpu_write_restart")
character(1_8,1),intent(in)::path(1_8:*)
integer(4),value::natom
real(8),intent(in)::x(1_8:*)
real(8),intent(in)::v(1_8:*)
integer(4),value::nshell
real(8),intent(in)::theta_corr(1_8:*)
integer(4)::qgpu_write_restart
end
end interface
interface
function qgpu_trj_create(input_path,trj_path,trj) bind(c,name="qgpu_trj_create")
import::c_ptr
character(1_8,1),intent(in)::input_path(1_8:*)
type(c_ptr),value::trj_path
type(c_ptr),intent(out)::trj
integer(4)::qgpu_trj_create
end
end interface
interface
function qgpu_trj_write(trj,x) bind(c,name="qgpu_trj_write")
import::c_ptr
type(c_ptr),value::trj
real(8),intent(in)::x(1_8:*)
integer(4)::qgpu_trj_write
end
end interface
interface
function qgpu_trj_close(trj) bind(c,name="qgpu_trj_close")
import::c_ptr
type(c_ptr),value::trj
integer(4)::qgpu_trj_close
end
end interface
interface
function qgpu_last_error() bind(c,name="qgpu_last_error")
import::c_ptr
type(c_ptr)::qgpu_last_error
end
end interface
contains
function qgpu_error_message() result(msg)
character(256_4,1)::msg
end
end
