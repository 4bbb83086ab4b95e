﻿!mod$ v1 sum:d08c980de4d7a69b
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module yk_bind
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
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::yk_pure_ion=0_4
integer(4),parameter::yk_ion_mixture=1_4
integer(4),parameter::yk_nucleon=2_4
integer(4),parameter::yk_bfield_test=3_4
integer(4),parameter::yk_fp64=0_4
integer(4),parameter::yk_mixed=1_4
integer(4),parameter::yk_flag_ref_f1=1_4
integer(4),parameter::yk_flag_kdk=2_4
integer(4),parameter::yk_flag_timing=4_4
integer(4),parameter::yk_flag_allpairs=8_4
integer(4),parameter::yk_flag_sw=16_4
integer(4),parameter::yk_ok=0_4
type,bind(c)::yk_params_t
integer(4)::sim_type=0_4
integer(4)::precision=0_4
integer(4)::flags=0_4
integer(4)::coulomb_on=1_4
integer(4)::nuclear_on=1_4
integer(4)::resort_every=0_4
real(8)::xl(1_8:3_8)=[REAL(8)::0._8,0._8,0._8]
real(8)::xmuc=0._8
real(8)::rccut=1.0000000000000000762976984109188700329496497094656e50_8
real(8)::frp=1._8
real(8)::vc=1.43996468081380069037322755320928990840911865234375_8
real(8)::xmass=9.31e2_8
real(8)::rho=0._8
real(8)::kt=0._8
real(8)::zi=0._8
real(8)::ai=0._8
real(8)::aa=1.1e2_8
real(8)::bb=-2.6e1_8
real(8)::cc=2.4e1_8
real(8)::xpacket=1.25_8
real(8)::rncut=0._8
real(8)::bfield=0._8
end type
type,bind(c)::yk_meas_t
real(8)::ev=0._8
real(8)::ek=0._8
real(8)::px=0._8
real(8)::vir=0._8
real(8)::pp(1_8:6_8)=[REAL(8)::0._8,0._8,0._8,0._8,0._8,0._8]
end type
type,bind(c)::yk_timing_t
real(8)::force_ms=0._8
integer(8)::force_launches=0_8
real(8)::step_ms=0._8
integer(8)::steps=0_8
integer(8)::pairs_swept=0_8
end type
interface
function yk_abi_version() bind(c,name="yk_abi_version")
integer(4)::yk_abi_version
end
end interface
private::yk_last_error_c
interface
function yk_last_error_c() bind(c,name="yk_last_error")
import::c_ptr
type(c_ptr)::yk_last_error_c
end
end interface
interface
function yk_create(params,n,zii,aii,sys) bind(c,name="yk_create")
import::c_ptr
import::yk_params_t
type(yk_params_t),intent(in)::params
integer(4),value::n
real(8),intent(in)::zii(1_8:*)
real(8),intent(in)::aii(1_8:*)
type(c_ptr),intent(out)::sys
integer(4)::yk_create
end
end interface
interface
subroutine yk_destroy(sys) bind(c,name="yk_destroy")
import::c_ptr
type(c_ptr),value::sys
end
end interface
interface
function yk_set_params(sys,params) bind(c,name="yk_set_params")
import::c_ptr
import::yk_params_t
type(c_ptr),value::sys
type(yk_params_t),intent(in)::params
integer(4)::yk_set_params
end
end interface
interface
function yk_put_state(sys,x,v) bind(c,name="yk_put_state")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(in)::x(1_8:3_8,1_8:*)
real(8),intent(in)::v(1_8:3_8,1_8:*)
integer(4)::yk_put_state
end
end interface
interface
function yk_put_accel(sys,a) bind(c,name="yk_put_accel")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(in)::a(1_8:3_8,1_8:*)
integer(4)::yk_put_accel
end
end interface
interface
function yk_get_state(sys,x,v,a) bind(c,name="yk_get_state")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(out)::x(1_8:3_8,1_8:*)
real(8),intent(out)::v(1_8:3_8,1_8:*)
real(8),intent(out)::a(1_8:3_8,1_8:*)
integer(4)::yk_get_state
end
end interface
interface
function yk_accel(sys,do_measurements,meas) bind(c,name="yk_accel")
import::c_ptr
import::yk_meas_t
type(c_ptr),value::sys
integer(4),value::do_measurements
type(yk_meas_t),intent(out)::meas
integer(4)::yk_accel
end
end interface
interface
function yk_newton(sys,dt,do_measurements,meas) bind(c,name="yk_newton")
import::c_ptr
import::yk_meas_t
type(c_ptr),value::sys
real(8),value::dt
integer(4),value::do_measurements
type(yk_meas_t),intent(out)::meas
integer(4)::yk_newton
end
end interface
interface
function yk_run(sys,dt,nsteps,nind,records) bind(c,name="yk_run")
import::c_ptr
import::yk_meas_t
type(c_ptr),value::sys
real(8),value::dt
integer(4),value::nsteps
integer(4),value::nind
type(yk_meas_t),intent(inout)::records(1_8:*)
integer(4)::yk_run
end
end interface
interface
function yk_ttot(sys,eka) bind(c,name="yk_ttot")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(out)::eka
integer(4)::yk_ttot
end
end interface
interface
function yk_get_timing(sys,t) bind(c,name="yk_get_timing")
import::c_ptr
import::yk_timing_t
type(c_ptr),value::sys
type(yk_timing_t),intent(out)::t
integer(4)::yk_get_timing
end
end interface
interface
function yk_reset_timing(sys) bind(c,name="yk_reset_timing")
import::c_ptr
type(c_ptr),value::sys
integer(4)::yk_reset_timing
end
end interface
private::yk_batch_last_error_c
interface
function yk_batch_last_error_c() bind(c,name="yk_batch_last_error")
import::c_ptr
type(c_ptr)::yk_batch_last_error_c
end
end interface
interface
function yk_batch_create(capacity,max_replicas,max_records,b) bind(c,name="yk_batch_create")
import::c_ptr
integer(8),value::capacity
integer(4),value::max_replicas
integer(4),value::max_records
type(c_ptr),intent(out)::b
integer(4)::yk_batch_create
end
end interface
interface
subroutine yk_batch_destroy(b) bind(c,name="yk_batch_destroy")
import::c_ptr
type(c_ptr),value::b
end
end interface
interface
function yk_batch_footprint(b,n,slots,bytes) bind(c,name="yk_batch_footprint")
import::c_ptr
type(c_ptr),value::b
integer(4),value::n
integer(8),intent(out)::slots
integer(8),intent(out)::bytes
integer(4)::yk_batch_footprint
end
end interface
interface
function yk_batch_add(b,params,n,zii,aii,x,v,dt,steps,id) bind(c,name="yk_batch_add")
import::c_ptr
import::yk_params_t
type(c_ptr),value::b
type(yk_params_t),intent(in)::params
integer(4),value::n
real(8),intent(in)::zii(1_8:*)
real(8),intent(in)::aii(1_8:*)
real(8),intent(in)::x(1_8:*)
real(8),intent(in)::v(1_8:*)
real(8),value::dt
integer(4),value::steps
integer(4),intent(out)::id
integer(4)::yk_batch_add
end
end interface
interface
function yk_batch_remove(b,id) bind(c,name="yk_batch_remove")
import::c_ptr
type(c_ptr),value::b
integer(4),value::id
integer(4)::yk_batch_remove
end
end interface
interface
function yk_batch_set_tuning(b,nbuckets,n_upper,chunk_tiles,waves) bind(c,name="yk_batch_set_tuning")
import::c_ptr
type(c_ptr),value::b
integer(4),value::nbuckets
integer(4),intent(in)::n_upper(1_8:*)
integer(4),intent(in)::chunk_tiles(1_8:*)
integer(4),intent(in)::waves(1_8:*)
integer(4)::yk_batch_set_tuning
end
end interface
interface
function yk_batch_set_resort(b,every) bind(c,name="yk_batch_set_resort")
import::c_ptr
type(c_ptr),value::b
integer(4),value::every
integer(4)::yk_batch_set_resort
end
end interface
interface
function yk_batch_run(b,nsteps,nind) bind(c,name="yk_batch_run")
import::c_ptr
type(c_ptr),value::b
integer(4),value::nsteps
integer(4),value::nind
integer(4)::yk_batch_run
end
end interface
interface
function yk_batch_status(b,id,steps_left,nrec) bind(c,name="yk_batch_status")
import::c_ptr
type(c_ptr),value::b
integer(4),value::id
integer(4),intent(out)::steps_left
integer(4),intent(out)::nrec
integer(4)::yk_batch_status
end
end interface
interface
function yk_batch_get(b,id,x,v,a,init,records,rec_cap,nrec) bind(c,name="yk_batch_get")
import::c_ptr
import::yk_meas_t
type(c_ptr),value::b
integer(4),value::id
real(8),intent(inout)::x(1_8:*)
real(8),intent(inout)::v(1_8:*)
real(8),intent(inout)::a(1_8:*)
type(yk_meas_t),intent(inout)::init
type(yk_meas_t),intent(inout)::records(1_8:*)
integer(4),value::rec_cap
integer(4),intent(out)::nrec
integer(4)::yk_batch_get
end
end interface
interface
function yk_batch_last_run(b,ms,steps,items) bind(c,name="yk_batch_last_run")
import::c_ptr
type(c_ptr),value::b
real(8),intent(out)::ms
integer(8),intent(out)::steps
integer(8),intent(out)::items
integer(4)::yk_batch_last_run
end
end interface
private::c_strlen
interface
function c_strlen(s) bind(c,name="strlen")
import::c_ptr
type(c_ptr),value::s
integer(8)::c_strlen
end
end interface
private::c_string
contains
function yk_error_message() result(msg)
character(:,1),allocatable::msg
end
function yk_batch_error_message() result(msg)
character(:,1),allocatable::msg
end
function c_string(p) result(msg)
type(c_ptr),intent(in)::p
character(:,1),allocatable::msg
end
end
