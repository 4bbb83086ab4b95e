﻿!mod$ v1 sum:48e667c7f08623b3
!need$ d08c980de4d7a69b n yk_bind
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module md_yk_shim_state
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
use yk_bind,only:yk_pure_ion
use yk_bind,only:yk_ion_mixture
use yk_bind,only:yk_nucleon
use yk_bind,only:yk_bfield_test
use yk_bind,only:yk_fp64
use yk_bind,only:yk_mixed
use yk_bind,only:yk_flag_ref_f1
use yk_bind,only:yk_flag_kdk
use yk_bind,only:yk_flag_timing
use yk_bind,only:yk_flag_allpairs
use yk_bind,only:yk_flag_sw
use yk_bind,only:yk_ok
use yk_bind,only:yk_params_t
use yk_bind,only:yk_meas_t
use yk_bind,only:yk_timing_t
use yk_bind,only:yk_abi_version
use yk_bind,only:yk_create
use yk_bind,only:yk_destroy
use yk_bind,only:yk_set_params
use yk_bind,only:yk_put_state
use yk_bind,only:yk_put_accel
use yk_bind,only:yk_get_state
use yk_bind,only:yk_accel
use yk_bind,only:yk_newton
use yk_bind,only:yk_run
use yk_bind,only:yk_ttot
use yk_bind,only:yk_get_timing
use yk_bind,only:yk_reset_timing
use yk_bind,only:yk_batch_create
use yk_bind,only:yk_batch_destroy
use yk_bind,only:yk_batch_footprint
use yk_bind,only:yk_batch_add
use yk_bind,only:yk_batch_remove
use yk_bind,only:yk_batch_set_tuning
use yk_bind,only:yk_batch_set_resort
use yk_bind,only:yk_batch_run
use yk_bind,only:yk_batch_status
use yk_bind,only:yk_batch_get
use yk_bind,only:yk_batch_last_run
use yk_bind,only:yk_error_message
use yk_bind,only:yk_batch_error_message
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
type(c_ptr),save::handle
integer(4),save::handle_n
character(20_4,1),save::handle_type
contains
subroutine shim_check(rc,where)
integer(4),intent(in)::rc
character(*,1),intent(in)::where
end
subroutine shim_params(p)
type(yk_params_t),intent(out)::p
end
subroutine shim_sync_in()
end
subroutine shim_take(m,do_measurements,from_newton)
type(yk_meas_t),intent(in)::m
logical(4),intent(in)::do_measurements
logical(4),intent(in)::from_newton
end
end
