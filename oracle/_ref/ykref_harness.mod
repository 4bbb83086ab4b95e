﻿!mod$ v1 sum:3c7f01c474468e05
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ 49f872476a916429 n md_comm
!need$ 770f10bc665f50df n md_globals
!need$ 1ea0d66338494540 n md_types
module ykref_harness
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
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
use md_types,only:dble
use md_types,only:kind
use md_types,only:species
use md_types,only:species_eq
use md_types,only:species_neq
use md_globals,only:code_name
use md_globals,only:code_version
use md_globals,only:maxspec
use md_globals,only:xnostrain
use md_globals,only:hbarc
use md_globals,only:alpha
use md_globals,only:alphai
use md_globals,only:me
use md_globals,only:pi
use md_globals,only:runmdin
use md_globals,only:start
use md_globals,only:suffix
use md_globals,only:tstart
use md_globals,only:time
use md_globals,only:dt
use md_globals,only:tend
use md_globals,only:iseed
use md_globals,only:nspec
use md_globals,only:ni
use md_globals,only:zi
use md_globals,only:ai
use md_globals,only:spec_file
use md_globals,only:spec_list
use md_globals,only:n
use md_globals,only:ztot
use md_globals,only:rho
use md_globals,only:xlf
use md_globals,only:aspect
use md_globals,only:xl0
use md_globals,only:xl
use md_globals,only:stretch
use md_globals,only:deps
use md_globals,only:kt
use md_globals,only:xmass
use md_globals,only:rmax
use md_globals,only:proton
use md_globals,only:neutron
use md_globals,only:nullspecies
use md_globals,only:allspecies
use md_globals,only:istep
use md_globals,only:nwgroup
use md_globals,only:nwsteps
use md_globals,only:ngroup
use md_globals,only:ntot
use md_globals,only:nind
use md_globals,only:nptensor
use md_globals,only:ncom
use md_globals,only:ntnorm
use md_globals,only:ngofr
use md_globals,only:nckpt
use md_globals,only:nout
use md_globals,only:ftype
use md_globals,only:append
use md_globals,only:gspec
use md_globals,only:nbin
use md_globals,only:nsbin
use md_globals,only:qmin
use md_globals,only:dq
use md_globals,only:sim_type
use md_globals,only:coulomb
use md_globals,only:xlambda
use md_globals,only:rccut
use md_globals,only:xmuc
use md_globals,only:vc
use md_globals,only:frp
use md_globals,only:nuclear
use md_globals,only:xpacket
use md_globals,only:aa
use md_globals,only:bb
use md_globals,only:cc
use md_globals,only:rncut
use md_globals,only:bfield
use md_globals,only:efield
use md_globals,only:q0
use md_globals,only:w0
use md_globals,only:tref
use md_globals,only:myrank
use md_globals,only:nprocs
use md_globals,only:x
use md_globals,only:v
use md_globals,only:vold
use md_globals,only:a
use md_globals,only:ctype
use md_globals,only:zii
use md_globals,only:aii
use md_globals,only:ek
use md_globals,only:ev
use md_globals,only:e
use md_globals,only:eva
use md_globals,only:ev2a
use md_globals,only:dev
use md_globals,only:px
use md_globals,only:pa
use md_globals,only:p2a
use md_globals,only:dp
use md_globals,only:pp
use md_globals,only:gg
use md_globals,only:cgg
use md_globals,only:ss
use md_globals,only:cs
use md_globals,only:irnd
use md_globals,only:ranname
use md_comm,only:mp_type
use md_comm,only:mpi_send
use md_comm,only:mpi_recv
use md_comm,only:mpi_bcast
use md_comm,only:mpi_reduce
use md_comm,only:mpi_allreduce
use md_comm,only:bcast_parms
use md_comm,only:mpi_reduce_da
use md_comm,only:mpi_allreduce_d
use md_comm,only:mpi_allreduce_da
use md_comm,only:mpi_allreduce_da2
use md_comm,only:mpi_init
use md_comm,only:mpi_finalize
use md_comm,only:mpi_comm_size
use md_comm,only:mpi_comm_rank
use md_comm,only:mpi_send_i
use md_comm,only:mpi_send_d
use md_comm,only:mpi_recv_i
use md_comm,only:mpi_recv_d
use md_comm,only:mpi_bcast_i1xn
use md_comm,only:mpi_bcast_d3xn
use md_comm,only:mpi_bcast_species1xn
use md_comm,only:mpi_bcast_d1xn
use md_comm,only:mpi_barrier
use,intrinsic::iso_c_binding,only:operator(==)
use md_types,only:operator(==)
use md_types,only:operator(==)
use md_types,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use md_types,only:operator(/=)
use md_types,only:operator(/=)
use md_types,only:operator(/=)
real(8)::t_md
real(8)::ts_md
real(8)::t_newton
real(8)::ts_newton
real(8)::t_accel
real(8)::ts_accel
real(8)::t_calc_a
real(8)::ts_calc_a
real(8)::t_vtot
real(8)::ts_vtot
real(8)::t_calc_v
real(8)::ts_calc_v
real(8)::t_g
real(8)::ts_g
real(8)::t_pressure
real(8)::ts_pressure
real(8)::t_calc_vir
real(8)::ts_calc_vir
real(8)::t_strain
real(8)::ts_strain
integer(4)::n_md
integer(4)::n_newton
integer(4)::n_accel
integer(4)::n_calc_a
integer(4)::n_vtot
integer(4)::n_calc_v
integer(4)::n_g
integer(4)::n_pressure
integer(4)::n_calc_vir
integer(4)::n_strain
logical(4),save::timers_ready
interface operator(==)
end interface
interface operator(/=)
end interface
common/timers/t_md,ts_md,t_newton,ts_newton,t_accel,ts_accel,t_calc_a,ts_calc_a,t_vtot,ts_vtot,t_calc_v,ts_calc_v,t_g,ts_g,t_pressure,ts_pressure,t_calc_vir,ts_calc_vir,t_strain,ts_strain
common/counters/n_md,n_newton,n_accel,n_calc_a,n_vtot,n_calc_v,n_g,n_pressure,n_calc_vir,n_strain
contains
function ykref_setup(kind,nn,coulomb_on,nuclear_on,prm,xin,vin,zin,ain) bind(c,name="ykref_setup")
integer(4),value::kind
integer(4),value::nn
integer(4),value::coulomb_on
integer(4),value::nuclear_on
real(8),intent(in)::prm(1_8:17_8)
real(8),intent(in)::xin(1_8:3_8,1_8:int(nn,kind=8))
real(8),intent(in)::vin(1_8:3_8,1_8:int(nn,kind=8))
real(8),intent(in)::zin(1_8:int(nn,kind=8))
real(8),intent(in)::ain(1_8:int(nn,kind=8))
integer(4)::ykref_setup
end
subroutine pack_meas(meas)
real(8),intent(out)::meas(1_8:10_8)
end
subroutine ykref_accel(do_meas,aout,meas) bind(c,name="ykref_accel")
integer(4),value::do_meas
real(8),intent(out)::aout(1_8:3_8,1_8:int(n+3_4,kind=8))
real(8),intent(out)::meas(1_8:10_8)
end
subroutine ykref_newton(nsteps,dtin,meas_last,meas) bind(c,name="ykref_newton")
integer(4),value::nsteps
real(8),value::dtin
integer(4),value::meas_last
real(8),intent(out)::meas(1_8:10_8)
end
subroutine ykref_newton_alt(which,nsteps,dtin,meas) bind(c,name="ykref_newton_alt")
integer(4),value::which
integer(4),value::nsteps
real(8),value::dtin
real(8),intent(out)::meas(1_8:10_8)
end
subroutine ykref_set_accel(ain_) bind(c,name="ykref_set_accel")
real(8),intent(in)::ain_(1_8:3_8,1_8:int(n,kind=8))
end
subroutine ykref_get(xout,vout,aout) bind(c,name="ykref_get")
real(8),intent(out)::xout(1_8:3_8,1_8:int(n,kind=8))
real(8),intent(out)::vout(1_8:3_8,1_8:int(n,kind=8))
real(8),intent(out)::aout(1_8:3_8,1_8:int(n,kind=8))
end
function ykref_ttot() bind(c,name="ykref_ttot")
real(8)::ykref_ttot
end
subroutine ykref_timers(t) bind(c,name="ykref_timers")
real(8),intent(out)::t(1_8:6_8)
end
subroutine ykref_reset_timers() bind(c,name="ykref_reset_timers")
end
end
