﻿!mod$ v1 sum:49f872476a916429
!need$ 770f10bc665f50df n md_globals
module md_comm
use md_globals,only:dble
use md_globals,only:kind
use md_globals,only:species
use md_globals,only:operator(.eq.)
use md_globals,only:operator(.ne.)
use md_globals,only:species_eq
use md_globals,only:species_neq
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
character(3_8,1)::mp_type
integer(4),private::ierr
interface mpi_send
procedure::mpi_send_i
procedure::mpi_send_d
end interface
interface mpi_recv
procedure::mpi_recv_i
procedure::mpi_recv_d
end interface
interface mpi_bcast
procedure::mpi_bcast_i1xn
procedure::mpi_bcast_d3xn
procedure::mpi_bcast_species1xn
procedure::mpi_bcast_d1xn
end interface
interface mpi_reduce
procedure::mpi_reduce_da
end interface
interface mpi_allreduce
procedure::mpi_allreduce_d
procedure::mpi_allreduce_da
procedure::mpi_allreduce_da2
end interface
contains
subroutine bcast_parms()
end
subroutine mpi_reduce_da(x,xx,count,datatype,op,root,comm,ierror)
real(8)::x(:,:)
real(8)::xx(:,:)
integer(4)::count
integer(4)::datatype
integer(4)::op
integer(4)::root
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_allreduce_d(x,xx,count,data_type,op,comm,ierror)
real(8)::x
real(8)::xx
integer(4)::count
integer(4)::data_type
integer(4)::op
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_allreduce_da(x,xx,count,data_type,op,comm,ierror)
real(8)::x(:)
real(8)::xx(:)
integer(4)::count
integer(4)::data_type
integer(4)::op
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_allreduce_da2(x,xx,count,data_type,op,comm,ierror)
real(8)::x(:,:)
real(8)::xx(:,:)
integer(4)::count
integer(4)::data_type
integer(4)::op
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_init(ierror)
integer(4)::ierror
end
subroutine mpi_finalize(ierror)
integer(4)::ierror
end
subroutine mpi_comm_size(comm,nprocs,ierror)
integer(4)::comm
integer(4)::nprocs
integer(4)::ierror
end
subroutine mpi_comm_rank(comm,myrank,ierror)
integer(4)::comm
integer(4)::myrank
integer(4)::ierror
end
subroutine mpi_send_i(buff,count,datatype,dest,tag,comm,ierror)
integer(4)::buff(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::dest
integer(4)::tag
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_send_d(buff,count,datatype,dest,tag,comm,ierror)
real(8)::buff(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::dest
integer(4)::tag
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_recv_i(buff,count,datatype,source,tag,comm,status,ierror)
integer(4)::buff(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::source
integer(4)::tag
integer(4)::comm
integer(4)::status(1_8:4_8)
integer(4)::ierror
end
subroutine mpi_recv_d(buff,count,datatype,source,tag,comm,status,ierror)
real(8)::buff(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::source
integer(4)::tag
integer(4)::comm
integer(4)::status(1_8:4_8)
integer(4)::ierror
end
subroutine mpi_bcast_i1xn(buf,count,datatype,root,comm,ierror)
integer(4)::buf(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::root
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_bcast_d3xn(buf,count,datatype,root,comm,ierror)
real(8)::buf(1_8:3_8,1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::root
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_bcast_species1xn(buf,count,datatype,root,comm,ierror)
use md_types,only:species
type(species)::buf(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::root
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_bcast_d1xn(buf,count,datatype,root,comm,ierror)
real(8)::buf(1_8:*)
integer(4)::count
integer(4)::datatype
integer(4)::root
integer(4)::comm
integer(4)::ierror
end
subroutine mpi_barrier(comm,ierror)
integer(4)::comm
integer(4)::ierror
end
end
