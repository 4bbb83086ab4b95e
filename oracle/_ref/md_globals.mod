﻿!mod$ v1 sum:770f10bc665f50df
!need$ 1ea0d66338494540 n md_types
module md_globals
use md_types,only:dble
use md_types,only:kind
use md_types,only:species
use md_types,only:operator(.eq.)
use md_types,only:operator(.ne.)
use md_types,only:species_eq
use md_types,only:species_neq
character(10_8,1),parameter::code_name="MD        "
character(8_8,1),parameter::code_version="6.2.0   "
integer(4),parameter::maxspec=1000_4
real(8),parameter::xnostrain=1.e3_8
real(8),parameter::hbarc=1.97326999999999998181010596454143524169921875e2_8
real(8),parameter::alpha=7.29735252050556053304841697126903454773128032684326171875e-3_8
real(8),parameter::alphai=1.3703600000000000136424205265939235687255859375e2_8
real(8),parameter::me=5.109989999999999810142981004901230335235595703125e-1_8
real(8),parameter::pi=3.141592653589793115997963468544185161590576171875_8
character(256_8,1)::runmdin
character(256_8,1)::start
character(40_8,1)::suffix
real(8)::tstart
real(8)::time
real(8)::dt
real(8)::tend
integer(4)::iseed
integer(4)::irnd
integer(4)::nspec
integer(4)::ni
real(8)::zi
real(8)::ai
character(256_8,1)::spec_file
type(species),save::spec_list(1_8:1000_8)
integer(4)::n
real(8)::ztot
real(8)::rho
real(8)::xlf
real(8)::aspect(1_8:3_8)
real(8)::xl0(1_8:3_8)
real(8)::xl(1_8:3_8)
real(8)::stretch(1_8:3_8)
real(8)::deps(1_8:3_8)
real(8)::kt
real(8)::xmass
real(8)::rmax
type(species),parameter::proton=species(n=0_4,z=1._8,a=1._8)
type(species),parameter::neutron=species(n=0_4,z=0._8,a=1._8)
type(species),parameter::nullspecies=species(n=0_4,z=0._8,a=0._8)
type(species),parameter::allspecies=species(n=0_4,z=-1._8,a=-1._8)
integer(4)::istep
integer(4)::nwgroup
integer(4)::nwsteps
integer(4)::ngroup
integer(4)::ntot
integer(4)::nind
integer(4)::nptensor
integer(4)::ncom
integer(4)::ntnorm
integer(4)::ngofr
integer(4)::nckpt
integer(4)::nout
character(6_8,1)::ftype
logical(4)::append
type(species),save::gspec
integer(4)::nbin
integer(4)::nsbin
real(8)::qmin
real(8)::dq
character(6_8,1)::ranname(0_8:4_8)
character(20_8,1)::sim_type
character(20_8,1)::coulomb
real(8)::xlambda
real(8)::rccut
real(8)::xmuc
real(8)::vc
real(8)::frp
character(20_8,1)::nuclear
real(8)::xpacket
real(8)::aa
real(8)::bb
real(8)::cc
real(8)::rncut
real(8)::bfield
real(8)::efield
real(8)::q0
real(8)::w0
real(8)::tref
integer(4)::myrank
integer(4)::nprocs
real(8),allocatable,target::x(:,:)
real(8),allocatable::v(:,:)
real(8),allocatable::vold(:,:)
real(8),allocatable::a(:,:)
character(6_8,1),allocatable::ctype(:)
real(8),allocatable::zii(:)
real(8),allocatable::aii(:)
real(8)::ek
real(8)::ev
real(8)::e
real(8),allocatable::eva(:)
real(8),allocatable::ev2a(:)
real(8),allocatable::dev(:)
real(8)::px
real(8),allocatable::pa(:)
real(8),allocatable::p2a(:)
real(8),allocatable::dp(:)
real(8)::pp(1_8:3_8,1_8:3_8)
real(8),allocatable::gg(:,:,:)
real(8),allocatable::cgg(:)
real(8),allocatable::ss(:,:)
real(8),allocatable::cs(:)
common/rndtype/irnd,ranname
end
