﻿!mod$ v1 sum:1ea0d66338494540
module md_types
integer(4),parameter::dble=8_4
intrinsic::kind
type::species
integer(4)::n=0_4
real(8)::z=0._8
real(8)::a=0._8
end type
interface operator(.eq.)
procedure::species_eq
end interface
interface operator(.ne.)
procedure::species_neq
end interface
contains
function species_eq(s1,s2)
type(species),intent(in)::s1
type(species),intent(in)::s2
logical(4)::species_eq
end
function species_neq(s1,s2)
type(species),intent(in)::s1
type(species),intent(in)::s2
logical(4)::species_neq
end
end
